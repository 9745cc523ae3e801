// C ABI of t3ns_amd (include/t3ns_amd.h, include/t3ns_host.h).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/t3ns_amd.h"
#include "../../include/t3ns_host.h"
#include "dev/davidson.hpp"
#include "dev/decompose_dev.hpp"
#include "dev/heff_dev.hpp"
#include "dev/sitetensor_dev.hpp"
#include "dev/renorm_dev.hpp"
#include "dev/resident_ops.hpp"
#include "heff_tune.hpp"
#include "host/context.hpp"
#include "host/heff_plan.hpp"
#include "host/tensors.hpp"

using namespace t3;

namespace {
thread_local std::string g_last_error;
}
namespace t3 {
void set_last_error(const std::string& msg) { g_last_error = msg; }
}
namespace {

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return T3NS_OK;
    } catch (const std::bad_alloc& e) {
        return fail(T3NS_ERR_NOMEM, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(T3NS_ERR_INVALID, e.what());
    } catch (const std::domain_error& e) {
        return fail(T3NS_ERR_SOLVER, e.what());
    } catch (const std::runtime_error& e) {
        const std::string m = e.what();
        return fail(m.rfind("HIP error", 0) == 0 ? T3NS_ERR_DEVICE : T3NS_ERR_INVALID, m);
    } catch (...) {
        return fail(T3NS_ERR_INVALID, "unknown error");
    }
}

Symsecs symsecs_from(const t3ns_symsecs_desc& d) {
    Symsecs s;
    s.bond = d.bond;
    s.nrSecs = d.nrSecs;
    s.irreps.resize(d.nrSecs);
    s.dims.resize(d.nrSecs);
    s.fcidims.assign(d.nrSecs, 1.0);
    s.totaldims = 0;
    for (int i = 0; i < d.nrSecs; ++i) {
        for (int k = 0; k < MAX_SYMMETRIES; ++k) s.irreps[i][k] = d.irreps[i * T3NS_MAX_SYMMETRIES + k];
        s.dims[i] = d.dims[i];
        s.totaldims += d.dims[i];
    }
    return s;
}

// copy = false: the operator data stays in the caller's arrays (tel_src)
ROperators rops_from(const t3ns_rops_desc& d, bool copy = true) {
    ROperators r;
    r.bond = d.bond;
    r.is_left = d.is_left;
    r.P_operator = d.P_operator;
    r.nrhss = d.nrhss;
    if (d.nrhss <= 0 || !d.begin_blocks_of_hss) throw std::invalid_argument("rops: bad nrhss");
    r.begin_blocks_of_hss.assign(d.begin_blocks_of_hss, d.begin_blocks_of_hss + d.nrhss + 1);
    const int64_t nq = (int64_t)r.begin_blocks_of_hss[d.nrhss] * (d.P_operator ? 3 : 1);
    r.qnumbers.assign(d.qnumbers, d.qnumbers + nq);
    r.nrops = d.nrops;
    r.hss_of_ops.assign(d.hss_of_ops, d.hss_of_ops + d.nrops);
    r.beginblock.resize(d.nrops);
    for (int o = 0; o < d.nrops; ++o) {
        const int h = r.hss_of_ops[o];
        if (h < 0 || h >= r.nrhss) throw std::invalid_argument("rops: bad hss_of_ops");
        const int nb = r.nblocks_hss(h);
        r.beginblock[o].assign(d.beginblock[o], d.beginblock[o] + nb + 1);
        const int64_t b0 = r.beginblock[o][0];
        for (auto& v : r.beginblock[o]) v -= b0;
    }
    r.finalize_offsets();
    if (!copy) {
        r.tel_src.resize(d.nrops, nullptr);
        for (int o = 0; o < d.nrops; ++o)
            if (r.beginblock[o].back()) r.tel_src[o] = d.tel[o] + d.beginblock[o][0];
        return r;
    }
    r.tel.resize(r.total_size());
    for (int o = 0; o < d.nrops; ++o) {
        const int64_t n = r.beginblock[o].back();
        if (n) std::memcpy(r.tel.data() + r.op_offset[o], d.tel[o] + d.beginblock[o][0], n * sizeof(double));
    }
    return r;
}

// Streams of destroyed Heff objects and finished solves, kept per device
// for the next owner: a sweep makes one Heff per step, and creating a
// stream costs about as much host time as a step's small kernel launches.
// A stream enters the pool idle (synchronised), so its next owner's work
// is ordered after nothing.
std::mutex g_stream_mu;
std::vector<std::pair<int, hipStream_t>> g_streams;
constexpr size_t kStreamPool = 32;

hipStream_t stream_acquire() {
    int dev = 0;
    T3_HIP_CHECK(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        for (size_t i = g_streams.size(); i-- > 0;)
            if (g_streams[i].first == dev) {
                hipStream_t s = g_streams[i].second;
                g_streams.erase(g_streams.begin() + i);
                return s;
            }
    }
    hipStream_t s = nullptr;
    T3_HIP_CHECK(hipStreamCreate(&s));
    return s;
}

void stream_release(hipStream_t s) {
    if (!s) return;
    hipDevice_t dev = 0;
    if (hipStreamSynchronize(s) == hipSuccess && hipStreamGetDevice(s, &dev) == hipSuccess) {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        if (g_streams.size() < kStreamPool) {
            g_streams.emplace_back((int)dev, s);
            return;
        }
    }
    (void)hipStreamDestroy(s);
}
}  // namespace

// Host copy of the boundary descriptors in the internal representation.
struct HeffHost {
    Symsecs symarr[4][3];
    Symsecs mpo;
    ROperators ops[3];
    SiteTensor site;
    InstructionSet iset;
    HeffInput in;
};

namespace {
void convert_heff_input(HeffHost* h, const t3ns_heff_desc* d, const t3ns_rops_desc* ops,
                        const t3ns_site_desc* site) {
    HeffInput& in = h->in;
    in.sym.nrSyms = d->nrSyms;
    for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) in.sym.sgs[i] = d->sgs[i];
    h->mpo = symsecs_from(d->mpo);
    in.mpo = &h->mpo;
    const int H = h->mpo.nrSecs;
    in.hermitian.assign(d->hermitian, d->hermitian + H);
    in.singlet.assign(d->singlet, d->singlet + (size_t)H * H * H);
    h->site.nrsites = site->nrsites;
    if (site->nrsites < 1 || site->nrsites > 4) throw std::invalid_argument("bad nrsites");
    for (int i = 0; i < 4; ++i) h->site.sites[i] = site->sites[i];
    h->site.nrblocks = site->nrblocks;
    h->site.qnumbers.assign(site->qnumbers, site->qnumbers + (size_t)site->nrblocks * site->nrsites);
    h->site.beginblock.assign(site->beginblock, site->beginblock + site->nrblocks + 1);
    in.site = &h->site;
    for (int i = 0; i < site->nrsites; ++i) {
        in.site_is_phys[i] = site->site_is_phys[i] != 0;
        for (int k = 0; k < 3; ++k) {
            h->symarr[i][k] = symsecs_from(d->symarr[i][k]);
            in.symarr[i][k] = &h->symarr[i][k];
        }
    }
    const bool isdmrg = (site->nrsites == 1 && site->site_is_phys[0]) ||
                        (site->nrsites == 2 && site->site_is_phys[0] && site->site_is_phys[1]);
    for (int k = 0; k < (isdmrg ? 2 : 3); ++k) {
        h->ops[k] = rops_from(ops[k], /*copy=*/false);  // uploaded from the caller's arrays
        in.ops[k] = &h->ops[k];
        in.rops_on_site[k] = d->rops_on_site[k];
    }
    if (isdmrg) {
        h->ops[2] = ROperators();
        h->ops[2].P_operator = 0;
        in.ops[2] = &h->ops[2];
    }
    // merge instruction set, sorted and grouped as fetch_merge does
    h->iset.step = d->merge_step;
    for (int i = 0; i < d->merge_nr; ++i) {
        Instruction ins;
        for (int k = 0; k < 3; ++k) ins.instr[k] = d->merge_instr[3 * i + k];
        ins.pref = d->merge_pref[i];
        if (ins.instr[0] < 0 || ins.instr[1] < 0 || ins.instr[2] < 0) continue;
        h->iset.instr.push_back(ins);
    }
    const int* hss[3] = {h->ops[0].hss_of_ops.data(), h->ops[1].hss_of_ops.data(),
                         isdmrg ? nullptr : h->ops[2].hss_of_ops.data()};
    sort_instructions_merge(h->iset, hss, H);
    in.iset = &h->iset;
}

int64_t write_plan_table(const HeffPlan& P, int* ints, double* prefs) {
    int64_t c = 0;
    for (const PlanPair& pp : P.pairs)
        for (const PlanEntry& e : pp.entries) {
            if (ints) {
                int* r = ints + 7 * c;
                r[0] = pp.newsb;
                r[1] = pp.oldsb;
                r[2] = pp.bestorder;
                r[3] = e.mpo_group;
                r[4] = e.sbops[0];
                r[5] = e.sbops[1];
                r[6] = P.isdmrg ? -1 : e.sbops[2];
            }
            if (prefs) prefs[c] = e.prefactor;
            ++c;
        }
    return c;
}
}  // namespace

struct t3ns_heff {
    HeffHost host;  // empty for a derived object (t3ns_heff_derive)
    std::unique_ptr<HeffDevice> dev;
    DevBuf<double> dvec, dres;
    hipStream_t stream = nullptr;
    bool derived = false;
};

extern "C" {

const char* t3ns_last_error(void) { return g_last_error.c_str(); }

int t3ns_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int t3ns_heff_init(t3ns_heff** out, const t3ns_heff_desc* d, const t3ns_rops_desc* ops,
                   const t3ns_site_desc* site, int precision) {
    if (!out || !d || !ops || !site) return fail(T3NS_ERR_INVALID, "null argument");
    if (precision != T3NS_FP64 && precision != T3NS_FP32 && precision != T3NS_BF16)
        return fail(T3NS_ERR_INVALID, "unknown precision");
    return guarded([&] {
        if (t3ns_device_count() < 1) throw std::runtime_error("HIP error: no device visible");
        auto h = std::make_unique<t3ns_heff>();
        convert_heff_input(&h->host, d, ops, site);
        HeffInput& in = h->host.in;
        h->stream = stream_acquire();
        h->dev = std::make_unique<HeffDevice>(in, precision, h->stream);
        h->dvec.alloc(h->dev->size());
        h->dres.alloc(h->dev->size());
        *out = h.release();
    });
}

int t3ns_heff_derive(t3ns_heff** out, const t3ns_heff* base, int precision) {
    if (!out || !base) return fail(T3NS_ERR_INVALID, "null argument");
    if (precision != T3NS_FP32 || base->dev->precision() != T3NS_FP64)
        return fail(T3NS_ERR_INVALID, "a derived Heff is FP32 on an FP64 one");
    return guarded([&] {
        auto h = std::make_unique<t3ns_heff>();
        h->derived = true;
        h->stream = stream_acquire();
        T3_HIP_CHECK(hipStreamSynchronize(base->stream));
        h->dev = std::make_unique<HeffDevice>(*base->dev, base->host.in, precision, h->stream);
        h->dvec.alloc(h->dev->size());
        h->dres.alloc(h->dev->size());
        *out = h.release();
    });
}

int t3ns_heff_matvec(t3ns_heff* h, const double* vec, double* result) {
    if (!h || !vec || !result) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        const size_t n = h->dev->size();
        T3_HIP_CHECK(hipMemcpyAsync(h->dvec.get(), vec, n * sizeof(double), hipMemcpyHostToDevice,
                                    h->stream));
        h->dev->matvec(h->dvec.get(), h->dres.get(), h->stream);
        T3_HIP_CHECK(hipMemcpyAsync(result, h->dres.get(), n * sizeof(double), hipMemcpyDeviceToHost,
                                    h->stream));
        T3_HIP_CHECK(hipStreamSynchronize(h->stream));
    });
}

int t3ns_heff_matvec_device(t3ns_heff* h, const double* dvec, double* dres, void* stream) {
    if (!h || !dvec || !dres) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] { h->dev->matvec(dvec, dres, (hipStream_t)stream); });
}

int t3ns_heff_use_graphs(t3ns_heff* h, int on) {
    if (!h) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] { h->dev->set_graphs(on != 0); });
}

struct t3ns_heff_batch {
    HeffBatch b;
    hipStream_t stream = nullptr;
};

int t3ns_heff_batch_create(t3ns_heff_batch** out) {
    if (!out) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        auto b = std::make_unique<t3ns_heff_batch>();
        T3_HIP_CHECK(hipStreamCreate(&b->stream));
        *out = b.release();
    });
}

int t3ns_heff_batch_destroy(t3ns_heff_batch* b) {
    if (!b) return T3NS_OK;
    return guarded([&] {
        if (b->stream) {
            (void)hipStreamSynchronize(b->stream);
            (void)hipStreamDestroy(b->stream);
        }
        delete b;
    });
}

int t3ns_heff_batch_matvec(t3ns_heff_batch* b, t3ns_heff* const* hs, const double* const* dx,
                           double* const* dy, int n, void* stream) {
    if (!b || n < 0 || (n > 0 && (!hs || !dx || !dy))) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        std::vector<HeffDevice*> d(n);
        for (int i = 0; i < n; ++i) {
            if (!hs[i] || !dx[i] || !dy[i]) throw std::invalid_argument("batched matvec: null entry");
            d[i] = hs[i]->dev.get();
        }
        b->b.matvec(d.data(), dx, dy, n, static_cast<hipStream_t>(stream));
    });
}

int t3ns_heff_batch_matvec_host(t3ns_heff_batch* b, t3ns_heff* const* hs, const double* const* x,
                                double* const* y, int n) {
    if (!b || n < 0 || (n > 0 && (!hs || !x || !y))) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        std::vector<HeffDevice*> d(n);
        std::vector<const double*> dx(n);
        std::vector<double*> dy(n);
        for (int i = 0; i < n; ++i) {
            if (!hs[i] || !x[i] || !y[i]) throw std::invalid_argument("batched matvec: null entry");
            for (int j = 0; j < i; ++j)
                if (hs[j] == hs[i]) throw std::invalid_argument("batched matvec: an object twice");
            d[i] = hs[i]->dev.get();
            const size_t sz = d[i]->size();
            T3_HIP_CHECK(hipMemcpyAsync(hs[i]->dvec.get(), x[i], sz * sizeof(double),
                                        hipMemcpyHostToDevice, b->stream));
            dx[i] = hs[i]->dvec.get();
            dy[i] = hs[i]->dres.get();
        }
        b->b.matvec(d.data(), dx.data(), dy.data(), n, b->stream);
        for (int i = 0; i < n; ++i)
            T3_HIP_CHECK(hipMemcpyAsync(y[i], dy[i], d[i]->size() * sizeof(double),
                                        hipMemcpyDeviceToHost, b->stream));
        T3_HIP_CHECK(hipStreamSynchronize(b->stream));
    });
}

int t3ns_heff_autotune(t3ns_heff* h, int max_trials, double* out, int n) {
    if (!h || (n > 0 && !out)) return fail(T3NS_ERR_INVALID, "null argument");
    if (h->derived) return fail(T3NS_ERR_INVALID, "a derived Heff keeps its base's work lists");
    if (max_trials < 1) return fail(T3NS_ERR_INVALID, "max_trials must be at least 1");
    return guarded([&] {
        const size_t sz = h->dev->size();
        std::vector<double> x(sz);
        for (size_t i = 0; i < sz; ++i) x[i] = 0.1 * (double)(i % 7 + 1);
        T3_HIP_CHECK(hipMemcpyAsync(h->dvec.get(), x.data(), sz * sizeof(double),
                                    hipMemcpyHostToDevice, h->stream));
        const TuneReport R = autotune_heff(*h->dev, h->host.in, h->stream, h->dvec.get(),
                                           h->dres.get(), max_trials);
        const double v[] = {R.best.chunk_steps, R.best.fuse_over, (double)R.best.fuse,
                            (double)R.best.sched,  R.ms, R.baseline_ms, (double)R.trials,
                            R.cached ? 1.0 : 0.0, (double)R.best.wpc};
        for (int i = 0; i < n && i < 9; ++i) out[i] = v[i];
    });
}

int t3ns_heff_set_tuning(t3ns_heff* h, const double* t) {
    if (!h || !t) return fail(T3NS_ERR_INVALID, "null argument");
    if (h->derived) return fail(T3NS_ERR_INVALID, "a derived Heff keeps its base's work lists");
    if (!(t[0] >= 0) || !(t[1] >= 0) || t[3] < 0 || t[3] > 3 ||
        (t[4] != 0 && t[4] != 1 && t[4] != 2 && t[4] != 4))
        return fail(T3NS_ERR_INVALID, "tuning out of range");
    return guarded([&] {
        HeffTuning q;
        q.chunk_steps = t[0];
        q.fuse_over = t[1];
        q.fuse = t[2] != 0;
        q.sched = (int)t[3];
        q.wpc = (int)t[4];
        h->dev->retune(h->host.in, q, h->stream);
    });
}

int t3ns_heff_get_tuning(const t3ns_heff* h, double* t) {
    if (!h || !t) return fail(T3NS_ERR_INVALID, "null argument");
    const HeffTuning& q = h->dev->tuning();
    t[0] = q.chunk_steps;
    t[1] = q.fuse_over;
    t[2] = q.fuse;
    t[3] = q.sched;
    t[4] = q.wpc;
    return T3NS_OK;
}

void t3ns_tune_cache_clear(void) { clear_tune_cache(); }

int t3ns_heff_diagonal(t3ns_heff* h, double* diag) {
    if (!h || !diag) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        const size_t n = h->dev->size();
        h->dev->diagonal(h->dres.get(), h->stream);
        T3_HIP_CHECK(hipMemcpyAsync(diag, h->dres.get(), n * sizeof(double), hipMemcpyDeviceToHost,
                                    h->stream));
        T3_HIP_CHECK(hipStreamSynchronize(h->stream));
    });
}

int t3ns_davidson(t3ns_heff* h, double* result, double* energy, int max_vecs, int keep_deflate,
                  double tol, int max_its, const double* diagonal, int* its) {
    if (!h || !result || !energy) return fail(T3NS_ERR_INVALID, "null argument");
    int status = 0;
    const int rc = guarded([&] {
        const int64_t n = h->dev->size();
        DevBuf<double> x(n), d(n);
        T3_HIP_CHECK(hipMemcpyAsync(x.get(), result, n * sizeof(double), hipMemcpyHostToDevice,
                                    h->stream));
        if (diagonal)
            T3_HIP_CHECK(hipMemcpyAsync(d.get(), diagonal, n * sizeof(double),
                                        hipMemcpyHostToDevice, h->stream));
        else
            h->dev->diagonal(d.get(), h->stream);
        DeviceDavidson dav(n, max_vecs, keep_deflate, h->stream);
        HeffDevice* dev = h->dev.get();
        const DavidsonResult R = dav.solve(
            x.get(), d.get(), [dev](const double* a, double* b, hipStream_t s) { dev->matvec(a, b, s); },
            tol, max_its);
        if (R.status < 0) throw std::domain_error("davidson: projected eigenproblem failed");
        T3_HIP_CHECK(hipMemcpyAsync(result, x.get(), n * sizeof(double), hipMemcpyDeviceToHost,
                                    h->stream));
        T3_HIP_CHECK(hipStreamSynchronize(h->stream));
        *energy = R.energy;
        if (its) *its = R.its;
        status = R.status;
    });
    return rc != T3NS_OK ? rc : status;
}

namespace {
// The reduced-precision phase of the refined solve stops at a residual of
// kFloorFactor * (measured relative matvec error) * |theta|.
constexpr double kFloorFactor = 2.0;
}  // namespace

int t3ns_davidson_refined(t3ns_heff* h_low, t3ns_heff* h64, double* result, double* energy,
                          int max_vecs, int keep_deflate, double tol, double switch_tol,
                          int max_its, const double* diagonal, int* its_low, int* its64) {
    if (!h_low || !h64 || !result || !energy) return fail(T3NS_ERR_INVALID, "null argument");
    if (h_low->dev->size() != h64->dev->size())
        return fail(T3NS_ERR_INVALID, "davidson_refined: operators of different size");
    int status = 0;
    const int rc = guarded([&] {
        const int64_t n = h64->dev->size();
        DevBuf<double> x(n), d(n);
        hipStream_t s = h64->stream;
        T3_HIP_CHECK(hipMemcpyAsync(x.get(), result, n * sizeof(double), hipMemcpyHostToDevice, s));
        if (diagonal)
            T3_HIP_CHECK(hipMemcpyAsync(d.get(), diagonal, n * sizeof(double),
                                        hipMemcpyHostToDevice, s));
        else
            h64->dev->diagonal(d.get(), s);
        HeffDevice* lo = h_low->dev.get();
        HeffDevice* hi = h64->dev.get();
        // The reduced-precision operator's relative error, measured on the
        // guess, sets how far its iterations can usefully go: an
        // eigenvector of it is as close to the FP64 one as that error
        // allows once the residual is below about kFloorFactor * error *
        // |theta|.  Phase 1: reduced-precision matvecs down to max(tol,
        // switch_tol) or that floor.  Then, if phase 1 met tol, one FP64
        // matvec gives the FP64 Rayleigh quotient and residual, the result
        // when that residual meets tol; otherwise phase 2 resumes the same
        // search space (its lowest Ritz vectors, images recomputed in FP64)
        // with FP64 matvecs down to tol.
        const DevMatvec mvlo = [lo](const double* a, double* b, hipStream_t st) { lo->matvec(a, b, st); };
        const DevMatvec mv64 = [hi](const double* a, double* b, hipStream_t st) { hi->matvec(a, b, st); };
        DeviceDavidson dv(n, max_vecs, keep_deflate, s);
        const double rel = dv.relative_difference(x.get(), mv64, mvlo);
        const DavidsonResult R1 = dv.solve(x.get(), d.get(), mvlo, std::max(tol, switch_tol), max_its,
                                           kFloorFactor * rel);
        if (R1.status < 0) throw std::domain_error("davidson: projected eigenproblem failed");
        DavidsonResult R2;
        bool done = false;
        if (R1.status == 0 && R1.residual <= tol) {
            DeviceDavidson chk(n, 2, 1, s);
            const auto rq = chk.rayleigh(x.get(), mv64);
            if (rq.second <= tol) {
                R2.energy = rq.first;
                R2.residual = rq.second;
                done = true;
            }
        }
        if (!done) {
            R2 = dv.resume(x.get(), d.get(), mv64, tol, std::max(1, max_its - R1.its));
            if (R2.status < 0) throw std::domain_error("davidson: projected eigenproblem failed");
        }
        T3_HIP_CHECK(hipMemcpyAsync(result, x.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
        *energy = R2.energy;
        if (its_low) *its_low = R1.its;
        if (its64) *its64 = R2.its;
        status = R2.status;
    });
    return rc != T3NS_OK ? rc : status;
}

int t3ns_davidson_dev(int64_t size, double* dx, const double* ddiag, t3ns_dev_matvec matvec,
                      void* vdata, void* stream, double* energy, int max_vecs, int keep_deflate,
                      double tol, int max_its, int* its) {
    if (size <= 0 || !dx || !ddiag || !matvec || !energy) return fail(T3NS_ERR_INVALID, "bad argument");
    int status = 0;
    const int rc = guarded([&] {
        hipStream_t s = (hipStream_t)stream;
        DeviceDavidson dav(size, max_vecs, keep_deflate, s);
        const DavidsonResult R = dav.solve(
            dx, ddiag,
            [matvec, vdata](const double* a, double* b, hipStream_t st) {
                matvec(a, b, (void*)st, vdata);
            },
            tol, max_its);
        if (R.status < 0) throw std::domain_error("davidson: projected eigenproblem failed");
        *energy = R.energy;
        if (its) *its = R.its;
        status = R.status;
    });
    return rc != T3NS_OK ? rc : status;
}

int t3ns_davidson_dev_noalloc(int64_t size, double* dx, const double* ddiag,
                              t3ns_dev_matvec matvec, void* vdata, void* stream, double* energy,
                              int max_vecs, int keep_deflate, double tol, int max_its, int* its) {
    if (size <= 0 || !dx || !ddiag || !matvec || !energy) return fail(T3NS_ERR_INVALID, "bad argument");
    int status = 0;
    const int rc = guarded([&] {
        const DavidsonResult R = davidson_noalloc(
            size, dx, ddiag,
            [matvec, vdata](const double* a, double* b, hipStream_t st) {
                matvec(a, b, (void*)st, vdata);
            },
            (hipStream_t)stream, tol, max_vecs, keep_deflate, max_its);
        if (R.status < 0) throw std::domain_error("davidson: projected eigenproblem failed");
        *energy = R.energy;
        if (its) *its = R.its;
        status = R.status;
    });
    return rc != T3NS_OK ? rc : status;
}

/* Reference-shaped entry points (include/Heff.h:166-192,
 * include/davidson.h:56-60, include/wrapper_solvers.h:56-60). */

void t3ns_matvecT3NS(const double* vec, double* result, void* vdata) {
    // void like the reference: a failure is fatal there (src/macros.c:49-57)
    const int rc = t3ns_heff_matvec(static_cast<t3ns_heff*>(vdata), vec, result);
    if (rc != T3NS_OK) {
        std::fprintf(stderr, "matvecT3NS: %s\n", g_last_error.c_str());
        std::exit(EXIT_FAILURE);
    }
}

void t3ns_heff_dev_matvec(const double* dx, double* dy, void* stream, void* vdata) {
    const int rc = t3ns_heff_matvec_device(static_cast<t3ns_heff*>(vdata), dx, dy, stream);
    if (rc != T3NS_OK) {
        std::fprintf(stderr, "t3ns_heff_dev_matvec: %s\n", g_last_error.c_str());
        std::exit(EXIT_FAILURE);
    }
}

double* t3ns_make_diagonal(t3ns_heff* h) {
    if (!h) {
        fail(T3NS_ERR_INVALID, "null argument");
        return nullptr;
    }
    double* d = static_cast<double*>(std::malloc(sizeof(double) * (size_t)h->dev->size()));
    if (!d) {
        fail(T3NS_ERR_NOMEM, "make_diagonal: host allocation failed");
        return nullptr;
    }
    if (t3ns_heff_diagonal(h, d) != T3NS_OK) {
        std::free(d);
        return nullptr;
    }
    return d;
}

int t3ns_davidson_cb(double* result, double* energy, int size, int max_vecs, int keep_deflate,
                     double davidson_tol, int max_its, const double* diagonal,
                     void (*matvec)(const double*, double*, void*), void* vdat) {
    if (!result || !energy || !diagonal || !matvec || size <= 0) {
        fail(T3NS_ERR_INVALID, "davidson: bad argument");
        return -1;
    }
    int status = 0;
    const int rc = guarded([&] {
        const int64_t n = size;
        hipStream_t s = stream_acquire();
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { stream_release(s); }
        } guard{s};
        DevBuf<double> x(n), d(n);
        T3_HIP_CHECK(hipMemcpyAsync(x.get(), result, n * sizeof(double), hipMemcpyHostToDevice, s));
        T3_HIP_CHECK(hipMemcpyAsync(d.get(), diagonal, n * sizeof(double), hipMemcpyHostToDevice, s));
        std::vector<double> hx(n), hy(n);
        // the caller's matvec works on host arrays: each search vector goes
        // to the host and its image comes back
        const DevMatvec mv = [&](const double* a, double* b, hipStream_t st) {
            T3_HIP_CHECK(hipMemcpyAsync(hx.data(), a, n * sizeof(double), hipMemcpyDeviceToHost, st));
            T3_HIP_CHECK(hipStreamSynchronize(st));
            matvec(hx.data(), hy.data(), vdat);
            T3_HIP_CHECK(hipMemcpyAsync(b, hy.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
            T3_HIP_CHECK(hipStreamSynchronize(st));
        };
        DeviceDavidson dav(n, max_vecs, keep_deflate, s);
        const DavidsonResult R = dav.solve(x.get(), d.get(), mv, davidson_tol, max_its);
        if (R.status < 0) throw std::domain_error("davidson: projected eigenproblem failed");
        T3_HIP_CHECK(hipMemcpyAsync(result, x.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
        *energy = R.energy;
        status = R.status;
    });
    return rc != T3NS_OK ? -1 : status;
}

int t3ns_sparse_eigensolve(double* result, double* energy, int size, int max_vecs,
                           int keep_deflate, double tol, int max_its, const double* diagonal,
                           void (*matvec)(const double*, double*, void*), void* vdat,
                           const char solver[]) {
    if (!solver || std::strcmp(solver, "D") != 0) {
        fail(T3NS_ERR_INVALID, "sparse_eigensolve: only the Davidson solver \"D\" is built");
        return -1;
    }
    return t3ns_davidson_cb(result, energy, size, max_vecs, keep_deflate, tol, max_its, diagonal,
                            matvec, vdat);
}

}  // extern "C"

struct t3ns_rops {
    ROperators r;
    RenormWork work;
    DevArena dev;  // the device copy of r.tel, registered as resident
    std::vector<int64_t> bb_flat;
    std::vector<const int64_t*> bb_ptr;
    std::vector<const double*> tel_ptr;
    void views() {
        bb_ptr.resize(r.nrops);
        tel_ptr.resize(r.nrops);
        for (int k = 0; k < r.nrops; ++k) {
            bb_ptr[k] = r.beginblock[k].data();
            tel_ptr[k] = r.tel.data() + r.op_offset[k];
        }
        // later updates and Heff set-ups on these operators copy them on
        // the device
        if (dev) resident_put(r.tel.data(), (int64_t)r.tel.size(), dev);
        dev.reset();
    }
    ~t3ns_rops() { resident_drop(r.tel.data()); }
};

// The operators of a descriptor, with the device copy they were made in
// when it is still resident (then without a host copy), else with a host
// copy for upload.
ROperators rops_resident(const t3ns_rops_desc& d, ResidentRef& ref) {
    ROperators r = rops_from(d, /*copy=*/false);
    ref = resident_find(r);
    if (ref) return r;
    return rops_from(d, /*copy=*/true);
}

namespace {
AppendInput append_input_from(const t3ns_append_desc* d, const ROperators* old) {
    AppendInput in;
    in.sym.nrSyms = d->nrSyms;
    for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) in.sym.sgs[i] = d->sgs[i];
    in.is_left = d->is_left;
    for (int k = 0; k < 3; ++k) {
        in.site_ss[k] = symsecs_from(d->site_ss[k]);
        in.new_ss[k] = symsecs_from(d->new_ss[k]);
        in.old_ss[k] = symsecs_from(d->old_ss[k]);
    }
    in.mpo = symsecs_from(d->mpo);
    const int H = in.mpo.nrSecs;
    if (d->nr_instr < 0 || (d->nr_instr && (!d->instr || !d->pref || !d->hss_of_new)))
        throw std::invalid_argument("append: bad instructions");
    int nops = 0;
    for (int i = 0; i < d->nr_instr; ++i) {
        Instruction ins;
        for (int k = 0; k < 3; ++k) ins.instr[k] = d->instr[3 * i + k];
        ins.pref = d->pref[i];
        if (ins.instr[0] < 0 || ins.instr[0] >= old->nrops || ins.instr[1] < 0 ||
            ins.instr[1] >= d->nr_siteops || ins.instr[2] < 0)
            throw std::invalid_argument("append: instruction out of range");
        nops = std::max(nops, ins.instr[2] + 1);
        in.instr.push_back(ins);
    }
    in.hss_of_new.assign(d->hss_of_new, d->hss_of_new + nops);
    in.siteop_hss.assign(d->siteop_hss, d->siteop_hss + d->nr_siteops);
    in.physdim = d->physdim;
    in.siteop_el.assign(d->siteop_el, d->siteop_el + (size_t)d->nr_siteops * d->physdim * d->physdim);
    in.tprod_begin.assign(d->tprod_begin, d->tprod_begin + H + 1);
    in.tprods.assign(d->tprods, d->tprods + 2 * (size_t)in.tprod_begin[H]);
    in.nrhss = d->nrhss;
    if (d->nrhss != H) throw std::invalid_argument("append: nrhss differs from the MPO sectors");
    in.begin_blocks_of_hss.assign(d->begin_blocks_of_hss, d->begin_blocks_of_hss + d->nrhss + 1);
    const int nb = in.begin_blocks_of_hss[d->nrhss];
    in.qnumbers.assign(d->qnumbers, d->qnumbers + 3 * (size_t)nb);
    in.hss_bb.resize(d->nrhss);
    for (int h = 0; h < d->nrhss; ++h) {
        const int n = in.begin_blocks_of_hss[h + 1] - in.begin_blocks_of_hss[h];
        in.hss_bb[h].assign(d->hss_beginblock[h], d->hss_beginblock[h] + n + 1);
    }
    in.old = old;
    return in;
}
}  // namespace

extern "C" {

int t3ns_append_phys(const t3ns_append_desc* d, const t3ns_rops_desc* old, t3ns_rops** out) {
    if (!d || !old || !out) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        ResidentRef ref;
        const ROperators o = rops_resident(*old, ref);
        const AppendInput in = append_input_from(d, &o);
        const AppendPlan plan = plan_append(in);
        auto res = std::make_unique<t3ns_rops>();
        hipStream_t s = hipStreamPerThread;
        res->r = run_append(plan, o, s, &res->work, ref.ptr(), &res->dev);
        res->views();
        *out = res.release();
    });
}

int t3ns_update_physical(const t3ns_pupdate_desc* d, const t3ns_rops_desc* pops, t3ns_rops** out) {
    if (!d || !pops || !out) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        ResidentRef ref;
        const ROperators po = rops_resident(*pops, ref);
        PUpdateInput in;
        in.sym.nrSyms = d->nrSyms;
        for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) in.sym.sgs[i] = d->sgs[i];
        in.is_left = d->is_left;
        for (int k = 0; k < 3; ++k) in.site_ss[k] = symsecs_from(d->site_ss[k]);
        in.iss = symsecs_from(d->internal_ss);
        in.mpo = symsecs_from(d->mpo);
        if (d->T_nrblocks < 0 || (d->T_nrblocks && (!d->T_qnumbers || !d->T_beginblock || !d->T_tel)))
            throw std::invalid_argument("pUpdate: bad site tensor");
        SiteTensor T;
        T.nrsites = 1;
        T.nrblocks = d->T_nrblocks;
        T.qnumbers.assign(d->T_qnumbers, d->T_qnumbers + T.nrblocks);
        T.beginblock.assign(d->T_beginblock, d->T_beginblock + T.nrblocks + 1);
        T.tel.assign(d->T_tel, d->T_tel + T.size());
        const int H = in.mpo.nrSecs;
        if (d->nrhss != H) throw std::invalid_argument("pUpdate: nrhss differs from the MPO sectors");
        in.nrhss = d->nrhss;
        in.begin_blocks_of_hss.assign(d->begin_blocks_of_hss, d->begin_blocks_of_hss + H + 1);
        const int nb = in.begin_blocks_of_hss[H];
        in.qnumbers.assign(d->qnumbers, d->qnumbers + nb);
        in.hss_bb.resize(H);
        for (int h = 0; h < H; ++h) {
            const int n = in.begin_blocks_of_hss[h + 1] - in.begin_blocks_of_hss[h];
            in.hss_bb[h].assign(d->hss_beginblock[h], d->hss_beginblock[h] + n + 1);
        }
        if (po.nrhss != H || po.is_left != d->is_left)
            throw std::invalid_argument("pUpdate: operators do not match the descriptor");
        in.T = &T;
        in.pops = &po;
        const PUpdatePlan plan = plan_pupdate(in);
        auto res = std::make_unique<t3ns_rops>();
        hipStream_t s = hipStreamPerThread;
        res->r = run_pupdate(plan, po, T, s, &res->work, ref.ptr(), &res->dev);
        res->views();
        *out = res.release();
    });
}

namespace {
// The branching-update input of a descriptor (the operators and tensor
// copied into o0, o1, T, which `in` points at).
void bupdate_input(const t3ns_bupdate_desc* d, const t3ns_rops_desc* ops, ROperators& o0,
                   ROperators& o1, SiteTensor& T, BUpdateInput& in, ResidentRef* ref0 = nullptr,
                   ResidentRef* ref1 = nullptr) {
        ResidentRef r0, r1;
        o0 = rops_resident(ops[0], ref0 ? *ref0 : r0);
        o1 = rops_resident(ops[1], ref1 ? *ref1 : r1);
        in.sym.nrSyms = d->nrSyms;
        for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) in.sym.sgs[i] = d->sgs[i];
        in.update_case = d->update_case;
        for (int k = 0; k < 3; ++k) in.site_ss[k] = symsecs_from(d->site_ss[k]);
        in.mpo = symsecs_from(d->mpo);
        const int H = in.mpo.nrSecs;
        if (d->nrhss != H || o0.nrhss != H || o1.nrhss != H)
            throw std::invalid_argument("bUpdate: MPO sectors differ");
        if (d->T_nrblocks < 0 || (d->T_nrblocks && (!d->T_qnumbers || !d->T_beginblock || !d->T_tel)))
            throw std::invalid_argument("bUpdate: bad site tensor");
        T.nrsites = 1;
        T.nrblocks = d->T_nrblocks;
        T.qnumbers.assign(d->T_qnumbers, d->T_qnumbers + T.nrblocks);
        T.beginblock.assign(1, 0);
        if (T.nrblocks) T.beginblock.assign(d->T_beginblock, d->T_beginblock + T.nrblocks + 1);
        T.tel.assign(d->T_tel, d->T_tel + T.size());
        if (d->nr_instr < 0 || (d->nr_instr && (!d->instr || !d->pref || !d->hss_of_new)))
            throw std::invalid_argument("bUpdate: bad instructions");
        int nops = 0;
        for (int i = 0; i < d->nr_instr; ++i) {
            Instruction ins;
            for (int k = 0; k < 3; ++k) ins.instr[k] = d->instr[3 * i + k];
            ins.pref = d->pref[i];
            if (ins.instr[2] < 0) throw std::invalid_argument("bUpdate: instruction out of range");
            nops = std::max(nops, ins.instr[2] + 1);
            in.instr.push_back(ins);
        }
        in.hss_of_new.assign(d->hss_of_new, d->hss_of_new + nops);
        in.tprod_begin.assign(d->tprod_begin, d->tprod_begin + H + 1);
        in.tprods.assign(d->tprods, d->tprods + 2 * (size_t)in.tprod_begin[H]);
        in.nrhss = H;
        in.begin_blocks_of_hss.assign(d->begin_blocks_of_hss, d->begin_blocks_of_hss + H + 1);
        const int nb = in.begin_blocks_of_hss[H];
        in.qnumbers.assign(d->qnumbers, d->qnumbers + nb);
        in.hss_bb.resize(H);
        for (int h = 0; h < H; ++h) {
            const int n = in.begin_blocks_of_hss[h + 1] - in.begin_blocks_of_hss[h];
            in.hss_bb[h].assign(d->hss_beginblock[h], d->hss_beginblock[h] + n + 1);
        }
        in.T = &T;
        in.ops[0] = &o0;
        in.ops[1] = &o1;
}
}  // namespace

int t3ns_update_branching(const t3ns_bupdate_desc* d, const t3ns_rops_desc* ops, t3ns_rops** out) {
    if (!d || !ops || !out) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        ROperators o0, o1;
        SiteTensor T;
        BUpdateInput in;
        ResidentRef ref0, ref1;
        bupdate_input(d, ops, o0, o1, T, in, &ref0, &ref1);
        const BUpdatePlan plan = plan_bupdate(in);
        auto res = std::make_unique<t3ns_rops>();
        res->r = run_bupdate(plan, T, o0, o1, hipStreamPerThread, &res->work, ref0.ptr(), ref1.ptr(),
                             &res->dev);
        res->views();
        *out = res.release();
    });
}

int t3ns_bupdate_plan_stats(const t3ns_bupdate_desc* d, const t3ns_rops_desc* ops, double* out,
                            int n) {
    if (!d || !ops || (n > 0 && !out)) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        const auto t0 = std::chrono::steady_clock::now();
        ROperators o0, o1;
        SiteTensor T;
        BUpdateInput in;
        bupdate_input(d, ops, o0, o1, T, in);
        const auto t1 = std::chrono::steady_clock::now();
        const BUpdatePlan plan = plan_bupdate(in);
        const auto t2 = std::chrono::steady_clock::now();
        const double v[] = {std::chrono::duration<double>(t1 - t0).count(),
                            std::chrono::duration<double>(t2 - t1).count(),
                            (double)plan.tasks.size(), (double)plan.dests.size()};
        for (int i = 0; i < n && i < 4; ++i) out[i] = v[i];
    });
}

struct t3ns_split {
    SplitResult r;
    std::vector<int> irreps, nsing;
    std::vector<double> sing;
};

int t3ns_split_site(const t3ns_split_desc* d, t3ns_split** out, t3ns_work* work) {
    if (!d || !out) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        SplitInput in;
        in.sym.nrSyms = d->nrSyms;
        for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) in.sym.sgs[i] = d->sgs[i];
        in.nrsites = d->nrsites;
        in.siteV = d->siteV;
        if (d->nrsites < 2 || d->nrsites > 4 || d->nrblocks < 0)
            throw std::invalid_argument("split: needs 2-4 sites");
        for (int i = 0; i < d->nrsites; ++i)
            for (int k = 0; k < 3; ++k) in.ss[i][k] = symsecs_from(d->ss[i][k]);
        in.qnumbers.assign(d->qnumbers, d->qnumbers + (size_t)d->nrblocks * d->nrsites);
        in.beginblock.assign(d->beginblock, d->beginblock + d->nrblocks + 1);
        in.sel.minD = d->minD;
        in.sel.maxD = d->maxD;
        in.sel.truncerr = d->truncerr;
        in.sel.truncType = d->truncType;
        const std::vector<double> A(d->tel, d->tel + in.beginblock[d->nrblocks]);
        const SplitPlan plan = plan_split(in);
        auto res = std::make_unique<t3ns_split>();
        RenormWork w;
        res->r = run_split(in, plan, A, hipStreamPerThread, &w);
        for (const auto& ir : res->r.bond.irreps) res->irreps.insert(res->irreps.end(), ir.begin(), ir.end());
        for (const auto& v : res->r.sing) {
            res->nsing.push_back((int)v.size());
            res->sing.insert(res->sing.end(), v.begin(), v.end());
        }
        if (work) {
            work->flops = w.flops;
            work->bytes = w.bytes;
            work->kernel_ms = w.kernel_ms;
        }
        *out = res.release();
    });
}

int t3ns_split_get(const t3ns_split* s, t3ns_split_view* v) {
    if (!s || !v) return fail(T3NS_ERR_INVALID, "null argument");
    const SplitResult& r = s->r;
    v->U_nrsites = r.U.nrsites;
    v->U_nrblocks = r.U.nrblocks;
    v->U_qnumbers = r.U.qnumbers.data();
    v->U_beginblock = r.U.beginblock.data();
    v->U_tel = r.U.tel.data();
    v->V_nrblocks = r.V.nrblocks;
    v->V_qnumbers = r.V.qnumbers.data();
    v->V_beginblock = r.V.beginblock.data();
    v->V_tel = r.V.tel.data();
    v->bond_nrSecs = r.bond.nrSecs;
    v->bond_irreps = s->irreps.data();
    v->bond_dims = r.bond.dims.data();
    v->nrsectors = (int)r.kept.size();
    v->nsing = s->nsing.data();
    v->kept = r.kept.data();
    v->sing = s->sing.data();
    for (int i = 0; i < 2; ++i) {
        v->norm[i] = r.res.norm[i];
        v->entropy[i] = r.res.entropy[i];
    }
    v->norm_U = r.norm_U;
    v->sweeps = r.sweeps;
    return T3NS_OK;
}

int t3ns_split_destroy(t3ns_split* s) {
    delete s;
    return T3NS_OK;
}

struct t3ns_qr {
    QRResult r;
    std::vector<int> nsing;
    std::vector<double> sing;
};

extern "C++" {  // C++ helper inside the C-ABI block
namespace {
SiteTensor site_from(int nrblocks, const int64_t* qn, const int64_t* bb, const double* tel) {
    if (nrblocks < 0 || (nrblocks && (!qn || !bb || !tel))) throw std::invalid_argument("bad site tensor");
    SiteTensor t;
    t.nrsites = 1;
    t.nrblocks = nrblocks;
    t.qnumbers.assign(qn, qn + nrblocks);
    t.beginblock.assign(1, 0);
    if (nrblocks) t.beginblock.assign(bb, bb + nrblocks + 1);
    t.tel.assign(tel, tel + t.size());
    return t;
}
}  // namespace
}  // extern "C++"

int t3ns_qr_step(const t3ns_qr_desc* d, t3ns_qr** out, t3ns_work* work) {
    if (!d || !out) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        QRInput in;
        in.sym.nrSyms = d->nrSyms;
        for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) in.sym.sgs[i] = d->sgs[i];
        for (int k = 0; k < 3; ++k) {
            in.ssA[k] = symsecs_from(d->ssA[k]);
            in.ssC[k] = symsecs_from(d->ssC[k]);
        }
        const SiteTensor A = site_from(d->A_nrblocks, d->A_qnumbers, d->A_beginblock, d->A_tel);
        const SiteTensor C = site_from(d->C_nrblocks, d->C_qnumbers, d->C_beginblock, d->C_tel);
        in.A = &A;
        in.C = &C;
        const QRPlan plan = plan_qr(in);
        auto res = std::make_unique<t3ns_qr>();
        RenormWork w;
        res->r = run_qr(in, plan, hipStreamPerThread, &w);
        for (const auto& v : res->r.sing) {
            res->nsing.push_back((int)v.size());
            res->sing.insert(res->sing.end(), v.begin(), v.end());
        }
        if (work) {
            work->flops = w.flops;
            work->bytes = w.bytes;
            work->kernel_ms = w.kernel_ms;
        }
        *out = res.release();
    });
}

int t3ns_qr_get(const t3ns_qr* q, t3ns_qr_view* v) {
    if (!q || !v) return fail(T3NS_ERR_INVALID, "null argument");
    v->nrblocks_Q = q->r.Q.nrblocks;
    v->Q_beginblock = q->r.Q.beginblock.data();
    v->Q_tel = q->r.Q.tel.data();
    v->nrblocks_B = q->r.B.nrblocks;
    v->B_beginblock = q->r.B.beginblock.data();
    v->B_tel = q->r.B.tel.data();
    v->nrsectors = (int)q->r.dims.size();
    v->dims = q->r.dims.data();
    v->nsing = q->nsing.data();
    v->sing = q->sing.data();
    v->entropy = q->r.entropy;
    return T3NS_OK;
}

int t3ns_qr_destroy(t3ns_qr* q) {
    delete q;
    return T3NS_OK;
}

int t3ns_make_site_tensor(const t3ns_makesite_desc* d, double* tel, t3ns_work* work) {
    if (!d || !tel) return fail(T3NS_ERR_INVALID, "null argument");
    return guarded([&] {
        MakesiteInput in;
        in.nrsites = d->nrsites;
        in.innersite = d->innersite;
        if (d->nrsites < 2 || d->nrsites > 4 || d->nrblocks < 0)
            throw std::invalid_argument("makesite: needs 2-4 sites");
        SiteTensor sites[4];
        for (int i = 0; i < d->nrsites; ++i) {
            for (int k = 0; k < 3; ++k) {
                in.new_ss[i][k] = symsecs_from(d->new_ss[i][k]);
                in.old_ss[i][k] = symsecs_from(d->old_ss[i][k]);
            }
            SiteTensor& t = sites[i];
            t.nrsites = 1;
            t.nrblocks = d->site_nrblocks[i];
            if (t.nrblocks < 0 || (t.nrblocks && (!d->site_qnumbers[i] || !d->site_beginblock[i] ||
                                                  !d->site_tel[i])))
                throw std::invalid_argument("makesite: bad one-site tensor");
            t.qnumbers.assign(d->site_qnumbers[i], d->site_qnumbers[i] + t.nrblocks);
            t.beginblock.assign(1, 0);
            if (t.nrblocks) t.beginblock.assign(d->site_beginblock[i], d->site_beginblock[i] + t.nrblocks + 1);
            t.tel.assign(d->site_tel[i], d->site_tel[i] + t.size());
            in.sites[i] = &sites[i];
        }
        in.qnumbers.assign(d->qnumbers, d->qnumbers + (size_t)d->nrblocks * d->nrsites);
        in.beginblock.assign(d->beginblock, d->beginblock + d->nrblocks + 1);
        const MakesitePlan plan = plan_makesite(in);
        RenormWork w;
        const std::vector<double> out = run_makesite(plan, in.sites, hipStreamPerThread, &w);
        std::copy(out.begin(), out.end(), tel);
        if (work) {
            work->flops = w.flops;
            work->bytes = w.bytes;
            work->kernel_ms = w.kernel_ms;
        }
    });
}

int t3ns_rops_work(const t3ns_rops* r, double* flops, double* bytes, double* kernel_ms) {
    if (!r) return fail(T3NS_ERR_INVALID, "null argument");
    if (flops) *flops = r->work.flops;
    if (bytes) *bytes = r->work.bytes;
    if (kernel_ms) *kernel_ms = r->work.kernel_ms;
    return T3NS_OK;
}

int t3ns_rops_view(const t3ns_rops* r, t3ns_rops_desc* d) {
    if (!r || !d) return fail(T3NS_ERR_INVALID, "null argument");
    d->bond = r->r.bond;
    d->is_left = r->r.is_left;
    d->P_operator = r->r.P_operator;
    d->nrhss = r->r.nrhss;
    d->begin_blocks_of_hss = r->r.begin_blocks_of_hss.data();
    d->qnumbers = r->r.qnumbers.data();
    d->nrops = r->r.nrops;
    d->hss_of_ops = r->r.hss_of_ops.data();
    d->beginblock = r->bb_ptr.data();
    d->tel = r->tel_ptr.data();
    return T3NS_OK;
}

int t3ns_rops_destroy(t3ns_rops* r) {
    delete r;
    return T3NS_OK;
}

int t3ns_heff_destroy(t3ns_heff* h) {
    if (!h) return T3NS_OK;
    return guarded([&] {
        hipStream_t s = h->stream;
        delete h;
        stream_release(s);
    });
}

int64_t t3ns_heff_size(const t3ns_heff* h) { return h ? h->dev->size() : -1; }

int t3ns_heff_stats(const t3ns_heff* h, int64_t* nblocks, int64_t* npairs, int64_t* nterms,
                    double* flops) {
    if (!h) return fail(T3NS_ERR_INVALID, "null argument");
    const HeffStats& s = h->dev->stats();
    if (nblocks) *nblocks = s.nblocks;
    if (npairs) *npairs = s.npairs;
    if (nterms) *nterms = s.nterms;
    if (flops) *flops = s.flops;
    return T3NS_OK;
}

int t3ns_heff_kernel_stats(const t3ns_heff* h, double* out, int n) {
    if (!h || !out) return fail(T3NS_ERR_INVALID, "null argument");
    const HeffStats& s = h->dev->stats();
    const double v[] = {(double)s.nchunks, (double)s.nruns, (double)s.pairs_global_x,
                        (double)s.pairs_global_w, s.mfma_steps, s.stage_steps[0],
                        s.stage_steps[1], s.stage_steps[2], (double)s.scratch_per_wg,
                        (double)s.nunits, (double)s.nsplit_blocks, (double)s.pairs_fused,
                        (double)h->dev->wpc()};
    const int m = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < m; ++i) out[i] = v[i];
    return m;
}

int64_t t3ns_heff_plan_table(const t3ns_heff* h, int* ints, double* prefs) {
    if (!h) return -1;
    return write_plan_table(h->dev->plan(), ints, prefs);
}

int t3ns_plan_block_dims(const t3ns_heff_desc* d, const t3ns_rops_desc* ops,
                         const t3ns_site_desc* site, int* dims) {
    if (!d || !ops || !site || !dims) return fail(T3NS_ERR_INVALID, "null argument");
    int n = -1;
    const int rc = guarded([&] {
        HeffHost host;
        convert_heff_input(&host, d, ops, site);
        HeffPlan P = build_heff_plan(host.in);
        for (int b = 0; b < P.nrblocks; ++b)
            for (int k = 0; k < 3; ++k) dims[3 * b + k] = P.dims[b][k];
        n = P.nrblocks;
    });
    return rc != T3NS_OK ? rc : n;
}

int64_t t3ns_plan_table(const t3ns_heff_desc* d, const t3ns_rops_desc* ops,
                        const t3ns_site_desc* site, int* ints, double* prefs, double* flops) {
    if (!d || !ops || !site) return fail(T3NS_ERR_INVALID, "null argument");
    int64_t n = -1;
    const int rc = guarded([&] {
        HeffHost host;
        convert_heff_input(&host, d, ops, site);
        HeffPlan P = build_heff_plan(host.in);
        n = write_plan_table(P, ints, prefs);
        if (flops) *flops = P.flops;
    });
    return rc == T3NS_OK ? n : rc;
}

}  // extern "C"
