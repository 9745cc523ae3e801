// Device-resident operator arenas (see resident_ops.hpp).
#include <hip/hip_runtime.h>

#include "resident_ops.hpp"

#include <cstdlib>
#include <list>
#include <mutex>

namespace t3 {

namespace {

struct Entry {
    const double* host;
    int64_t n;
    int device;
    std::shared_ptr<DevBuf<double>> dev;
};

std::mutex g_mu;
std::list<Entry>& entries() {  // oldest first
    static std::list<Entry> l;
    return l;
}
int64_t g_bytes = 0;

int64_t cap_bytes() {
    static const int64_t cap = [] {
        double gb = 32;
        if (const char* e = std::getenv("T3NS_RESIDENT_GB")) gb = std::atof(e);
        return (int64_t)(gb * (1 << 30));
    }();
    return cap;
}

int current_device() {
    int d = 0;
    T3_HIP_CHECK(hipGetDevice(&d));
    return d;
}

}  // namespace

void resident_put(const double* host, int64_t n, std::shared_ptr<DevBuf<double>> dev) {
    const int64_t bytes = n * (int64_t)sizeof(double);
    if (!host || n <= 0 || !dev || bytes > cap_bytes()) return;
    const int d = current_device();
    std::lock_guard<std::mutex> lk(g_mu);
    auto& l = entries();
    while (!l.empty() && g_bytes + bytes > cap_bytes()) {
        g_bytes -= l.front().n * (int64_t)sizeof(double);
        l.pop_front();
    }
    l.push_back({host, n, d, std::move(dev)});
    g_bytes += bytes;
}

void resident_drop(const double* host) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto& l = entries();
    for (auto it = l.begin(); it != l.end(); ++it)
        if (it->host == host) {
            g_bytes -= it->n * (int64_t)sizeof(double);
            l.erase(it);
            return;
        }
}

ResidentRef resident_find(const ROperators& r) {
    ResidentRef out;
    if ((int)r.tel_src.size() != r.nrops || r.total_size() <= 0) return out;
    int first = -1;
    for (int o = 0; o < r.nrops && first < 0; ++o)
        if (r.op_offset[o + 1] > r.op_offset[o]) first = o;
    if (first < 0 || !r.tel_src[first]) return out;
    const double* p0 = r.tel_src[first] - r.op_offset[first];  // where operator 0 would start
    std::lock_guard<std::mutex> lk(g_mu);
    if (entries().empty()) return out;
    const int d = current_device();
    for (const Entry& e : entries()) {
        if (e.device != d || p0 < e.host || p0 + r.total_size() > e.host + e.n) continue;
        bool same = true;  // every operator where the layout puts it
        for (int o = 0; o < r.nrops && same; ++o)
            if (r.op_offset[o + 1] > r.op_offset[o]) same = r.tel_src[o] == p0 + r.op_offset[o];
        if (!same) continue;
        out.buf = e.dev;
        out.offset = p0 - e.host;
        return out;
    }
    return out;
}

}  // namespace t3
