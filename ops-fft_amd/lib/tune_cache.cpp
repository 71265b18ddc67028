// tune_cache.cpp - see tune_cache.h.
#include "tune_cache.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <sys/stat.h>
#include <unistd.h>

namespace ofx {

namespace {

std::mutex g_mu;

std::string cache_path() {
    const char* p = std::getenv("ACLFFT_TUNE_CACHE");
    if (p && *p) return p;
    const char* h = std::getenv("HOME");
    return std::string(h && *h ? h : "/tmp") + "/.cache/ops_fft_amd/tune.txt";
}

std::string key_of(const PlanDesc& d, size_t step) {
    std::ostringstream os;
    os << (int)d.type << ' ' << d.rank << ' ' << d.n0 << ' ' << d.n1 << ' ' << d.n2 << ' ' << d.batch << ' '
       << (d.vertical ? 1 : 0) << ' ' << step;
    if (d.uneven) {
        // uneven plans: the groups, as an FNV-1a hash of the (length, batch) list
        unsigned long long h = 1469598103934665603ULL;
        auto mix = [&h](long long v) {
            for (int b = 0; b < 8; ++b) {
                h ^= (unsigned long long)((v >> (8 * b)) & 0xff);
                h *= 1099511628211ULL;
            }
        };
        for (size_t g = 0; g < d.ulen.size(); ++g) {
            mix(d.ulen[g]);
            mix(d.ubatch[g]);
        }
        os << " u" << std::hex << h;
    }
    return os.str();
}

std::map<std::string, std::string> load() {
    std::map<std::string, std::string> m;
    std::ifstream f(cache_path());
    std::string line;
    while (std::getline(f, line)) {
        // "<key fields> <kernel signature>": the signature is the last field
        const size_t sp = line.find_last_of(' ');
        if (sp == std::string::npos || sp == 0 || sp + 1 >= line.size()) continue;
        m[line.substr(0, sp)] = line.substr(sp + 1);
    }
    return m;
}

}  // namespace

std::string kernel_signature(const ofx_kernel_info& k) {
    std::ostringstream os;
    os << k.kind << ':' << k.n << ':' << k.threads_per_transform << ':' << k.transforms_per_block << ':' << k.staged
       << ':' << k.iov << ':';
    for (int p = 0; p < k.npass; ++p) os << (p ? "x" : "") << k.radices[p];
    if (k.stream) os << ":s" << (k.stream == 1 ? std::string() : std::to_string(k.stream));
    return os.str();
}

namespace {

// the registry entry a step's kernel index refers to, or null
const ofx_kernel_info* entry_of(const Step& st, int id, const Registries& reg) {
    if (st.type == Step::KERNEL) return (id >= 0 && id < reg.nrow) ? &reg.row[id] : nullptr;
    if (st.type == Step::COLUMN) return (id >= 0 && id < reg.ncol) ? &reg.col[id] : nullptr;
    return nullptr;
}

// Signature of candidate `id` of a step ("" when it names no entry).  The LDS footprint tells
// apart configurations that differ only in their LDS layout (the column kernels' split exchange,
// the row kernels' pad shift); MFMA steps are named by their factorisation and MFMA tile.
std::string candidate_signature(const Step& st, int id, const Registries& reg) {
    if (st.type == Step::MFMA) {
        ofx_mfma_info mi;
        if (ofx_mfma_kernel_info_get(id, &mi) != 0) return std::string();
        return "mfma:" + std::to_string(mi.n) + ":" + std::to_string(mi.r1) + "x" + std::to_string(mi.r2) + ":" +
               std::to_string(mi.shape) + (mi.tiled ? ":t" : "") + ":m" + std::to_string(st.mfma_mode);
    }
    const ofx_kernel_info* k = entry_of(st, id, reg);
    if (k == nullptr) return std::string();
    return kernel_signature(*k) + (st.type == Step::COLUMN ? ":c" : ":l") + std::to_string(k->lds_bytes);
}

bool tunable(const Step& st) {
    return (st.type == Step::KERNEL || st.type == Step::COLUMN || st.type == Step::MFMA) && st.candidates.size() >= 2;
}

}  // namespace

// write the whole cache to a private file, then rename it over the cache (called with g_mu held)
static void write_cache(const std::map<std::string, std::string>& m) {
    const std::string path = cache_path();
    // create the directory chain of the location (none for a bare file name)
    const size_t slash = path.find_last_of('/');
    std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash);
    std::string acc;
    std::istringstream parts(dir);
    std::string part;
    while (std::getline(parts, part, '/')) {
        if (part.empty()) {
            acc = "/";
            continue;
        }
        acc += (acc.empty() || acc.back() == '/' ? "" : "/") + part;
        mkdir(acc.c_str(), 0755);
    }
    // write a private file and rename it over the cache, so that a reader (or another process
    // writing at the same time) never sees a partly written file
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    {
        std::ofstream f(tmp, std::ios::trunc);
        for (auto& kv : m) f << kv.first << ' ' << kv.second << '\n';
        if (!f.good()) {
            std::remove(tmp.c_str());
            return;
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

void apply_tune_cache(const PlanDesc& d, std::vector<Step>& steps, const Registries& reg) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto m = load();
    if (m.empty()) return;
    for (size_t i = 0; i < steps.size(); ++i) {
        Step& st = steps[i];
        if (!tunable(st)) continue;
        auto it = m.find(key_of(d, i));
        if (it == m.end()) continue;
        for (size_t c = 0; c < st.candidates.size(); ++c) {
            const std::string sig = candidate_signature(st, st.candidates[c], reg);
            if (!sig.empty() && sig == it->second) {
                st.kernel = st.candidates[c];
                st.tw_table = st.candidate_tables[c];
                break;
            }
        }
    }
}

void store_tune_cache(const PlanDesc& d, const std::vector<Step>& steps, const Registries& reg) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto m = load();
    for (size_t i = 0; i < steps.size(); ++i) {
        if (!tunable(steps[i])) continue;
        const std::string sig = candidate_signature(steps[i], steps[i].kernel, reg);
        if (!sig.empty()) m[key_of(d, i)] = sig;
    }
    write_cache(m);
}

namespace {

// the plan-level key: the shape (as for steps, the group hash for uneven plans) and the requested
// precision (automatic or a fixed mode)
std::string plan_key(const PlanDesc& d) {
    return key_of(d, 0) + (d.auto_precision ? std::string(" plan auto") : " plan p" + std::to_string(d.precision));
}

}  // namespace

bool load_plan_choice(const PlanDesc& d, PlanDesc* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto m = load();
    auto it = m.find(plan_key(d));
    if (it == m.end()) return false;
    // "form:sched:bpc:set:precision:stream:prologue" (older entries end early: the defaults)
    int v[7] = {0, -1, 0, 0, 0, -1, -1};
    const int got =
        std::sscanf(it->second.c_str(), "%d:%d:%d:%d:%d:%d:%d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6]);
    if (got < 5) return false;
    *out = d;
    out->uneven_form = v[0];
    out->sched = v[1];
    out->blocks_per_cu = v[2];
    out->class_set = v[3];
    out->precision = v[4];
    out->stream = v[5];
    out->prologue = v[6];
    return true;
}

void store_plan_choice(const PlanDesc& c) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto m = load();
    m[plan_key(c)] = std::to_string(c.uneven_form) + ":" + std::to_string(c.sched) + ":" +
                     std::to_string(c.blocks_per_cu) + ":" + std::to_string(c.class_set) + ":" +
                     std::to_string(c.precision) + ":" + std::to_string(c.stream) + ":" + std::to_string(c.prologue);
    write_cache(m);
}

}  // namespace ofx
