// tune_cache.cpp - see tune_cache.h.
#include "tune_cache.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <sys/stat.h>

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
    return os.str();
}

std::map<std::string, std::string> load() {
    std::map<std::string, std::string> m;
    std::ifstream f(cache_path());
    std::string line;
    while (std::getline(f, line)) {
        // key = first 7 fields, value = the rest
        std::istringstream is(line);
        std::string k[7], v;
        bool ok = true;
        for (auto& s : k)
            if (!(is >> s)) ok = false;
        if (!ok || !(is >> v)) continue;
        std::string key = k[0];
        for (int i = 1; i < 7; ++i) key += ' ' + k[i];
        m[key] = v;
    }
    return m;
}

}  // namespace

std::string kernel_signature(const ofx_kernel_info& k) {
    std::ostringstream os;
    os << k.kind << ':' << k.n << ':' << k.threads_per_transform << ':' << k.transforms_per_block << ':' << k.staged
       << ':' << k.iov << ':';
    for (int p = 0; p < k.npass; ++p) os << (p ? "x" : "") << k.radices[p];
    return os.str();
}

void apply_tune_cache(const PlanDesc& d, std::vector<Step>& steps, const ofx_kernel_info* reg, int nreg) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto m = load();
    if (m.empty()) return;
    for (size_t i = 0; i < steps.size(); ++i) {
        Step& st = steps[i];
        if (st.type != Step::KERNEL || st.candidates.size() < 2) continue;  // row kernels only
        auto it = m.find(key_of(d, i));
        if (it == m.end()) continue;
        for (size_t c = 0; c < st.candidates.size(); ++c) {
            const int id = st.candidates[c];
            if (id >= 0 && id < nreg && kernel_signature(reg[id]) == it->second) {
                st.kernel = id;
                st.tw_table = st.candidate_tables[c];
                break;
            }
        }
    }
}

void store_tune_cache(const PlanDesc& d, const std::vector<Step>& steps, const ofx_kernel_info* reg) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto m = load();
    for (size_t i = 0; i < steps.size(); ++i) {
        if (steps[i].type != Step::KERNEL || steps[i].candidates.size() < 2) continue;
        m[key_of(d, i)] = kernel_signature(reg[steps[i].kernel]);
    }
    const std::string path = cache_path();
    // create the directory chain of the default location
    std::string dir = path.substr(0, path.find_last_of('/'));
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
    std::ofstream f(path, std::ios::trunc);
    for (auto& kv : m) f << kv.first << ' ' << kv.second << '\n';
}

}  // namespace ofx
