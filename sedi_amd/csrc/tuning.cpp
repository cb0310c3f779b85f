// tuning.cpp -- the autotuner's table (see tuning.h).  Format, one entry per
// line, '#' starts a comment:
//     <key> <n1> <n2> <n3> <value>
#include "tuning.h"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <vector>

namespace sedi {

namespace {

struct Entry {
    std::string key;
    long long n1, n2, n3;
    int value;
};

std::once_flag g_once;
std::vector<Entry> g_table;
std::string g_path;

std::string default_path() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&tuned_value), &info) && info.dli_fname) {
        std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/tuning.txt";
    }
    return "";
}

void load() {
    const char* env = std::getenv("SEDI_TUNING_FILE");
    const std::string path = env ? std::string(env) : default_path();
    if (path.empty()) return;
    std::ifstream f(path);
    if (!f) return;
    g_path = path;
    std::string line;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream in(line);
        Entry e;
        if (in >> e.key >> e.n1 >> e.n2 >> e.n3 >> e.value) g_table.push_back(e);
    }
}

const Entry* find(const std::string& key, const Shape& s, bool any_grid) {
    for (const Entry& e : g_table) {
        if (e.key != key) continue;
        if (any_grid ? (e.n1 == 0 && e.n2 == 0 && e.n3 == 0)
                     : (e.n1 == s.n1 && e.n2 == s.n2 && e.n3 == s.n3))
            return &e;
    }
    return nullptr;
}

}  // namespace

const std::string& tuning_path() {
    std::call_once(g_once, load);
    return g_path;
}

int tuned_value(const std::string& key, const Shape& s, int fallback, bool nearest) {
    std::call_once(g_once, load);
    if (g_table.empty()) return fallback;
    // per-kernel keys ("zc:<kernel>", "sched:<kernel>", "wpc:<kernel>") fall back to "<kind>:*"
    const size_t colon = key.find(':');
    const bool per_kernel = key.compare(0, 3, "zc:") == 0 || key.compare(0, 6, "sched:") == 0 ||
                            key.compare(0, 4, "wpc:") == 0;
    const std::string wild = per_kernel ? key.substr(0, colon + 1) + "*" : key;
    for (const std::string& k : {key, wild}) {
        if (const Entry* e = find(k, s, false)) return e->value;
        if (const Entry* e = find(k, s, true)) return e->value;
    }
    if (nearest) {
        const double n = (double)s.n1 * s.n2 * s.n3;
        const Entry* best = nullptr;
        double bd = 0;
        for (const Entry& e : g_table) {
            if (e.key != key || e.n1 <= 0) continue;
            const double d = std::fabs(std::log((double)e.n1 * e.n2 * e.n3 / n));
            if (!best || d < bd) {
                best = &e;
                bd = d;
            }
        }
        if (best) return best->value;
    }
    return fallback;
}

}  // namespace sedi
