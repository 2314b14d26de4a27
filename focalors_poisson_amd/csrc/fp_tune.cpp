// Record of the autotuners' choices (chunk schedule of the x sweeps, thread
// counts of the y transforms): kept for the process and, when the environment
// variable FP_TUNE_FILE names a file, read from it at the first lookup and
// appended to at every new choice, so that a later process (a service
// restart, a profiled run whose timings the profiler would distort) takes the
// same choices without timing anything.  One line per choice:
//   <key>|<value> <value> ...
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "fp_internal.h"

namespace fp {

namespace {
std::mutex                                   g_mu;
std::map<std::string, std::vector<int>>      g_rec;
bool                                         g_loaded = false;

const char* tune_file()
{
    const char* f = std::getenv("FP_TUNE_FILE");
    return (f && *f) ? f : nullptr;
}

void load_locked()
{
    if (g_loaded) return;
    g_loaded = true;
    const char* path = tune_file();
    if (!path) return;
    std::ifstream in(path);
    std::string   line;
    while (std::getline(in, line)) {
        const size_t bar = line.find('|');
        if (bar == std::string::npos) continue;
        std::istringstream vs(line.substr(bar + 1));
        std::vector<int>   v;
        int                x;
        while (vs >> x) v.push_back(x);
        if (!v.empty()) g_rec[line.substr(0, bar)] = v;
    }
}
}  // namespace

bool tune_lookup(const std::string& key, std::vector<int>& vals)
{
    std::lock_guard<std::mutex> lk(g_mu);
    load_locked();
    auto it = g_rec.find(key);
    if (it == g_rec.end()) return false;
    vals = it->second;
    return true;
}

void tune_store(const std::string& key, const std::vector<int>& vals)
{
    std::lock_guard<std::mutex> lk(g_mu);
    load_locked();
    g_rec[key] = vals;
    const char* path = tune_file();
    if (!path) return;
    std::ofstream out(path, std::ios::app);
    out << key << '|';
    for (size_t q = 0; q < vals.size(); q++) out << (q ? " " : "") << vals[q];
    out << '\n';
}

}  // namespace fp
