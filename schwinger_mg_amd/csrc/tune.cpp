// tune.cpp -- the autotuner's table (host only; the timing searches live beside the
// objects they time, in runtime_b.cpp).
#include "tune.h"

#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "smg_common.h"

namespace smg {

namespace {
std::mutex g_tune_mu;
std::map<TuneKey, int>& table() {
    static auto* t = new std::map<TuneKey, int>();
    return *t;
}
bool valid(const TuneKey& k, int choice) {
    if (k[0] < 0 || k[0] >= TUNE_FAMILIES || choice < 0) return false;
    if (k[0] == TUNE_COARSE_BF16) return (choice >> 1) < kTuneNGrids;
    if (k[0] == TUNE_DSLASH_STORE) return choice <= 1;
    if (k[0] == TUNE_TRANSFER) return choice <= 2;
    if (k[0] == TUNE_DDAGD) return choice <= 2 * kTuneNGrids;   // 0, or 1 + 2 g + nt
    return choice < kTuneNGrids;
}
}  // namespace

int tune_lookup(const TuneKey& k) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = table().find(k);
    return it == table().end() ? -1 : it->second;
}
void tune_store(const TuneKey& k, int choice) {
    SMG_REQUIRE(valid(k, choice), "tune: no such family or choice");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    table()[k] = choice;
}
void tune_clear() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    table().clear();
}
size_t tune_size() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return table().size();
}
void tune_save(const std::string& path) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::ofstream f(path);
    SMG_REQUIRE(f.good(), "tune: cannot write " + path);
    f << "# smg launch-shape table: family a b c d e choice\n";
    for (const auto& kv : table()) {
        for (int w : kv.first) f << w << ' ';
        f << kv.second << '\n';
    }
    SMG_REQUIRE(f.good(), "tune: write failed on " + path);
}
void tune_load(const std::string& path) {
    std::ifstream f(path);
    SMG_REQUIRE(f.good(), "tune: cannot read " + path);
    std::map<TuneKey, int> add;
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        lineno++;
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        TuneKey k;
        int choice = -1;
        for (int& w : k) is >> w;
        is >> choice;
        std::string rest;
        SMG_REQUIRE(!is.fail() && !(is >> rest) && valid(k, choice),
                    "tune: bad entry at " + path + ":" + std::to_string(lineno));
        add[k] = choice;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto& kv : add) table()[kv.first] = kv.second;
}

}  // namespace smg
