// tune.cpp — the tile table: rows per pass tile by instance shape class.
//
// The fused pass works on tiles of ti_rows x 64 fibers. The best height
// depends on z (fiber registers, LDS per wave) and on the instance size
// (tiles per CU, partial-sum traffic ~ 1/ti_rows of S). The autotuner
// (bnmf_hip.autotune) times candidate heights per class on the device and
// stores the winners here; bnmf_tune_save/load keep them in a text file,
// and the library loads bnmf-algs_amd/tune/gfx950.txt (or $BNMF_TUNE_FILE)
// on first use. A class without an entry uses the built-in rule.
//
// The table is read when an instance is planned and depends on its shape
// only, so a given table gives every instance the same tiles (and the same
// summation order) alone, batched or served.

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>

#include "bnmf_internal.hpp"

namespace bnmf {

namespace {

constexpr int kZClasses = 8;
constexpr int kSClasses = 5;
// upper z of each z class
constexpr uint32_t kZHi[kZClasses] = {4, 8, 12, 16, 24, 32, 48, 64};

struct TileTable {
    std::mutex mu;
    uint32_t cap[2][kZClasses][kSClasses] = {};
    std::string source;
};

TileTable& table() {
    static TileTable t;
    return t;
}

std::string default_path() {
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&default_path), &info) == 0 || !info.dli_fname) return "";
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
    return dir + "/../tune/gfx950.txt";
}

// Parses "f32|f64 zclass sclass cap" lines; returns false if unreadable.
bool load_locked(TileTable& t, const std::string& path, std::string* err) {
    std::ifstream in(path);
    if (!in) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    uint32_t cap[2][kZClasses][kSClasses] = {};
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ls(line);
        std::string dt;
        int zc = -1, sc = -1;
        long v = -1;
        if (!(ls >> dt)) continue;
        if (!(ls >> zc >> sc >> v) || (dt != "f32" && dt != "f64") || zc < 0 || zc >= kZClasses ||
            sc < 0 || sc >= kSClasses || v < 4 || v > 1024) {
            if (err) *err = path + ":" + std::to_string(lineno) + ": bad entry";
            return false;
        }
        cap[dt == "f32" ? 0 : 1][zc][sc] = (uint32_t)v;
    }
    std::memcpy(t.cap, cap, sizeof(cap));
    t.source = path;
    return true;
}

void ensure_loaded(TileTable& t) {
    static std::once_flag once;
    std::call_once(once, [&t] {
        const char* env = std::getenv("BNMF_TUNE_FILE");
        const std::string path = env ? std::string(env) : default_path();
        if (!path.empty()) (void)load_locked(t, path, nullptr);
    });
}

}  // namespace

int tune_zclass(uint32_t z) {
    for (int c = 0; c < kZClasses; ++c)
        if (z <= kZHi[c]) return c;
    return kZClasses - 1;
}

int tune_sclass(uint64_t fibers) {
    if (fibers < (1ull << 14)) return 0;
    if (fibers < (1ull << 16)) return 1;
    if (fibers < (1ull << 18)) return 2;
    if (fibers < (1ull << 20)) return 3;
    return 4;
}

uint32_t tune_ti_cap(int dtype, uint32_t z, uint64_t fibers) {
    TileTable& t = table();
    ensure_loaded(t);
    std::lock_guard<std::mutex> lk(t.mu);
    return t.cap[dtype == BNMF_F32 ? 0 : 1][tune_zclass(z)][tune_sclass(fibers)];
}

void tune_set(int dtype, uint32_t z, uint64_t fibers, uint32_t cap) {
    TileTable& t = table();
    ensure_loaded(t);
    std::lock_guard<std::mutex> lk(t.mu);
    t.cap[dtype == BNMF_F32 ? 0 : 1][tune_zclass(z)][tune_sclass(fibers)] = cap;
}

void tune_load(const std::string& path) {
    TileTable& t = table();
    ensure_loaded(t);
    std::lock_guard<std::mutex> lk(t.mu);
    std::string err;
    if (!load_locked(t, path, &err)) throw Error(BNMF_ERR_INVALID_ARGUMENT, err);
}

void tune_save(const std::string& path) {
    TileTable& t = table();
    ensure_loaded(t);
    std::lock_guard<std::mutex> lk(t.mu);
    std::ofstream out(path);
    if (!out) throw Error(BNMF_ERR_INVALID_ARGUMENT, "cannot write " + path);
    out << "# rows per pass tile: dtype zclass sclass rows\n";
    out << "# zclass upper z: 4 8 12 16 24 32 48 64; sclass fibers x*y below 2^14 2^16 2^18 2^20, then above\n";
    for (int d = 0; d < 2; ++d)
        for (int zc = 0; zc < kZClasses; ++zc)
            for (int sc = 0; sc < kSClasses; ++sc)
                if (t.cap[d][zc][sc]) out << (d == 0 ? "f32 " : "f64 ") << zc << ' ' << sc << ' ' << t.cap[d][zc][sc] << '\n';
    if (!out) throw Error(BNMF_ERR_DEVICE, "write failed: " + path);
}

std::string tune_source() {
    TileTable& t = table();
    ensure_loaded(t);
    std::lock_guard<std::mutex> lk(t.mu);
    return t.source;
}

}  // namespace bnmf
