// CPU unit test of the serving HBM arena (bnmf-algs_amd/csrc/arena.hpp):
// first-fit placement, 256-byte rounding, merging of freed neighbours and
// exhaustion, checked against a brute-force occupancy map.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "arena.hpp"

#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::printf("FAILED %s at line %d\n", #c, __LINE__);    \
            return 1;                                               \
        }                                                           \
    } while (0)

int main() {
    using bnmf::Arena;
    {
        Arena a(1000);  // rounds up to 1024
        CHECK(a.capacity() == 1024);
        size_t o1, o2, o3, o4;
        CHECK(a.alloc(1, &o1) && o1 == 0);
        CHECK(a.alloc(300, &o2) && o2 == 256);
        CHECK(a.alloc(256, &o3) && o3 == 768);
        CHECK(!a.alloc(1, &o4));
        CHECK(!a.fits(1));
        a.release(o2, 300);
        CHECK(a.fits(512) && !a.fits(513));
        a.release(o1, 1);  // merges with the hole at 256
        CHECK(a.free_blocks() == 1 && a.fits(768));
        a.release(o3, 256);
        CHECK(a.used() == 0 && a.free_blocks() == 1 && a.fits(1024));
    }
    {
        // random alloc/free against a byte map
        const size_t cap = 256 * 4096;
        Arena a(cap);
        std::vector<int> owner(cap / 256, -1);
        std::vector<std::pair<size_t, size_t>> live;
        std::mt19937 rng(7);
        for (int step = 0; step < 20000; ++step) {
            if (live.empty() || rng() % 3) {
                const size_t n = 1 + rng() % (256 * 40);
                size_t off;
                const bool ok = a.alloc(n, &off);
                // brute force: is there a free run of the rounded size?
                const size_t units = (n + 255) / 256;
                bool exists = false;
                size_t first = 0;
                for (size_t u = 0, run = 0; u < owner.size(); ++u) {
                    run = owner[u] < 0 ? run + 1 : 0;
                    if (run == units) {
                        exists = true;
                        first = u + 1 - units;
                        break;
                    }
                }
                CHECK(ok == exists);
                if (ok) {
                    CHECK(off == first * 256);  // first fit
                    for (size_t u = 0; u < units; ++u) {
                        CHECK(owner[off / 256 + u] < 0);
                        owner[off / 256 + u] = step;
                    }
                    live.push_back({off, n});
                }
            } else {
                const size_t k = rng() % live.size();
                a.release(live[k].first, live[k].second);
                for (size_t u = 0; u < (live[k].second + 255) / 256; ++u) owner[live[k].first / 256 + u] = -1;
                live.erase(live.begin() + k);
            }
            size_t used = 0;
            for (int o : owner) used += o >= 0 ? 256 : 0;
            CHECK(a.used() == used);
        }
    }
    std::printf("arena ok\n");
    return 0;
}
