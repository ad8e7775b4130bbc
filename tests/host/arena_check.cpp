// Randomised check of sun::Arena against a byte-map model: blocks never overlap,
// stay inside the capacity, are 256-byte aligned, in_use() is the sum of the live
// blocks, first fit is honoured, and after every block is released the arena is
// one free block of the whole capacity again (neighbours merged).
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "../../sun_amd/csrc/sun_arena.hpp"

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);       \
            return 1;                                                     \
        }                                                                 \
    } while (0)

int main() {
    const size_t cap = 1 << 20;
    sun::Arena a;
    a.reset(cap + 100);  // the unaligned tail is dropped
    CHECK(a.capacity() == cap);
    CHECK(a.largest_free() == cap);
    std::mt19937_64 rng(12345);
    std::map<size_t, size_t> live;  // offset -> rounded length
    std::vector<unsigned char> owner(cap / 256, 0);
    for (int it = 0; it < 200000; it++) {
        const bool do_alloc = live.empty() || (rng() % 100) < 55;
        if (do_alloc) {
            const size_t n = 1 + rng() % (cap / 16);
            const size_t r = sun::Arena::round_up(n);
            const size_t off = a.alloc(n);
            // the model's first fit: lowest offset with r free bytes
            size_t model = SIZE_MAX, run = 0;
            for (size_t u = 0; u < owner.size(); u++) {
                run = owner[u] ? 0 : run + 1;
                if (run * 256 >= r) {
                    model = (u + 1) * 256 - run * 256;
                    break;
                }
            }
            if (off == SIZE_MAX) {
                CHECK(model == SIZE_MAX);
                continue;
            }
            CHECK(off == model);
            CHECK(off % 256 == 0 && off + r <= cap);
            for (size_t u = off / 256; u < (off + r) / 256; u++) {
                CHECK(owner[u] == 0);
                owner[u] = 1;
            }
            live[off] = r;
        } else {
            auto it2 = live.begin();
            std::advance(it2, rng() % live.size());
            CHECK(a.release(it2->first));
            CHECK(!a.release(it2->first));  // double release is refused
            for (size_t u = it2->first / 256; u < (it2->first + it2->second) / 256; u++) owner[u] = 0;
            live.erase(it2);
        }
        size_t sum = 0;
        for (auto& kv : live) sum += kv.second;
        CHECK(a.in_use() == sum);
        CHECK(a.blocks() == live.size());
    }
    for (auto& kv : live) CHECK(a.release(kv.first));
    CHECK(a.in_use() == 0);
    CHECK(a.largest_free() == cap);
    CHECK(a.alloc(cap) == 0);
    CHECK(a.alloc(1) == SIZE_MAX);
    std::printf("OK\n");
    return 0;
}
