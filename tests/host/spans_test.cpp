// Host check of eik::span_union (eik_spans.h), the busy time of a call's round launches over one or
// two streams: hand-computed cases, then random sets of integer intervals against a brute-force
// count of the unit cells they cover.  Prints "ok" or the first mismatch.
#include <cstdio>
#include <random>
#include <vector>

#include "eik_spans.h"

using V = std::vector<std::pair<float, float>>;

static int check(const char* name, const V& v, double want) {
    const double got = eik::span_union(v);
    if (got != want) {
        std::printf("%s: got %.9g want %.9g\n", name, got, want);
        return 1;
    }
    return 0;
}

int main() {
    int bad = 0;
    bad += check("empty", {}, 0.0);
    bad += check("single", {{1.0f, 3.5f}}, 2.5);
    bad += check("disjoint", {{0.0f, 1.0f}, {2.0f, 4.0f}}, 3.0);
    bad += check("overlap", {{0.0f, 2.0f}, {1.0f, 3.0f}}, 3.0);
    bad += check("nested", {{0.0f, 10.0f}, {2.0f, 3.0f}, {4.0f, 9.0f}}, 10.0);
    bad += check("touching", {{0.0f, 1.0f}, {1.0f, 2.0f}}, 2.0);
    bad += check("unsorted", {{5.0f, 6.0f}, {0.0f, 1.0f}, {0.5f, 2.0f}}, 3.0);
    bad += check("two streams", {{0.0f, 4.0f}, {6.0f, 8.0f}, {3.0f, 7.0f}}, 8.0);
    bad += check("reversed counts empty", {{3.0f, 1.0f}, {5.0f, 6.0f}}, 1.0);
    bad += check("zero length", {{2.0f, 2.0f}}, 0.0);
    std::mt19937 rng(20261);
    for (int trial = 0; trial < 20000 && !bad; trial++) {
        const int n = (int)(rng() % 12);
        V v;
        std::vector<int> cell(72, 0);  // intervals end at 59 + 7 at most
        for (int k = 0; k < n; k++) {
            const int a = (int)(rng() % 60), len = (int)(rng() % 8);
            v.emplace_back((float)a, (float)(a + len));
            for (int c = a; c < a + len; c++) cell[c] = 1;
        }
        double want = 0.0;
        for (int c : cell) want += c;
        char name[32];
        std::snprintf(name, sizeof(name), "random %d", trial);
        bad += check(name, v, want);
    }
    if (!bad) std::printf("ok\n");
    return bad ? 1 : 0;
}
