// Host-side planning code of the library (minijclsquant_amd/csrc/plan.cpp)
// under AddressSanitizer and UndefinedBehaviorSanitizer: the tile-plan
// breadth-first renumbering and its validator, the realization-batch
// row-chunk tables, and the wavefront plan and its validator (including
// corrupted plans it must refuse), on random, lattice-shaped and degenerate
// inputs.  Built
// and run by tests/test_sanitize.py; exits non-zero on the first failed
// check (the sanitizers abort on their own findings).
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "plan.h"

using namespace jkpm;

static int failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAIL %s:%d ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);             \
            std::fprintf(stderr, "\n");                    \
            ++failures;                                    \
        }                                                  \
    } while (0)

// slot-major [k][n] columns of a few graph shapes
static std::vector<int> random_cols(int n, int k, std::mt19937 &g) {
    std::vector<int> c((size_t)n * k);
    std::uniform_int_distribution<int> d(0, n - 1);
    for (auto &x : c) x = d(g);
    return c;
}

static std::vector<int> square_cols(int L, bool periodic) {
    const int n = L * L, k = 5;
    std::vector<int> c((size_t)n * k);
    for (int i = 0; i < n; ++i) {
        const int x = i % L, y = i / L;
        auto at = [&](int xx, int yy) {
            if (!periodic && (xx < 0 || xx >= L || yy < 0 || yy >= L)) return i;  // pad: self
            return ((yy + L) % L) * L + (xx + L) % L;
        };
        const int nb[5] = {i, at(x + 1, y), at(x - 1, y), at(x, y + 1), at(x, y - 1)};
        for (int s = 0; s < k; ++s) c[(size_t)s * n + i] = nb[s];
    }
    return c;
}

static void check_tiles(int n, int k, const std::vector<int> &c, int cap, const char *what) {
    TilePlanHost p;
    std::string err;
    const int rc = build_tile_plan(n, k, c, cap, 8, &p, &err);
    if (cap < 2 * k + 2) {
        CHECK(rc != 0, "%s: cap %d below 2k+2 accepted", what, cap);
        return;
    }
    CHECK(rc == 0, "%s: build failed: %s", what, err.c_str());
    if (rc) return;
    CHECK(validate_tile_plan(n, k, c, cap, p, &err) == 0, "%s: %s", what, err.c_str());
    CHECK(p.max_ext <= cap, "%s: max_ext %d > cap %d", what, p.max_ext, cap);
    CHECK(p.npatch < 8 || p.npatch % 8 == 0, "%s: patch count %d not padded", what, p.npatch);
    // a corrupted plan is refused
    if (n >= 2) {
        TilePlanHost q = p;
        std::swap(q.perm[0], q.perm[1]);
        CHECK(validate_tile_plan(n, k, c, cap, q, &err) != 0, "%s: corrupted perm passed", what);
        // every local column of a patch names a distinct row, so moving one
        // always names another row (or leaves the patch + halo range)
        TilePlanHost w = p;
        w.lcol[0] = (short)(w.lcol[0] + 1);
        CHECK(validate_tile_plan(n, k, c, cap, w, &err) != 0, "%s: corrupted local column passed",
              what);
    }
}

static void check_chunks(std::mt19937 &g) {
    std::uniform_int_distribution<int> dG(1, 70), dn(1, 5000), dt(1, 4096), coin(0, 2);
    for (int trial = 0; trial < 300; ++trial) {
        const int G = dG(g);
        std::vector<long long> off{0};
        std::vector<int> cr(G, 0);
        for (int i = 0; i < G; ++i) {
            off.push_back(off.back() + dn(g));
            if (coin(g) == 0) cr[i] = 1 + (int)(g() % 700);
        }
        const long n = (long)off.back();
        const long target = dt(g);
        std::vector<long> rows;
        std::vector<int> gch;
        std::string err;
        const bool strided = trial % 2 == 0;
        const int rc = plan_chunks(n, G, off.data(), strided ? cr.data() : nullptr, target, 8,
                                   &rows, &gch, &err);
        CHECK(rc == 0, "plan_chunks failed: %s", err.c_str());
        if (rc) continue;
        const long nch = (long)rows.size() - 1;
        CHECK(rows.front() == 0 && rows.back() == n, "chunk ends");
        CHECK(nch < 8 || nch % 8 == 0, "chunk count %ld not padded", nch);
        for (long c = 0; c < nch; ++c) CHECK(rows[c] <= rows[c + 1], "chunk order at %ld", c);
        CHECK((int)gch.size() == G + 1 && gch[0] == 0, "group table size");
        for (int i = 0; i < G; ++i) {
            const int c0 = gch[i], c1 = gch[i + 1];
            CHECK(c1 > c0, "realization %d has no chunk", i);
            CHECK(rows[c0] == off[i] && rows[c1] == off[i + 1], "realization %d not covered", i);
            for (int c = c0; c < c1; ++c) {
                const long len = rows[c + 1] - rows[c];
                CHECK(len > 0, "empty chunk inside realization %d", i);
                if (strided && cr[i] > 0 && c + 1 < c1)
                    CHECK(len == cr[i], "stride chunk of %ld rows, want %d", len, cr[i]);
                if (!(strided && cr[i] > 0) && c + 1 < c1)
                    CHECK(len % 64 == 0, "proportional chunk of %ld rows", len);
            }
        }
    }
    // bad offsets are refused
    std::vector<long> rows;
    std::vector<int> gch;
    std::string err;
    const long long bad1[3] = {0, 10, 10};
    CHECK(plan_chunks(10, 2, bad1, nullptr, 64, 8, &rows, &gch, &err) != 0, "repeated offset");
    const long long bad2[3] = {1, 5, 10};
    CHECK(plan_chunks(10, 2, bad2, nullptr, 64, 8, &rows, &gch, &err) != 0, "offset start");
    const long long bad3[2] = {0, 9};
    CHECK(plan_chunks(10, 1, bad3, nullptr, 64, 8, &rows, &gch, &err) != 0, "offset end");
    CHECK(plan_chunks(10, 0, bad3, nullptr, 64, 8, &rows, &gch, &err) != 0, "G = 0");
}

// Wavefront plans of a matrix for a few chunkings; every built plan must pass
// the independent check, and plans corrupted in the ways a wrong build would
// be must fail it.
static void check_wave(int n, int k, const std::vector<int> &c, const char *what) {
    for (int nchunks : {8, 16, 64})
        for (int rpw : {8, 16}) {
            const int blk = 4 * rpw;
            long rpc = ((long)n + nchunks - 1) / nchunks;
            rpc = (rpc + blk - 1) / blk * blk;
            WavePlanHost p;
            std::string err;
            const int rc = build_wave_plan(n, k, c, nchunks, (int)rpc, blk, rpw, 8, 1, 64, &p, &err);
            if (rc) continue;  // refused (a chunk without rows, too many chunks per block, ...)
            CHECK(validate_wave_plan(n, k, c, 8, 1, 64, p, &err) == 0, "%s: %s", what, err.c_str());
            size_t main_rows = 0;
            for (const auto &d : p.deps)
                for (int w = 0; w < 4; ++w) main_rows += (size_t)rpw - __builtin_popcount(d.mask[w]);
            (void)main_rows;
            // corruptions the check must catch
            for (size_t q = 0; q < p.deps.size(); ++q) {
                if (p.deps[q].cnt == 0) continue;
                WavePlanHost bad = p;
                bad.deps[q].need[0] = 0;  // a requirement below a block its rows read
                CHECK(validate_wave_plan(n, k, c, 8, 1, 64, bad, &err) != 0, "%s: zero need passed", what);
                bad = p;
                bad.deps[q].mask[0] ^= 1u;  // a row moved out of (or into) the main loop
                CHECK(validate_wave_plan(n, k, c, 8, 1, 64, bad, &err) != 0, "%s: mask flip passed", what);
                break;
            }
            if (!p.fix_rows.empty()) {
                WavePlanHost bad = p;
                bad.fix_rows.pop_back();
                bad.fix_group_off.back() -= 1;
                CHECK(validate_wave_plan(n, k, c, 8, 1, 64, bad, &err) != 0, "%s: lost fix-up row passed", what);
            }
            for (int t = 0; t < (int)p.tail_dep.size(); ++t)
                if (p.tail_dep[t] >= 0) {
                    WavePlanHost bad = p;
                    bad.tail_dep[t] = -1;
                    CHECK(validate_wave_plan(n, k, c, 8, 1, 64, bad, &err) != 0, "%s: dropped tail chunk passed", what);
                    break;
                }
        }
}

int main() {
    std::mt19937 g(2024);
    for (int n : {1, 2, 3, 7, 64, 500, 2000})
        for (int k : {1, 2, 4, 5})
            for (int cap : {2 * k + 1, 2 * k + 2, 16, 64, 200, 512}) {
                auto c = random_cols(n, k, g);
                check_tiles(n, k, c, cap, "random");
            }
    for (int L : {2, 3, 8, 33})
        for (bool per : {true, false})
            for (int cap : {12, 40, 128, 512}) {
                auto c = square_cols(L, per);
                check_tiles(L * L, 5, c, cap, per ? "square periodic" : "square open");
            }
    // bad shapes are refused
    {
        TilePlanHost p;
        std::string err;
        std::vector<int> c(10, 0);
        CHECK(build_tile_plan(4, 3, c, 64, 8, &p, &err) != 0, "size mismatch accepted");
    }
    check_chunks(g);
    for (int L : {16, 32, 64, 100})
        check_wave(L * L, 5, square_cols(L, true), "wave square periodic");
    for (int n : {64, 4096}) {
        auto c = random_cols(n, 4, g);
        check_wave(n, 4, c, "wave random");
    }
    {
        // chunk = one lattice row of a 256 x 256 periodic square (8 blocks of
        // 32 rows): lag 1; the x = 0 site of each row reads x = 255, 7 blocks
        // ahead (a tail row); the 16 chunks at the 8 XCD boundaries read
        // another XCD's chunk from every row (fix-up rows, their x = 0 sites
        // included)
        const int L = 256;
        auto c = square_cols(L, true);
        WavePlanHost p;
        std::string err;
        CHECK(build_wave_plan(L * L, 5, c, 256, 256, 32, 8, 8, 1, 64, &p, &err) == 0,
              "wave rows: %s", err.c_str());
        CHECK(p.lag == 1, "wave rows: lag %d", p.lag);
        CHECK(p.tail_rows.size() == 240, "wave rows: %zu tail rows", p.tail_rows.size());
        CHECK(p.fix_rows.size() == 16 * 256, "wave rows: %zu fix-up rows", p.fix_rows.size());
        CHECK(validate_wave_plan(L * L, 5, c, 8, 1, 64, p, &err) == 0, "wave rows: %s", err.c_str());
    }
    if (failures) {
        std::fprintf(stderr, "%d failed checks\n", failures);
        return 1;
    }
    std::printf("plan checks passed\n");
    return 0;
}
