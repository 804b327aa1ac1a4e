// Host-only planning code (see plan.h).  Reference counterpart: none -- the
// reference launches one fixed grid per step over one matrix
// (jclsquant/cuda_cython/cuda/recursions.cu:19-20).
#include "plan.h"

#include <algorithm>
#include <queue>

namespace jkpm {

// Breadth-first patches with at most `cap` rows of patch + halo each.
int build_tile_plan(int n, int k, const std::vector<int> &colh, int cap, int xcd,
                    TilePlanHost *out, std::string *err) {
    auto fail = [&](const char *what) {
        *err = what;
        return 1;
    };
    if (n < 0 || k < 1 || (long)colh.size() != (long)n * k) return fail("tile plan: bad shape");
    if (cap < 2 * k + 2) return fail("tile capacity too small for the ELL width");
    std::vector<int> owner(n, -1), hmark(n, -1), qmark(n, -1);
    std::vector<int> perm;
    perm.reserve(n);
    std::vector<int> pstart{0};
    std::vector<int> queue;
    queue.reserve(4 * cap);
    int p = 0;
    for (int seed = 0; seed < n; ++seed) {
        if (owner[seed] >= 0) continue;
        int owned = 0, halo = 0;
        queue.clear();
        size_t qh = 0;
        queue.push_back(seed);
        qmark[seed] = p;
        while (qh < queue.size()) {
            const int r = queue[qh++];
            if (owner[r] >= 0) continue;
            // rows r would add to the halo (an upper bound: repeated slots count twice)
            int add = 0;
            for (int s = 0; s < k; ++s) {
                const int j = colh[(size_t)s * n + r];
                if (j != r && owner[j] != p && hmark[j] != p) ++add;
            }
            const int ext_after = owned + 1 + halo - (hmark[r] == p ? 1 : 0) + add;
            if (owned > 0 && ext_after > cap) break;
            owner[r] = p;
            perm.push_back(r);
            ++owned;
            if (hmark[r] == p) {
                --halo;
                hmark[r] = -1;
            }
            for (int s = 0; s < k; ++s) {
                const int j = colh[(size_t)s * n + r];
                if (j == r) continue;
                if (owner[j] != p && hmark[j] != p) {
                    hmark[j] = p;
                    ++halo;
                }
                if (owner[j] < 0 && qmark[j] != p) {
                    qmark[j] = p;
                    queue.push_back(j);
                }
            }
        }
        pstart.push_back((int)perm.size());
        ++p;
    }
    const int P = p;
    // pad the patch count to a multiple of the XCD count with empty patches
    int Ppad = P;
    if (Ppad >= xcd && Ppad % xcd) Ppad += xcd - Ppad % xcd;
    while ((int)pstart.size() < Ppad + 1) pstart.push_back(n);

    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) inv[perm[i]] = i;
    // exact halos (new numbering, sorted) and patch-local columns
    std::vector<int> hoff{0}, hrows;
    std::vector<short> lcol((size_t)n * k);
    std::vector<int> mark(n, -1), hl;
    int max_ext = 0;
    for (int q = 0; q < Ppad; ++q) {
        const int a = pstart[q], bnd = pstart[q + 1];
        hl.clear();
        for (int i = a; i < bnd; ++i) {
            const int r = perm[i];
            for (int s = 0; s < k; ++s) {
                const int jn = inv[colh[(size_t)s * n + r]];
                if ((jn < a || jn >= bnd) && mark[jn] != q) {
                    mark[jn] = q;
                    hl.push_back(jn);
                }
            }
        }
        std::sort(hl.begin(), hl.end());
        const int nown = bnd - a, ext = nown + (int)hl.size();
        if (ext > cap) return fail("internal: patch exceeds the tile capacity");
        max_ext = std::max(max_ext, ext);
        for (int i = a; i < bnd; ++i) {
            const int r = perm[i];
            for (int s = 0; s < k; ++s) {
                const int jn = inv[colh[(size_t)s * n + r]];
                int loc;
                if (jn >= a && jn < bnd) {
                    loc = jn - a;
                } else {
                    loc = nown + (int)(std::lower_bound(hl.begin(), hl.end(), jn) - hl.begin());
                }
                lcol[(size_t)s * n + i] = (short)loc;
            }
        }
        hrows.insert(hrows.end(), hl.begin(), hl.end());
        hoff.push_back((int)hrows.size());
    }
    out->npatch = Ppad;
    out->max_ext = max_ext;
    out->perm = std::move(perm);
    out->inv = std::move(inv);
    out->pstart = std::move(pstart);
    out->hoff = std::move(hoff);
    out->hrows = std::move(hrows);
    out->lcol = std::move(lcol);
    return 0;
}

// Full consistency check of a plan against the columns it was built from:
// perm is a bijection, patches tile 0..n, halos are sorted, unique and
// outside their patch, and every patch-local column names the right row.
int validate_tile_plan(int n, int k, const std::vector<int> &colh, int cap,
                       const TilePlanHost &pl, std::string *err) {
    auto bad = [&](const char *what) {
        *err = std::string("tile plan: ") + what;
        return 1;
    };
    if ((int)pl.perm.size() != n || (int)pl.inv.size() != n) return bad("size");
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; ++i) {
        const int r = pl.perm[i];
        if (r < 0 || r >= n || seen[r]) return bad("perm is not a permutation");
        seen[r] = 1;
        if (pl.inv[r] != i) return bad("inv");
    }
    const int P = pl.npatch;
    if ((int)pl.pstart.size() != P + 1 || (int)pl.hoff.size() != P + 1) return bad("tables");
    if (pl.pstart[0] != 0 || pl.pstart[P] != n || pl.hoff[0] != 0 ||
        pl.hoff[P] != (int)pl.hrows.size())
        return bad("table ends");
    if ((int)pl.lcol.size() != n * k) return bad("lcol size");
    for (int q = 0; q < P; ++q) {
        const int a = pl.pstart[q], b = pl.pstart[q + 1];
        const int h0 = pl.hoff[q], h1 = pl.hoff[q + 1];
        if (b < a || h1 < h0) return bad("table order");
        const int nown = b - a, ext = nown + (h1 - h0);
        if (ext > cap || ext > pl.max_ext) return bad("patch exceeds the tile");
        for (int h = h0; h < h1; ++h) {
            const int g = pl.hrows[h];
            if (g < 0 || g >= n || (g >= a && g < b)) return bad("halo row");
            if (h > h0 && pl.hrows[h - 1] >= g) return bad("halo order");
        }
        for (int i = a; i < b; ++i)
            for (int s = 0; s < k; ++s) {
                const int loc = pl.lcol[(size_t)s * n + i];
                if (loc < 0 || loc >= ext) return bad("local column out of range");
                const int g = loc < nown ? a + loc : pl.hrows[h0 + loc - nown];
                if (g != pl.inv[colh[(size_t)s * n + pl.perm[i]]]) return bad("local column");
            }
    }
    return 0;
}

int plan_chunks(long n_total, int G, const long long *row_offsets, const int *chunk_rows,
                long target, int xcd, std::vector<long> *rows_out, std::vector<int> *gch_out,
                std::string *err) {
    if (G < 1 || !row_offsets) {
        *err = "batch: G must be positive and row_offsets given";
        return 1;
    }
    if (row_offsets[0] != 0 || row_offsets[G] != n_total) {
        *err = "batch: row_offsets must run from 0 to the handle's row count";
        return 1;
    }
    for (int g = 0; g < G; ++g)
        if (row_offsets[g + 1] <= row_offsets[g]) {
            *err = "batch: row_offsets must be strictly increasing";
            return 1;
        }
    target = std::max<long>(target, G);
    std::vector<long> cnt(G, 1);
    auto rows_of = [&](int g) { return (long)(row_offsets[g + 1] - row_offsets[g]); };
    auto size_of = [&](int g) { return (double)rows_of(g) / (double)cnt[g]; };
    auto cmp = [&](int a, int b) { return size_of(a) < size_of(b); };
    std::priority_queue<int, std::vector<int>, decltype(cmp)> q(cmp);
    for (int g = 0; g < G; ++g) q.push(g);
    long total = G;
    while (total < target && !q.empty()) {
        const int g = q.top();
        q.pop();
        if (rows_of(g) / 64 <= cnt[g]) continue;  // chunks of at least 64 rows: g is done
        ++cnt[g];
        ++total;
        q.push(g);
    }
    std::vector<long> rows{0};
    std::vector<int> gch{0};
    for (int g = 0; g < G; ++g) {
        const long n0 = row_offsets[g], ng = rows_of(g);
        long per = (ng + cnt[g] - 1) / cnt[g];
        per = (per + 63) / 64 * 64;
        if (chunk_rows && chunk_rows[g] > 0) per = chunk_rows[g];
        for (long r0 = 0; r0 < ng; r0 += per) rows.push_back(n0 + std::min(ng, r0 + per));
        gch.push_back((int)rows.size() - 1);
    }
    // Pad with empty chunks to a multiple of the XCD count; they belong to no group.
    while ((long)(rows.size() - 1) % xcd != 0 && (long)rows.size() - 1 >= xcd)
        rows.push_back(n_total);
    *rows_out = std::move(rows);
    *gch_out = std::move(gch);
    return 0;
}

}  // namespace jkpm
