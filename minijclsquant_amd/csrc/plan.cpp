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

// ---- wavefront plan ----

namespace {
// 0: main loop, 1: tail, 2: fix-up
int wave_row_kind(int n, int k, const std::vector<int> &colh, int i, int rpc, int blk,
                  int per_xcd, int max_lag) {
    const int c = i / rpc, q = (i - c * rpc) / blk;
    int kind = 0;
    for (int s = 0; s < k; ++s) {
        const int j = colh[(size_t)s * n + i];
        const int cj = j / rpc;
        if (cj / per_xcd != c / per_xcd) return 2;
        if ((j - cj * rpc) / blk - q > max_lag) kind = 1;
    }
    return kind;
}

// Adds chunk cj to a list of at most four (-1 = free); false when full.
bool add_chunk(int *list, int cj) {
    for (int t = 0; t < 4; ++t) {
        if (list[t] == cj) return true;
        if (list[t] < 0) {
            list[t] = cj;
            return true;
        }
    }
    return false;
}
}  // namespace

int build_wave_plan(int n, int k, const std::vector<int> &colh, int nchunks, int rpc, int blk,
                    int rpw, int xcd, int max_lag, int fix_group, WavePlanHost *out,
                    std::string *err) {
    auto fail = [&](const std::string &m) {
        *err = m;
        return 1;
    };
    if (n < 1 || k < 1) return fail("wavefront plan: empty matrix");
    if ((long)colh.size() != (long)n * k) return fail("wavefront plan: column array size mismatch");
    for (int j : colh)
        if (j < 0 || j >= n) return fail("wavefront plan: column index out of range");
    if (xcd < 1 || nchunks < xcd || nchunks % xcd != 0)
        return fail("wavefront plan: the chunk count must be a positive multiple of the XCD count");
    if (rpw < 1 || rpw > 32 || blk != 4 * rpw)
        return fail("wavefront plan: a block is four waves of at most 32 rows");
    if (rpc < blk || rpc % blk != 0)
        return fail("wavefront plan: rows per chunk must be a positive multiple of the block");
    if ((long)(nchunks - 1) * rpc >= n || (long)nchunks * rpc < n)
        return fail("wavefront plan: every chunk must hold rows");
    if (max_lag < 1) return fail("wavefront plan: the lag must be at least 1");
    if (fix_group < 1) return fail("wavefront plan: the fix-up group must hold rows");
    const int nblk = rpc / blk;
    if (nblk + 4 >= 4096) return fail("wavefront plan: more than 4091 blocks per chunk");
    const int per_xcd = nchunks / xcd;
    WavePlanHost P;
    P.nchunks = nchunks;
    P.rpc = rpc;
    P.nblk = nblk;
    P.blk = blk;
    P.rpw = rpw;
    WaveDep zero{};
    for (int t = 0; t < 4; ++t) zero.chunk[t] = -1;
    P.deps.assign((size_t)nchunks * nblk, zero);
    P.tail_dep.assign((size_t)nchunks * 4, -1);
    std::vector<char> kind(n, 0);
    long lag = 0, nout = 0;
    for (int i = 0; i < n; ++i) {
        const int c = i / rpc;
        const int q = (i - c * rpc) / blk;
        kind[i] = (char)wave_row_kind(n, k, colh, i, rpc, blk, per_xcd, max_lag);
        if (kind[i] == 2) {
            ++nout;
            continue;
        }
        if (kind[i] == 1) {
            ++nout;
            for (int s = 0; s < k; ++s)
                if (!add_chunk(&P.tail_dep[(size_t)c * 4], colh[(size_t)s * n + i] / rpc))
                    return fail("wavefront plan: a chunk's tail rows reference more than four chunks");
            continue;
        }
        WaveDep &d = P.deps[(size_t)c * nblk + q];
        for (int s = 0; s < k; ++s) {
            const int j = colh[(size_t)s * n + i];
            const int cj = j / rpc;
            const int qj = (j - cj * rpc) / blk;
            lag = std::max(lag, (long)qj - q);
            int t = 0;
            while (t < d.cnt && d.chunk[t] != cj) ++t;
            if (t == d.cnt) {
                if (d.cnt == 4)
                    return fail("wavefront plan: a block references more than four chunks");
                d.chunk[t] = cj;
                d.need[t] = 0;
                ++d.cnt;
            }
            d.need[t] = std::max(d.need[t], (unsigned)(qj + 1));
        }
    }
    P.lag = (int)std::max(1L, lag);
    if (4 * nout > n)
        return fail("wavefront plan: " + std::to_string(nout) + " of " + std::to_string(n) +
                    " rows read another XCD's chunks or blocks too far ahead");
    P.tail_off.push_back(0);
    P.fix_group_off.push_back(0);
    P.fold_off.push_back(0);
    for (int c = 0; c < nchunks; ++c) {
        const int r1 = std::min(n, (c + 1) * rpc);
        int in_group = 0;
        const size_t groups_before = P.fix_group_chunk.size();
        for (int i = c * rpc; i < r1; ++i) {
            if (!kind[i]) continue;
            const int off = i - c * rpc;
            WaveDep &d = P.deps[(size_t)c * nblk + off / blk];
            d.mask[(off % blk) / rpw] |= 1u << (off % rpw);
            if (kind[i] == 1) {
                P.tail_rows.push_back(i);
                continue;
            }
            if (in_group == 0) P.fix_group_chunk.push_back(c);
            P.fix_rows.push_back(i);
            if (++in_group == fix_group) {
                P.fix_group_off.push_back((int)P.fix_rows.size());
                in_group = 0;
            }
        }
        if (in_group > 0) P.fix_group_off.push_back((int)P.fix_rows.size());
        P.tail_off.push_back((int)P.tail_rows.size());
        if (P.fix_group_chunk.size() > groups_before) {
            P.fold_chunk.push_back(c);
            P.fold_off.push_back((int)P.fix_group_chunk.size());
        }
    }
    *out = std::move(P);
    return 0;
}

int validate_wave_plan(int n, int k, const std::vector<int> &colh, int xcd, int max_lag,
                       int fix_group, const WavePlanHost &P, std::string *err) {
    auto fail = [&](const std::string &m) {
        *err = "wavefront plan check: " + m;
        return 1;
    };
    if ((long)colh.size() != (long)n * k) return fail("column array size mismatch");
    for (int j : colh)
        if (j < 0 || j >= n) return fail("column index out of range");
    if (P.nchunks < xcd || P.nchunks % xcd != 0 || P.blk != 4 * P.rpw || P.rpw > 32 ||
        P.rpc != P.nblk * P.blk || (long)(P.nchunks - 1) * P.rpc >= n ||
        (long)P.nchunks * P.rpc < n || P.lag < 1 || P.lag > max_lag)
        return fail("inconsistent geometry");
    if (P.deps.size() != (size_t)P.nchunks * P.nblk) return fail("requirement table size");
    if (P.tail_off.size() != (size_t)P.nchunks + 1 || P.tail_dep.size() != (size_t)P.nchunks * 4 ||
        P.tail_off[0] != 0 || P.tail_off.back() != (int)P.tail_rows.size())
        return fail("tail table shape");
    const size_t ng = P.fix_group_chunk.size();
    if (P.fix_group_off.size() != ng + 1 || P.fix_group_off[0] != 0 ||
        P.fix_group_off.back() != (int)P.fix_rows.size() ||
        P.fold_off.size() != P.fold_chunk.size() + 1 || P.fold_off[0] != 0 ||
        P.fold_off.back() != (int)ng)
        return fail("fix-up table shape");
    const int per_xcd = P.nchunks / xcd;
    std::vector<char> seen(n, 0);  // 1: tail list, 2: fix-up list
    for (int c = 0; c < P.nchunks; ++c) {
        if (P.tail_off[c + 1] < P.tail_off[c]) return fail("tail offsets");
        for (int t = P.tail_off[c]; t < P.tail_off[c + 1]; ++t) {
            const int i = P.tail_rows[t];
            if (i < c * P.rpc || i >= std::min(n, (c + 1) * P.rpc)) return fail("tail row outside its chunk");
            if (t > P.tail_off[c] && i <= P.tail_rows[t - 1]) return fail("tail rows not ascending");
            seen[i] = 1;
        }
    }
    for (size_t g = 0; g < ng; ++g) {
        const int c = P.fix_group_chunk[g];
        const int lo = P.fix_group_off[g], hi = P.fix_group_off[g + 1];
        if (c < 0 || c >= P.nchunks || hi <= lo || hi - lo > fix_group) return fail("fix-up group");
        if (g > 0 && c < P.fix_group_chunk[g - 1]) return fail("fix-up groups not by chunk");
        for (int t = lo; t < hi; ++t) {
            const int i = P.fix_rows[t];
            if (i < c * P.rpc || i >= std::min(n, (c + 1) * P.rpc)) return fail("fix-up row outside its chunk");
            if (t > 0 && i <= P.fix_rows[t - 1] && t > lo) return fail("fix-up rows not ascending");
            if (seen[i]) return fail("row listed twice");
            seen[i] = 2;
        }
    }
    for (size_t f = 0; f < P.fold_chunk.size(); ++f) {
        const int c = P.fold_chunk[f];
        if (f > 0 && c <= P.fold_chunk[f - 1]) return fail("folds not ascending");
        if (P.fold_off[f + 1] <= P.fold_off[f]) return fail("empty fold");
        for (int g = P.fold_off[f]; g < P.fold_off[f + 1]; ++g)
            if (P.fix_group_chunk[g] != c) return fail("fold covers another chunk's group");
    }
    for (int i = 0; i < n; ++i) {
        const int c = i / P.rpc, off = i - c * P.rpc, q = off / P.blk;
        const WaveDep &d = P.deps[(size_t)c * P.nblk + q];
        const bool masked = (d.mask[(off % P.blk) / P.rpw] >> (off % P.rpw)) & 1u;
        const int kind = wave_row_kind(n, k, colh, i, P.rpc, P.blk, per_xcd, max_lag);
        if (kind != seen[i]) return fail("row " + std::to_string(i) + " sorted wrongly");
        if (masked != (kind != 0)) return fail("mask wrong at row " + std::to_string(i));
        if (kind == 2) continue;
        for (int s = 0; s < k; ++s) {
            const int j = colh[(size_t)s * n + i];
            const int cj = j / P.rpc, qj = (j - cj * P.rpc) / P.blk;
            if (kind == 1) {
                bool listed = false;
                for (int t = 0; t < 4; ++t) listed |= P.tail_dep[(size_t)c * 4 + t] == cj;
                if (!listed) return fail("tail row " + std::to_string(i) + " reads an unlisted chunk");
                continue;
            }
            int t = 0;
            while (t < d.cnt && t < 4 && d.chunk[t] != cj) ++t;
            if (t >= d.cnt || t >= 4) return fail("row " + std::to_string(i) + " reads an unlisted chunk");
            if (d.need[t] < (unsigned)(qj + 1)) return fail("requirement below a read block");
        }
    }
    for (int c = 0; c < P.nchunks; ++c) {
        for (int t = 0; t < 4; ++t) {
            const int cj = P.tail_dep[(size_t)c * 4 + t];
            if (cj < -1 || cj >= P.nchunks || (cj >= 0 && cj / per_xcd != c / per_xcd))
                return fail("tail requirement");
        }
        for (int q = 0; q < P.nblk; ++q) {
            const WaveDep &d = P.deps[(size_t)c * P.nblk + q];
            if (d.cnt < 0 || d.cnt > 4) return fail("requirement count");
            for (int t = 0; t < d.cnt; ++t) {
                if (d.chunk[t] < 0 || d.chunk[t] >= P.nchunks) return fail("requirement chunk range");
                if (d.chunk[t] / per_xcd != c / per_xcd) return fail("requirement on another XCD");
                if (d.need[t] < 1 || d.need[t] > (unsigned)P.nblk) return fail("requirement range");
                // the bound every wait relies on to end: a wave that waits
                // for block q has finished the first step of q + lag blocks
                if ((long)d.need[t] > (long)q + P.lag + 1) return fail("requirement beyond the lag");
            }
        }
    }
    return 0;
}

}  // namespace jkpm
