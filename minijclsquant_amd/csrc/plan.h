// Host-only planning of the library's launches: the row-chunk tables of a
// realization batch and the patch plan of the two-steps-per-launch path.
// Plain C++17 with no HIP dependency, so it also builds with a host compiler
// alone (the sanitizer build of the CPU tests does that).
#pragma once

#include <string>
#include <vector>

namespace jkpm {

// Patches of the two-steps-per-launch path (see tiles.hip): rows renumbered
// patch by patch, each patch's halo rows, and the patch-local columns.
struct TilePlanHost {
    int npatch = 0, max_ext = 0;
    std::vector<int> perm, inv, pstart, hoff, hrows;  // new -> old, old -> new, tables
    std::vector<short> lcol;                          // [k][n] patch-local columns
};

// Patches of at most `cap` patch + halo rows, grown breadth-first on the
// graph of the slot-major [k][n] column array colh; the patch count is padded
// with empty patches to a multiple of `xcd`.  Returns 0, or 1 with *err set.
int build_tile_plan(int n, int k, const std::vector<int> &colh, int cap, int xcd,
                    TilePlanHost *out, std::string *err);

// Full consistency check of a plan against the columns it was built from.
// Returns 0, or 1 with *err set.
int validate_tile_plan(int n, int k, const std::vector<int> &colh, int cap,
                       const TilePlanHost &plan, std::string *err);

// Row-chunk table of a block-diagonal batch of G realizations with rows
// [row_offsets[g], row_offsets[g+1]) (strictly increasing from 0): about
// `target` chunks in all, none straddling two realizations (every
// realization starts with one; the rest go one at a time to the realization
// whose chunks are currently the largest, chunks of at least 64 rows); a
// positive chunk_rows[g] instead cuts realization g into chunks of exactly
// that many rows (the last one shorter).  The chunk count is padded with
// empty chunks at n_total to a multiple of `xcd` once it reaches xcd.
// rows: chunk boundaries (chunk c = [rows[c], rows[c+1])); gch: chunks of
// realization g are gch[g] .. gch[g+1].  Returns 0, or 1 with *err set.
int plan_chunks(long n_total, int G, const long long *row_offsets, const int *chunk_rows,
                long target, int xcd, std::vector<long> *rows, std::vector<int> *gch,
                std::string *err);

}  // namespace jkpm
