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

// ---- two steps per launch by a wavefront over row chunks (wave.hip) ----

// Readiness requirements of one (chunk, block): before the second step of the
// block's rows, the chunks chunk[0 .. cnt) must have finished the first step
// of at least need[k] blocks; mask[w] has bit i set when row i of wave w's
// share of the block is not done in the main loop (a tail or fix-up row).
// 64 bytes, the layout the kernel reads.
struct WaveDep {
    int cnt;
    int pad[3];
    int chunk[4];
    unsigned need[4];
    unsigned mask[4];
};

struct WavePlanHost {
    int nchunks = 0, rpc = 0, nblk = 0, blk = 0, rpw = 0, lag = 0;
    std::vector<WaveDep> deps;         // [nchunks][nblk]
    // Tail rows: their columns stay on the chunk's XCD but reach a block more
    // than max_lag ahead (the periodic wrap of a lattice row); the chunk's
    // work-group does them after its main loop, once the chunks they read
    // (tail_dep, -1 = unused) have finished every first-step block.
    std::vector<int> tail_off;         // [nchunks + 1] into tail_rows
    std::vector<int> tail_rows;
    std::vector<int> tail_dep;         // [nchunks][4]
    // Fix-up rows: their columns reach another XCD's chunk.  Groups of at
    // most fix_group rows of one chunk, done after the launch; fold lists,
    // per chunk, the run of groups whose partials are added to its own.
    std::vector<int> fix_group_chunk;  // [groups]
    std::vector<int> fix_group_off;    // [groups + 1] into fix_rows
    std::vector<int> fix_rows;
    std::vector<int> fold_chunk;       // chunks that own groups
    std::vector<int> fold_off;         // [fold_chunk.size() + 1] into the groups
};

// Plan of the wavefront path for the slot-major [k][n] columns colh cut into
// `nchunks` chunks of `rpc` rows (the last one shorter), blocks of `blk` rows
// (4 waves of `rpw` rows), chunks assigned XCD-major to `xcd` XCDs (chunk c
// on XCD c / (nchunks / xcd)).  A row whose columns reach a chunk of another
// XCD is a fix-up row; one whose columns reach a block more than max_lag
// blocks ahead of its own is a tail row; every other row adds its columns'
// (chunk, block) to the requirements of its own (chunk, block).  The lag is
// the largest (referenced block - own block) over those rows (at least 1).
// Plans whose blocks or tails reference more than four chunks, or that leave
// more than a quarter of the rows out of the main loop, are refused.
// Returns 0, or 1 with *err set.
int build_wave_plan(int n, int k, const std::vector<int> &colh, int nchunks, int rpc, int blk,
                    int rpw, int xcd, int max_lag, int fix_group, WavePlanHost *out,
                    std::string *err);

// Independent check of a plan against its columns: every requirement is met
// by the rows each main-loop row reads, the rows are sorted into main, tail
// and fix-up exactly by the rule above, the lag bounds every requirement,
// every tail row's chunks are in its chunk's tail list, and the fix-up groups
// and folds cover exactly the fix-up rows.  Returns 0, or 1 with *err set.
int validate_wave_plan(int n, int k, const std::vector<int> &colh, int xcd, int max_lag,
                       int fix_group, const WavePlanHost &plan, std::string *err);

}  // namespace jkpm
