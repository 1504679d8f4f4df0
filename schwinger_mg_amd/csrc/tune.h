// tune.h -- the launch-shape autotuner's table.
//
// A choice is kept per (kernel family, shape): which variant of a kernel runs and the
// workgroup count of its reducing launches.  The workgroup count of a reduction fixes
// its summation order, so a choice, once made, is reused verbatim for that shape: runs
// with the same table give bit-identical results.  Shapes without an entry use the
// built-in defaults.  The table is per process; it can be written to and read from a
// text file (one "family a b c d e choice" line per entry).
#pragma once
#include <array>
#include <string>

namespace smg {

enum TuneFamily : int {
    // k_b_coarse_bf16 (FP16 links, matrix cores); shape (nx, ny, k, N, 4*hv + epi);
    // choice = herm + 2 * grid index (herm: 1 reads 5 of 9 link blocks, 0 all 9;
    // grid index: the workgroups of the launch with fused dot products)
    TUNE_COARSE_BF16 = 0,
    // k_b_dslash with fused dot products; shape (L, N, prec, 0, 0); choice = grid index
    TUNE_DSLASH_DOTS = 1,
    // k_b_dslash stores; shape (L, N, prec, epi, 0); choice 1: non-temporal stores, 0: ordinary
    TUNE_DSLASH_STORE = 2,
    // k_h_cg_step (the BF16-vector CG's vector update, no reduction: the grid does not
    // change the results); shape (n, N, 0, 0, 0); choice = grid index
    TUNE_CG_UPDATE = 3,
    // the batched transfers; shape (nx, ny, 128 nd + k, N, 2 (FP64) + (1 prolong, 0 restrict))
    // of the finer lattice; choice 1: a workgroup per aggregate with its blocks staged in
    // LDS, 0: one work-item per output reading through the caches (the same sums), 2
    // (FP32): a wave per 16 x 16 tile on the matrix cores
    TUNE_TRANSFER = 4,
    // D^dag D of the batched path; shape (L, N, prec bytes, epi, 0); choice 0: two stencil
    // launches, 1 + 2 g + nt: one fused launch (kTuneGrids[g] workgroups with dot
    // products; non-temporal stores when nt)
    TUNE_DDAGD = 5,
    TUNE_FAMILIES = 6,
};
using TuneKey = std::array<int, 6>;   // family, then five shape words

// candidate workgroup counts of reducing launches (all fit the partial-sum buffer)
constexpr int kTuneGrids[] = {128, 256, 512, 1024, 2048};
constexpr int kTuneNGrids = 5;

int tune_lookup(const TuneKey& k);   // the stored choice, or -1
void tune_store(const TuneKey& k, int choice);
void tune_clear();
void tune_save(const std::string& path);
void tune_load(const std::string& path);
size_t tune_size();

}  // namespace smg
