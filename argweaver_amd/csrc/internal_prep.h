// internal_prep.h -- host-prepared inputs of the threading HMM for an
// internal branch (the ARG in subtree/maintree form), in the lane-run
// layout of the device kernels.
#pragma once

#include <vector>

#include "aw_math.h"
#include "local_trees.h"

namespace aw {

struct HostPrep {
    int R = 1;                   // states per lane
    std::vector<int> S;          // [B] states as stored (a tree without states keeps one)
    std::vector<int> Sreal;      // [B] states of the reference (0 for a complete tree)
    std::vector<double> planes;  // [B * NCOEF * 64R] coefficient planes (see Coef)
    std::vector<int> info, det;  // [B * 64R] state info, switch targets of block b-1's states
    std::vector<double> tm, tm2s;   // [B * (T-1)^2], [B * TC]
    std::vector<int> srcs;       // [2B] recombination / recoalescence source of each switch
    std::vector<double> prior;   // [64R] first-column prior
    std::vector<int> cls;        // [L] -1 invariant, -2 masked (or no states), >= 0 variant row
    std::vector<double> vemit;   // [nvar * 64R] emissions of the variant columns
    int nvar = 0;
};

// Prepare every block of tr (every tree in subtree/maintree form or
// complete); seqs holds the sequences, leaf i reads row seqids[i].  minage:
// the iterator's minimum state age (0 for sample_arg_thread_internal).
// Returns 0 on success.
int prepare_internal(const ArgTrees &tr, const Model &m, char **seqs, int minage, HostPrep &hp);

// recombinations of an internal thread (state index per column)
void sample_recombinations_internal(const ArgTrees &tr, const Model &m, int minage,
                                    const int *path, std::vector<int> &rpos,
                                    std::vector<int> &rnode, std::vector<int> &rtime);

}  // namespace aw
