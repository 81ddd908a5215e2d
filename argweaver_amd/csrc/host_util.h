// host_util.h -- host-side helpers shared by the ARG statistics and the
// dense-matrix exports: a local tree as flat arrays (TreeView) and the log
// probability of one SPR.
#pragma once

#include <vector>

#include "aw_math.h"
#include "local_trees.h"

namespace aw {

// flat arrays of one local tree, viewed as a TreeView
struct FlatTree {
    std::vector<int> par, c0, c1, age;
    TreeView view;
    explicit FlatTree(const LTree &t)
    {
        const int N = t.nnodes;
        par.resize(N);
        c0.resize(N);
        c1.resize(N);
        age.resize(N);
        for (int i = 0; i < N; i++) {
            par[i] = t.nodes[i].parent;
            c0[i] = t.nodes[i].child[0];
            c1[i] = t.nodes[i].child[1];
            age[i] = t.nodes[i].age;
        }
        view = TreeView{N, t.root, par.data(), c0.data(), c1.data(), age.data()};
    }
};

// calc_spr_prob (total_prob.cpp:213-257): log probability of the SPR that
// leads out of tree t (recombination point, then re-coalescence), given the
// tree's length
double spr_logprob(const TreeView &t, const Model &m, const Spr &spr, double treelen);

}  // namespace aw
