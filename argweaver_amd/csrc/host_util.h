// host_util.h -- host-side helpers shared by the ARG statistics and the
// dense-matrix exports: a local tree as flat arrays (TreeView) and the log
// probability of one SPR.
#pragma once

#include <cstdlib>
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

// irand (common.h:109-113): uniform integer in [lo, hi) from the
// process-global glibc stream, drawn through a float as the reference does
inline int irand(int lo, int hi)
{
    const int i = lo + int(rand() / float(RAND_MAX) * (hi - lo));
    return (i == hi) ? hi - 1 : i;
}
inline int irand(int max) { return irand(0, max); }

// relative-or-absolute closeness used by the reference's checks (fequal,
// common.h)
bool close_enough(double a, double b, double rel, double eabs);

// likelihood_sites for state statei of a complete tree t2 (emit.cpp:328-397):
// the invariant-site likelihood for columns marked invariant (1 for a masked
// column), a full Felsenstein pass over the tree for the others; rows[i] is
// leaf i's sequence
void emission_column_slow(const LTree &t2, const Model &m, const char *const *rows, int seqlen,
                          const std::vector<char> &invariant, int statei, double **emit);

}  // namespace aw
