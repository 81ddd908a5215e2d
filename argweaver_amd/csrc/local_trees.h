// local_trees.h -- host-side ARG representation (a list of local trees with
// the SPR between neighbours) and the leaf-thread surgery the threading
// sampler needs.  Node naming, child-slot order and block splitting follow
// the reference exactly, because callers read the resulting trees back with
// get_local_trees_ptrees and compare node numbering.
//   data model      src/argweaver/local_tree.h:138-674
//   apply_spr       src/argweaver/local_tree.cpp:223-314
//   add/remove leaf src/argweaver/thread.cpp:9-554
#pragma once

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "aw_math.h"

namespace aw {

struct LNode {
    int parent, child[2], age;
};

struct LTree {
    int nnodes = 0;
    int root = -1;
    std::vector<LNode> nodes;   // size() is the capacity

    void ensure_capacity(int cap)
    {
        if ((int)nodes.size() < cap) nodes.resize(cap);
    }
    int num_leaves() const { return (nnodes + 1) / 2; }
    int sibling(int node) const
    {
        int p = nodes[node].parent;
        if (p == -1) return -1;
        const int *c = nodes[p].child;
        return c[0] == node ? c[1] : c[0];
    }
};

struct LBlock {
    LTree tree;
    Spr spr;                    // SPR from the previous tree (-1s when null)
    std::vector<int> mapping;   // previous-tree node -> this-tree node; empty for block 0
    int blocklen = 0;
};

struct ArgTrees {
    int start_coord = 0, end_coord = 0, nnodes = 0;
    std::string chrom = "chr";  // chromosome name of the region (.smc REGION line)
    std::vector<LBlock> blocks;
    std::vector<int> seqids;    // leaf -> sequence row

    int num_leaves() const { return (nnodes + 1) / 2; }
    int length() const { return end_coord - start_coord; }
};

inline bool spr_is_null(const Spr &s) { return s.rn == -1; }

// f(lo, hi) over consecutive ranges covering [0, n), on up to 16 host
// threads (on the calling thread alone when n < 256)
template <class F>
void parallel_ranges(int n, F f)
{
    const int nthr = n < 256 ? 1 : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (nthr == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nthr; t++)
        pool.emplace_back(f, (int)((long long)n * t / nthr), (int)((long long)n * (t + 1) / nthr));
    for (auto &th : pool) th.join();
}

// Construct from the caller's arrays (LocalTrees ctor, local_tree.cpp:321-350).
ArgTrees *trees_from_arrays(int **ptrees, int **ages, int **sprs, const int *blocklens,
                            int ntrees, int nnodes, int start_coord);
// Trunk genealogy of one leaf (LocalTrees::make_trunk).
void make_trunk(ArgTrees &tr, int start, int end, int capacity);
void tree_from_parents(LTree &t, const int *ptree, const int *ages, int nnodes, int capacity);

void apply_spr(LTree &t, const Spr &spr);
void add_tree_branch(LTree &t, int node, int time);
void remove_tree_branch(LTree &t, int remove_leaf, int *displace);

// add_spr_branch (thread.cpp:176-266): rewrite the SPR into tree t (both
// trees carrying the new branch) and extend the node mapping for it
void add_spr_branch(const LTree &t, const LTree &last, int st_node, int st_time, int last_node,
                    Spr &spr, int *mapping, int newleaf, int displaced, int newcoal);

// Splice a sampled thread into the ARG (thread.cpp:272-418).  path holds the
// state index of every column (relative to start_coord); recomb_pos are
// absolute coordinates.
void add_arg_thread(ArgTrees &tr, int ntimes, const int *path, int seqid,
                    const std::vector<int> &recomb_pos, const std::vector<int> &recomb_node,
                    const std::vector<int> &recomb_time);
// Remove the leaf holding sequence `seqid` (thread.cpp:425-554).
void remove_arg_thread(ArgTrees &tr, int seqid);

// Merge every block whose SPR is null into the block before it
// (remove_null_sprs, local_tree.cpp:409-461).
void remove_null_sprs(ArgTrees &tr);

// Structural checks (assert_trees, local_tree.cpp:1763-1809); thread
// insertion and removal run them when check_edits() (AWAMD_CHECK_TREES=1).
bool check_trees(const ArgTrees &tr);
bool check_edits();

// Flat, device-friendly copy of the blocks: per block nnodes ints of
// parent/child0/child1/age, 4 spr ints, nnodes mapping ints, start, length.
struct FlatTrees {
    int nblocks = 0, nnodes = 0, start_coord = 0, seqlen = 0;
    std::vector<int> parent, child0, child1, age, spr, mapping, start, len;
};
void flatten(const ArgTrees &tr, FlatTrees &f);

// state list of a tree (node-major, time ascending)
void tree_states(const LTree &t, int ntimes, std::vector<int> &node, std::vector<int> &time);

}  // namespace aw
