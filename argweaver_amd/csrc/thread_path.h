// thread_path.h -- internal-branch re-threading on the host (removal paths,
// removing a branch along a path, adding a sampled thread back).
#pragma once

#include <vector>

#include "local_trees.h"

namespace aw {

// the node of t that no node of the previous tree maps to (get_recoal_node)
int recoal_node(const LTree &t, const int *mapping);
void next_removal_nodes(const LTree &t1, const Spr &spr2, const int *mapping2, int node,
                        int out[2], int recoal);
void prev_removal_nodes(const LTree &t1, const Spr &spr2, const int *mapping2,
                        const int *inv_mapping, int node, int out[2]);

// removal paths: one node per block (thread.cpp:672-1020)
void removal_leaf_path(const ArgTrees &tr, int node, int *path);
void removal_path_at(const ArgTrees &tr, int node, int pos, int *path, double prob_switch);
void removal_path_from_start(const ArgTrees &tr, int node, int *path);
void removal_path_recomb(const ArgTrees &tr, double recomb_preference, int *path);
// uniform over all paths; returns the log number of paths
double removal_path_uniform(const ArgTrees &tr, int *path);
double count_removal_paths(const ArgTrees &tr);

// the states of a tree in subtree/maintree form (node-major, time ascending)
void internal_states(const LTree &t, int ntimes, int minage, std::vector<int> &node,
                     std::vector<int> &time);

// remove the branch along path (maxtime: age of the virtual roots); when
// original_thread is given it receives the removed branch as internal state
// indices per column
void remove_thread_path(ArgTrees &tr, const int *path, int maxtime, int *original_thread);
// add_spr_branch for an internal branch (thread.cpp:1204-1309): rewrite
// the SPR and extend the mapping into tree t once both trees carry the
// subtree (st_node: its node in t, -1 when t is complete; last_node: its
// node in last)
void add_spr_branch_internal(const LTree &t, const LTree &last, int st_node, int st_time,
                             int last_node, Spr &spr, std::vector<int> &mapping, int sub,
                             int last_sub);
// add the sampled thread (internal state index per column) and its
// recombinations (node -1 is not used: a recombination on the thread names
// the subtree root); 0 on success
int add_thread_path(ArgTrees &tr, int ntimes, int minage, const int *thread_path,
                    const std::vector<int> &recomb_pos, const std::vector<int> &recomb_node,
                    const std::vector<int> &recomb_time);

}  // namespace aw
