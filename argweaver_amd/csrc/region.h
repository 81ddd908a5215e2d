// region.h -- cutting an ARG at a position and joining the pieces again
// (the host surgery of region re-threading).
#pragma once

#include <vector>

#include "local_trees.h"

namespace aw {

// tr keeps [start, pos) (its last tree cut at pos: a block of length 0 when
// pos is a block start); the returned ARG holds [pos, end).
ArgTrees partition_trees(ArgTrees &tr, int pos);
// t2 appended to tr (t2 left empty); the two trees that meet are joined
// by a congruent mapping when t2's first SPR is null (the redundant block
// removed), else by repairing that SPR
void append_trees(ArgTrees &tr, ArgTrees &t2);
void map_congruent_trees(const LTree &t1, const std::vector<int> &seqids1, const LTree &t2,
                         const std::vector<int> &seqids2, std::vector<int> &mapping);
void infer_mapping(const LTree &t1, const LTree &t2, int recomb_node, std::vector<int> &mapping);
void repair_spr(const LTree &last, const LTree &t, Spr &spr, std::vector<int> &mapping);
// the state (node of partial, time) at which the subtree of partial (a tree
// in subtree/maintree form) hangs in full; node -1 when partial is complete
void find_state_sub_tree_internal(const LTree &full, const LTree &partial, int maxtime,
                                  int &node, int &time);

}  // namespace aw
