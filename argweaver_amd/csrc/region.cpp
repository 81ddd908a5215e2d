// region.cpp -- cutting an ARG at a position and joining the pieces again,
// the host surgery of region re-threading (resample_arg_region): the
// region's trees are re-threaded on their own, conditioned on the states
// the removed branch held at the region's ends, and spliced back.
//   partition_local_trees  src/argweaver/local_tree.cpp:502-584
//   map_congruent_trees    src/argweaver/local_tree.cpp:588-641
//   infer_mapping          src/argweaver/local_tree.cpp:646-703
//   repair_spr             src/argweaver/local_tree.cpp:708-735
//   append_local_trees     src/argweaver/local_tree.cpp:745-785
//   remove_null_spr        src/argweaver/local_tree.cpp:409-452
//   find_state_sub_tree_internal  src/argweaver/sample_arg.cpp:372-419
#include "region.h"

#include <utility>

namespace aw {
namespace {

// children before parents
void postorder_nodes(const LTree &t, std::vector<int> &order)
{
    order.clear();
    std::vector<std::pair<int, int>> stack{{t.root, 0}};
    while (!stack.empty()) {
        auto &top = stack.back();
        const LNode &n = t.nodes[top.first];
        if (n.child[0] != -1 && top.second < 2) {
            const int c = n.child[top.second++];
            stack.push_back({c, 0});
            continue;
        }
        order.push_back(top.first);
        stack.pop_back();
    }
}

bool is_leaf(const LTree &t, int n) { return t.nodes[n].child[0] == -1; }

}  // namespace

ArgTrees partition_trees(ArgTrees &tr, int pos)
{
    ArgTrees t2;
    t2.nnodes = tr.nnodes;
    t2.chrom = tr.chrom;
    t2.seqids = tr.seqids;
    if (pos == tr.start_coord) {
        t2.start_coord = pos;
        t2.end_coord = tr.end_coord;
        t2.blocks = std::move(tr.blocks);
        tr.blocks.clear();
        tr.end_coord = pos;
        return t2;
    }
    if (pos == tr.end_coord) {
        t2.start_coord = t2.end_coord = pos;
        return t2;
    }
    int end = tr.start_coord, b = 0, bstart = 0;
    for (; b < (int)tr.blocks.size(); b++) {
        bstart = end;
        end += tr.blocks[b].blocklen;
        if (bstart <= pos && pos < end) break;
    }
    t2.start_coord = pos;
    t2.end_coord = tr.end_coord;
    for (size_t i = b; i < tr.blocks.size(); i++) t2.blocks.push_back(std::move(tr.blocks[i]));
    tr.blocks.resize(b);
    // the first tree stays on this side too, cut to [bstart, pos) (a block of
    // length 0 when pos is a block start)
    LBlock &f = t2.blocks[0];
    LBlock left;
    left.tree = f.tree;
    left.spr = f.spr;
    left.mapping = f.mapping;
    left.blocklen = pos - bstart;
    tr.blocks.push_back(std::move(left));
    f.mapping.clear();
    f.spr = Spr{-1, -1, -1, -1};
    tr.end_coord = pos;
    f.blocklen -= pos - bstart;
    return t2;
}

void map_congruent_trees(const LTree &t1, const std::vector<int> &seqids1, const LTree &t2,
                         const std::vector<int> &seqids2, std::vector<int> &mapping)
{
    const int n1 = t1.num_leaves(), n2 = t2.num_leaves();
    mapping.assign(t1.nnodes, -1);
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++)
            if (seqids2[j] == seqids1[i]) {
                mapping[i] = j;
                break;
            }
    std::vector<int> order;
    postorder_nodes(t1, order);
    for (int j : order) {
        if (is_leaf(t1, j)) continue;
        const int a = mapping[t1.nodes[j].child[0]], b = mapping[t1.nodes[j].child[1]];
        if (a != -1 && b != -1) mapping[j] = t2.nodes[a].parent;
        else if (a != -1) mapping[j] = a;
        else if (b != -1) mapping[j] = b;
        else mapping[j] = -1;
    }
}

void infer_mapping(const LTree &t1, const LTree &t2, int recomb_node, std::vector<int> &mapping)
{
    const int nleaves = t1.num_leaves();
    mapping.assign(t1.nnodes, -1);
    for (int i = 0; i < nleaves; i++) mapping[i] = i;
    std::vector<int> order;
    postorder_nodes(t1, order);
    for (int j : order) {
        if (is_leaf(t1, j)) continue;
        const int a = mapping[t1.nodes[j].child[0]], b = mapping[t1.nodes[j].child[1]];
        if (a != -1 && b != -1 && t2.nodes[a].parent == t2.nodes[b].parent)
            mapping[j] = t2.nodes[a].parent;
    }
    const int broken = t1.nodes[recomb_node].parent;
    const int other = t1.sibling(recomb_node);
    const int recoal = t2.nodes[mapping[recomb_node]].parent;
    for (int j : order) {
        if (is_leaf(t1, j) || j == broken) continue;
        int a = t1.nodes[j].child[0], b = t1.nodes[j].child[1];
        if (a == broken) a = other;
        if (b == broken) b = other;
        int c = t2.nodes[mapping[a]].parent;
        if (c == recoal) c = t2.nodes[c].parent;
        mapping[j] = c;
    }
    mapping[broken] = -1;
}

void repair_spr(const LTree &last, const LTree &t, Spr &spr, std::vector<int> &mapping)
{
    infer_mapping(last, t, spr.rn, mapping);
    const int broken = last.nodes[spr.rn].parent;
    const int recomb = mapping[spr.rn];
    const int recoal = t.nodes[recomb].parent;
    spr.ct = t.nodes[recoal].age;
    const int other = t.sibling(recomb);
    std::vector<int> inv(t.nnodes, -1);
    for (int i = 0; i < (int)mapping.size(); i++)
        if (mapping[i] != -1) inv[mapping[i]] = i;
    spr.cn = inv[other];
    if (spr.cn == broken) spr.cn = last.sibling(spr.rn);
    const int parent = last.nodes[spr.cn].parent;
    if (parent != -1 && spr.ct > last.nodes[parent].age) spr.cn = parent;
}

// block i merged into block i+1 when i+1's SPR is null
static void remove_null_spr_at(ArgTrees &tr, size_t i)
{
    if (i + 1 >= tr.blocks.size() || !spr_is_null(tr.blocks[i + 1].spr)) return;
    LBlock &B = tr.blocks[i], &N = tr.blocks[i + 1];
    if (B.mapping.empty()) {
        N.mapping.clear();
    } else {
        const int nn = N.tree.nnodes;
        std::vector<int> m(nn);
        for (int j = 0; j < nn; j++) m[j] = (B.mapping[j] != -1) ? N.mapping[B.mapping[j]] : -1;
        N.mapping = m;
        N.spr = B.spr;
    }
    N.blocklen += B.blocklen;
    tr.blocks.erase(tr.blocks.begin() + i);
}

void append_trees(ArgTrees &tr, ArgTrees &t2)
{
    const size_t n1 = tr.blocks.size(), n2 = t2.blocks.size();
    for (LBlock &b : t2.blocks) tr.blocks.push_back(std::move(b));
    t2.blocks.clear();
    tr.end_coord = t2.end_coord;
    t2.end_coord = t2.start_coord;
    if (n1 > 0 && n2 > 0) {
        LBlock &prev = tr.blocks[n1 - 1], &next = tr.blocks[n1];
        if (spr_is_null(next.spr)) {
            // no SPR between the two trees: a congruent mapping, then the
            // redundant block goes
            map_congruent_trees(prev.tree, tr.seqids, next.tree, t2.seqids, next.mapping);
            remove_null_spr_at(tr, n1 - 1);
        } else {
            repair_spr(prev.tree, next.tree, next.spr, next.mapping);
        }
    }
}

void find_state_sub_tree_internal(const LTree &full, const LTree &partial, int maxtime,
                                  int &node, int &time)
{
    node = time = -1;
    if (partial.nodes[partial.root].age < maxtime) return;   // complete tree
    const int sub = partial.nodes[partial.root].child[0];
    // the node by its path length from its left-most leaf
    int count = 0, leaf = sub;
    while (!is_leaf(partial, leaf)) {
        leaf = partial.nodes[leaf].child[0];
        count++;
    }
    int ptr = leaf;
    for (; count > 0; count--) ptr = full.nodes[ptr].parent;
    const int sib = full.sibling(ptr);
    const int parent = full.nodes[ptr].parent;
    time = full.nodes[parent].age;
    count = 0;
    leaf = sib;
    while (!is_leaf(full, leaf)) {
        leaf = full.nodes[leaf].child[0];
        count++;
    }
    ptr = leaf;
    for (; count > 0; count--) ptr = partial.nodes[ptr].parent;
    node = ptr;
}

}  // namespace aw
