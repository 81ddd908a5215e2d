// local_trees.cpp -- ARG data model and leaf-thread surgery (host).
#include "local_trees.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace aw {

void tree_from_parents(LTree &t, const int *ptree, const int *ages, int nnodes, int capacity)
{
    t.nnodes = nnodes;
    t.nodes.assign(std::max(capacity, nnodes), LNode{-1, {-1, -1}, 0});
    t.root = -1;
    for (int i = 0; i < nnodes; i++) {
        t.nodes[i].parent = ptree[i];
        t.nodes[i].child[0] = t.nodes[i].child[1] = -1;
        t.nodes[i].age = ages ? ages[i] : 0;
    }
    for (int i = 0; i < nnodes; i++) {
        int p = ptree[i];
        if (p != -1) {
            int *c = t.nodes[p].child;
            if (c[0] == -1) c[0] = i;
            else c[1] = i;
        } else {
            t.root = i;
        }
    }
}

ArgTrees *trees_from_arrays(int **ptrees, int **ages, int **sprs, const int *blocklens,
                            int ntrees, int nnodes, int start_coord)
{
    ArgTrees *tr = new ArgTrees();
    tr->start_coord = start_coord;
    tr->end_coord = start_coord;
    tr->nnodes = nnodes;
    tr->blocks.resize(ntrees);
    for (int i = 0; i < ntrees; i++) {
        LBlock &B = tr->blocks[i];
        tree_from_parents(B.tree, ptrees[i], ages[i], nnodes, nnodes);
        B.spr = Spr{sprs[i][0], sprs[i][1], sprs[i][2], sprs[i][3]};
        B.blocklen = blocklens[i];
        if (i > 0) {
            // make_node_mapping (local_tree.h:767-776)
            B.mapping.resize(nnodes);
            for (int j = 0; j < nnodes; j++) B.mapping[j] = j;
            B.mapping[ptrees[i - 1][sprs[i][0]]] = -1;
        }
        tr->end_coord += blocklens[i];
    }
    const int nleaves = tr->num_leaves();
    tr->seqids.resize(nleaves);
    for (int i = 0; i < nleaves; i++) tr->seqids[i] = i;
    return tr;
}

void make_trunk(ArgTrees &tr, int start, int end, int capacity)
{
    tr.blocks.clear();
    tr.start_coord = start;
    tr.end_coord = end;
    tr.nnodes = 1;
    LBlock B;
    int p = -1, a = 0;
    tree_from_parents(B.tree, &p, &a, 1, std::max(capacity, 1));
    B.spr = Spr{-1, -1, -1, -1};
    B.blocklen = end - start;
    tr.blocks.push_back(std::move(B));
    tr.seqids.assign(1, 0);
}

// apply_spr (local_tree.cpp:223-314): the broken node is reused as the
// re-coalescence node.
void apply_spr(LTree &t, const Spr &spr)
{
    LNode *nodes = t.nodes.data();
    if (spr.rn == t.root) return;

    int recoal = nodes[spr.rn].parent;
    int *c = nodes[recoal].child;
    int other = (c[0] == spr.rn ? 1 : 0);
    int recomb_sib = c[other];
    int broke_parent = nodes[recoal].parent;

    nodes[recomb_sib].parent = broke_parent;
    int x = 0;
    if (broke_parent != -1) {
        c = nodes[broke_parent].child;
        x = (c[0] == recoal ? 0 : 1);
        nodes[broke_parent].child[x] = recomb_sib;
    }

    if (spr.cn == recoal) {
        nodes[recoal].child[other] = recomb_sib;
        nodes[recoal].parent = nodes[recomb_sib].parent;
        nodes[recomb_sib].parent = recoal;
        if (broke_parent != -1) nodes[broke_parent].child[x] = recoal;
    } else {
        nodes[recoal].child[other] = spr.cn;
        nodes[recoal].parent = nodes[spr.cn].parent;
        nodes[spr.cn].parent = recoal;
        int parent = nodes[recoal].parent;
        if (parent != -1) {
            c = nodes[parent].child;
            if (c[0] == spr.cn) c[0] = recoal;
            else c[1] = recoal;
        }
    }
    nodes[recoal].age = spr.ct;

    int root;
    if (spr.cn == t.root) root = recoal;
    else if (recoal == t.root) root = (spr.cn == recomb_sib) ? recoal : recomb_sib;
    else root = t.root;
    t.root = root;
}

// rename_node (thread.cpp:9-38)
static void rename_node(LTree &t, int src, int dest)
{
    if (src == dest) return;
    LNode *nodes = t.nodes.data();
    nodes[dest] = nodes[src];
    int parent = nodes[dest].parent;
    if (parent != -1) {
        int *c = nodes[parent].child;
        if (c[0] == src) c[0] = dest;
        else c[1] = dest;
    }
    int *c = nodes[dest].child;
    if (c[0] != -1) nodes[c[0]].parent = dest;
    if (c[1] != -1) nodes[c[1]].parent = dest;
}

// add_tree_branch (thread.cpp:47-101): new leaf = nleaves, displaced node =
// nnodes, new coalescence node = nnodes+1.
void add_tree_branch(LTree &t, int node, int time)
{
    LNode *nodes = t.nodes.data();
    int nleaves = t.num_leaves();
    int nnodes = t.nnodes;
    int newleaf = nleaves, displaced = nnodes, newcoal = nnodes + 1;
    int node2 = (node != newleaf ? node : displaced);
    int parent = nodes[node].parent;
    int parent2 = (parent != newleaf ? parent : displaced);

    if (newleaf < displaced) rename_node(t, newleaf, displaced);

    nodes[newleaf].parent = newcoal;
    nodes[newleaf].child[0] = nodes[newleaf].child[1] = -1;
    nodes[newleaf].age = 0;

    nodes[newcoal].parent = parent2;
    nodes[newcoal].child[0] = newleaf;
    nodes[newcoal].child[1] = node2;
    nodes[newcoal].age = time;

    nodes[node2].parent = newcoal;
    if (parent2 != -1) {
        int *c = nodes[parent2].child;
        if (c[0] == node2) c[0] = newcoal;
        else c[1] = newcoal;
    }

    t.nnodes = nnodes + 2;
    if (nodes[newcoal].parent == -1) t.root = newcoal;
    else t.root = (t.root != newleaf ? t.root : displaced);
}

// remove_tree_branch (thread.cpp:107-166)
void remove_tree_branch(LTree &t, int remove_leaf, int *displace)
{
    LNode *nodes = t.nodes.data();
    int nnodes = t.nnodes;
    int last_leaf = t.num_leaves() - 1;

    int remove_coal = nodes[remove_leaf].parent;
    int *c = nodes[remove_coal].child;
    int coal_child = (c[0] == remove_leaf ? c[1] : c[0]);
    int coal_parent = nodes[remove_coal].parent;
    nodes[coal_child].parent = coal_parent;
    if (coal_parent != -1) {
        c = nodes[coal_parent].child;
        if (c[0] == remove_coal) c[0] = coal_child;
        else c[1] = coal_child;
    }

    if (displace) {
        for (int i = 0; i < nnodes; i++) displace[i] = i;
        displace[remove_leaf] = -1;
        displace[remove_coal] = -1;
    }

    if (last_leaf != remove_leaf) {
        if (displace) displace[last_leaf] = remove_leaf;
        rename_node(t, last_leaf, remove_leaf);
    }

    int hole = last_leaf;
    if (remove_coal != nnodes - 2) {
        if (displace) displace[nnodes - 2] = hole;
        rename_node(t, nnodes - 2, hole);
        hole = remove_coal;
    }
    if (remove_coal != nnodes - 1) {
        if (displace) displace[nnodes - 1] = hole;
        rename_node(t, nnodes - 1, hole);
    }

    t.nnodes -= 2;
    int root = t.root;
    if (t.root == remove_coal) root = coal_child;
    if (root == nnodes - 2) root = last_leaf;
    if (root == nnodes - 1) root = hole;
    t.root = root;
}

// add_spr_branch (thread.cpp:176-266)
void add_spr_branch(const LTree &t, const LTree &last, int st_node, int st_time,
                           int last_node, Spr &spr, int *mapping, int newleaf, int displaced,
                           int newcoal)
{
    const LNode *nodes = t.nodes.data();
    const LNode *last_nodes = last.nodes.data();

    int last_node2 = (last_node != newleaf ? last_node : displaced);
    int node2 = (st_node != newleaf ? st_node : displaced);

    mapping[displaced] = mapping[newleaf];
    mapping[newleaf] = newleaf;
    mapping[newcoal] = newcoal;
    for (int i = newleaf + 1; i < t.nnodes; i++)
        if (mapping[i] == newleaf) mapping[i] = displaced;

    if (spr.rn == newleaf) spr.rn = displaced;
    if (spr.cn == newleaf) spr.cn = displaced;

    int recoal = nodes[mapping[spr.rn]].parent;
    if (recoal == newcoal) {
        if (mapping[last_node2] == node2) {
            spr.rn = newcoal;
        } else {
            int state_node = (st_node != newleaf) ? st_node : displaced;
            if (state_node == mapping[spr.rn]) {
                spr.cn = newleaf;
                spr.ct = st_time;
            }
        }
    } else {
        const int *c = nodes[recoal].child;
        if (c[0] == newcoal || c[1] == newcoal) {
            if (last_nodes[spr.cn].parent == newcoal) spr.cn = newcoal;
        }
    }

    if (last_nodes[spr.rn].parent == newcoal) {
        mapping[newcoal] = -1;
        int p = last_nodes[newcoal].parent;
        if (p != -1) mapping[p] = newcoal;
    } else {
        int x = newcoal;
        while (true) {
            int y = last_nodes[x].child[0];
            if (y == spr.cn || y == spr.rn) y = last_nodes[x].child[1];
            x = y;
            if (mapping[x] != -1) break;
        }
        mapping[newcoal] = nodes[mapping[x]].parent;
    }
}

void tree_states(const LTree &t, int ntimes, std::vector<int> &node, std::vector<int> &time)
{
    node.clear();
    time.clear();
    for (int i = 0; i < t.nnodes; i++) {
        int p = t.nodes[i].parent;
        int top = (p == -1) ? ntimes - 2 : t.nodes[p].age;
        for (int a = t.nodes[i].age; a <= top; a++) {
            node.push_back(i);
            time.push_back(a);
        }
    }
}

// AWAMD_CHECK_TREES=1 in the environment: every thread insertion and
// removal is followed by the structural checks of the whole ARG (the
// reference's assert_trees), which abort on a failure
bool check_edits()
{
    static const bool on = [] {
        const char *e = getenv("AWAMD_CHECK_TREES");
        return e && e[0] == '1';
    }();
    return on;
}

// add_arg_thread (thread.cpp:272-418).  Every input block becomes its own
// run of output blocks (the block with the new branch, then one block per
// recombination inside it), so the blocks are rebuilt in parallel; the SPR
// and mapping of each run's first block, which read the last tree of the
// run before, are rewritten in a second parallel pass once every tree is
// final.
void add_arg_thread(ArgTrees &tr, int ntimes, const int *path, int seqid,
                    const std::vector<int> &recomb_pos, const std::vector<int> &recomb_node,
                    const std::vector<int> &recomb_time)
{
    const int nleaves = tr.num_leaves();
    const int nnodes = tr.nnodes;
    const int nnodes2 = nnodes + 2;
    const int newleaf = nleaves, displaced = nnodes, newcoal = nnodes + 1;
    const int sc = tr.start_coord;
    const int nb = (int)tr.blocks.size();

    tr.seqids.push_back(seqid);
    tr.nnodes = nnodes2;

    // per input block: its columns, its recombinations [rfirst, rfirst+1)
    // and its first output block
    std::vector<int> bstart(nb + 1, sc), rfirst(nb + 1, 0), ofirst(nb + 1, 0);
    for (int b = 0; b < nb; b++) bstart[b + 1] = bstart[b] + tr.blocks[b].blocklen;
    for (int b = 0, r = 0; b <= nb; b++) {
        while (b < nb && r < (int)recomb_pos.size() && recomb_pos[r] < bstart[b]) r++;
        rfirst[b] = (b < nb) ? r : (int)recomb_pos.size();
        ofirst[b] = b + rfirst[b];
    }
    // first state (node, time) of each block and node of its last state
    // (newleaf renamed), both in the tree before the branch is added
    std::vector<int> n0(nb), t0(nb), lastn(nb);
    std::vector<LBlock> out(nb + recomb_pos.size());

    parallel_ranges(nb, [&](int lo, int hi) {
        // (node, time) of a state: off[node] is the node's first state
        std::vector<int> off, age0;
        auto state_node = [&](int st) {
            return (int)(std::upper_bound(off.begin(), off.end(), st) - off.begin()) - 1;
        };
        auto state_time = [&](int st) {
            const int n = state_node(st);
            return age0[n] + (st - off[n]);
        };
        for (int bi = lo; bi < hi; bi++) {
            const int start0 = bstart[bi], end = bstart[bi + 1];
            LBlock *blk = &out[ofirst[bi]];
            *blk = std::move(tr.blocks[bi]);
            {
                const LTree &t = blk->tree;
                off.resize(t.nnodes + 1);
                age0.resize(t.nnodes);
                off[0] = 0;
                for (int i = 0; i < t.nnodes; i++) {
                    const int p = t.nodes[i].parent;
                    const int top = (p == -1) ? ntimes - 2 : t.nodes[p].age;
                    age0[i] = t.nodes[i].age;
                    off[i + 1] = off[i] + std::max(0, top - age0[i] + 1);
                }
            }
            blk->tree.ensure_capacity(nnodes2);
            if (!blk->mapping.empty()) blk->mapping.resize(std::max<size_t>(blk->mapping.size(), nnodes2));
            const int s0 = path[start0 - sc];
            n0[bi] = state_node(s0);
            t0[bi] = state_time(s0);
            const int ln = state_node(path[end - 1 - sc]);
            lastn[bi] = (ln == newleaf) ? displaced : ln;
            add_tree_branch(blk->tree, n0[bi], t0[bi]);

            int start = start0;
            for (int ir = rfirst[bi]; ir < rfirst[bi + 1]; ir++) {
                const int pos = recomb_pos[ir];
                const LTree &tree = blk->tree;
                const int sn = state_node(path[pos - sc]), stt = state_time(path[pos - sc]);
                const int pl = path[pos - 1 - sc];
                const int lnode = state_node(pl), lt = state_time(pl);

                Spr spr2;
                spr2.rn = recomb_node[ir];
                spr2.rt = recomb_time[ir];
                if (spr2.rn == newleaf) spr2.rn = displaced;
                if (spr2.rn == -1) {
                    spr2.rn = newleaf;
                    spr2.cn = sn;
                    if (spr2.cn == newleaf) spr2.cn = displaced;
                    if (sn == lnode && stt > lt) spr2.cn = newcoal;
                } else {
                    if (stt > lt) spr2.cn = tree.nodes[newleaf].parent;
                    else spr2.cn = newleaf;
                }
                spr2.ct = stt;

                LBlock &nbk = out[ofirst[bi] + 1 + (ir - rfirst[bi])];
                nbk.mapping.resize(tree.nodes.size());
                for (int j = 0; j < nnodes2; j++) nbk.mapping[j] = j;
                nbk.mapping[tree.nodes[spr2.rn].parent] = -1;
                nbk.tree = tree;
                nbk.tree.nnodes = nnodes2;
                apply_spr(nbk.tree, spr2);
                nbk.spr = spr2;

                const int block_end = (ir + 1 < (int)recomb_pos.size()) ? std::min(recomb_pos[ir + 1], end) : end;
                blk->blocklen = pos - start;
                nbk.blocklen = block_end - pos;
                blk = &nbk;
                start = pos;
            }
        }
    });
    parallel_ranges(nb, [&](int lo, int hi) {
        for (int bi = std::max(lo, 1); bi < hi; bi++) {
            LBlock &blk = out[ofirst[bi]];
            if (blk.mapping.empty()) continue;
            add_spr_branch(blk.tree, out[ofirst[bi] - 1].tree, n0[bi], t0[bi], lastn[bi - 1], blk.spr,
                           blk.mapping.data(), newleaf, displaced, newcoal);
        }
    });
    tr.blocks = std::move(out);
    if (check_edits() && !check_trees(tr)) {
        fprintf(stderr, "argweaver_amd: add_arg_thread produced an invalid ARG\n");
        abort();
    }
}

// remove_null_spr / remove_null_sprs (local_tree.cpp:409-469)
void remove_null_sprs(ArgTrees &tr)
{
    std::vector<LBlock> out;
    out.reserve(tr.blocks.size());
    for (size_t i = 0; i < tr.blocks.size(); i++) {
        LBlock &B = tr.blocks[i];
        bool has_next = (i + 1 < tr.blocks.size());
        if (has_next && spr_is_null(tr.blocks[i + 1].spr)) {
            // merge B into the next block
            LBlock &N = tr.blocks[i + 1];
            if (B.mapping.empty()) {
                N.mapping.clear();
            } else {
                int nn = N.tree.nnodes;
                std::vector<int> m(nn);
                for (int j = 0; j < nn; j++) m[j] = (B.mapping[j] != -1) ? N.mapping[B.mapping[j]] : -1;
                for (int j = 0; j < nn; j++) N.mapping[j] = m[j];
                N.spr = B.spr;
            }
            N.blocklen += B.blocklen;
            continue;
        }
        out.push_back(std::move(B));
    }
    tr.blocks = std::move(out);
}

// remove_arg_thread (thread.cpp:425-554)
void remove_arg_thread(ArgTrees &tr, int remove_seqid)
{
    const int nnodes = tr.nnodes;
    const int nleaves = tr.num_leaves();
    const int last_leaf = nleaves - 1;

    int remove_leaf = -1;
    for (size_t i = 0; i < tr.seqids.size(); i++)
        if (tr.seqids[i] == remove_seqid) {
            remove_leaf = (int)i;
            break;
        }
    if (remove_leaf == -1) {
        fprintf(stderr, "argweaver_amd: remove_arg_thread: unknown sequence %d\n", remove_seqid);
        abort();
    }

    if (nnodes == 3) {
        int seqid = tr.seqids[1 - remove_leaf];
        int cap = (int)tr.blocks.front().tree.nodes.size();
        make_trunk(tr, tr.start_coord, tr.end_coord, cap);
        tr.seqids[0] = seqid;
        return;
    }

    // Each tree loses the branch on its own (first pass, in parallel); the
    // renaming it implies (displace) then rewrites the keys of the next
    // block's mapping and that block's SPR, and the values of the block's
    // own mapping, in that order (second pass, each block reading the
    // displacement of the block before)
    const int nb = (int)tr.blocks.size();
    struct Removed {
        int coal, coal_time, coal_child;
    };
    std::vector<Removed> rem(nb);
    std::vector<int> disp((size_t)nb * nnodes);
    parallel_ranges(nb, [&](int lo, int hi) {
        for (int i = lo; i < hi; i++) {
            LTree &tree = tr.blocks[i].tree;
            const LNode *nodes = tree.nodes.data();
            Removed &r = rem[i];
            r.coal = nodes[remove_leaf].parent;
            r.coal_time = nodes[r.coal].age;
            const int *c = nodes[r.coal].child;
            r.coal_child = (c[0] == remove_leaf ? c[1] : c[0]);
            remove_tree_branch(tree, remove_leaf, &disp[(size_t)i * nnodes]);
        }
    });
    parallel_ranges(nb, [&](int lo, int hi) {
        for (int i = lo; i < hi; i++) {
            std::vector<int> &mapping = tr.blocks[i].mapping;
            if (i > 0) {
                // the previous block's renaming: keys of this mapping, this SPR
                const int *displace = &disp[(size_t)(i - 1) * nnodes];
                const Removed &r = rem[i - 1];
                const LNode *nodes = tr.blocks[i - 1].tree.nodes.data();
                std::vector<int> &m2 = mapping;
                if (displace[last_leaf] != -1) m2[displace[last_leaf]] = m2[last_leaf];
                if (displace[nnodes - 2] != -1) m2[displace[nnodes - 2]] = m2[nnodes - 2];
                if (displace[nnodes - 1] != -1) m2[displace[nnodes - 1]] = m2[nnodes - 1];

                Spr &spr = tr.blocks[i].spr;
                const int coal_child = displace[r.coal_child];
                if (spr.rn == remove_leaf) {
                    spr = Spr{-1, -1, -1, -1};
                } else {
                    if (spr.rn == r.coal) spr.rn = coal_child;
                    else spr.rn = displace[spr.rn];
                    if (spr.rn == coal_child && nodes[coal_child].parent == -1) {
                        spr = Spr{-1, -1, -1, -1};
                    } else {
                        if (spr.cn == remove_leaf) {
                            spr.cn = coal_child;
                            spr.ct = r.coal_time;
                        } else if (spr.cn == r.coal) {
                            spr.cn = coal_child;
                        } else {
                            spr.cn = displace[spr.cn];
                        }
                        if (spr.rn == spr.cn) spr = Spr{-1, -1, -1, -1};
                    }
                }
            }
            // this block's renaming: values of its mapping
            const int *displace = &disp[(size_t)i * nnodes];
            if (!mapping.empty())
                for (int j = 0; j < nnodes - 2; j++)
                    mapping[j] = (mapping[j] != -1) ? displace[mapping[j]] : -1;
        }
    });

    tr.seqids[remove_leaf] = tr.seqids[last_leaf];
    tr.seqids.resize(nleaves - 1);
    tr.nnodes -= 2;
    remove_null_sprs(tr);
    if (check_edits() && !check_trees(tr)) {
        fprintf(stderr, "argweaver_amd: remove_arg_thread produced an invalid ARG\n");
        abort();
    }
}

static bool check_tree(const LTree &t)
{
    const LNode *nodes = t.nodes.data();
    for (int i = 0; i < t.nnodes; i++) {
        for (int x = 0; x < 2; x++) {
            int c = nodes[i].child[x];
            if (c != -1 && (c < 0 || c >= t.nnodes || nodes[c].parent != i)) return false;
        }
        if (nodes[i].parent == -1) {
            if (t.root != i) return false;
        } else if (nodes[i].parent < 0 || nodes[i].parent >= t.nnodes) {
            return false;
        }
    }
    return t.root >= 0 && nodes[t.root].parent == -1;
}

// assert_spr (local_tree.cpp:1689-1759) as a predicate
static bool check_spr(const LTree &last, const LTree &t, const Spr &spr, const int *mapping)
{
    const LNode *ln = last.nodes.data();
    const LNode *nd = t.nodes.data();
    if (spr.rn == -1) return false;
    if (mapping[spr.rn] == -1) return false;
    if (spr.rt > spr.ct) return false;
    if (ln[spr.rn].parent == -1) return false;
    if (spr.rt > ln[ln[spr.rn].parent].age || spr.rt < ln[spr.rn].age) return false;
    if (spr.ct < ln[spr.cn].age) return false;
    if (ln[spr.cn].parent != -1 && spr.ct > ln[ln[spr.cn].parent].age) return false;
    int recoal = nd[mapping[spr.rn]].parent;
    if (recoal == -1) return false;
    const int *c = nd[recoal].child;
    int other = (c[0] == mapping[spr.rn] ? c[1] : c[0]);
    if (mapping[spr.cn] != -1) {
        if (other != mapping[spr.cn]) return false;
    } else {
        int broken = ln[spr.rn].parent;
        const int *c2 = ln[broken].child;
        int last_other = (c2[0] == spr.rn ? c2[1] : c2[0]);
        if (mapping[last_other] == -1) return false;
        if (nd[mapping[last_other]].parent != recoal) return false;
    }
    for (int i = 0; i < last.nnodes; i++) {
        int i2 = mapping[i];
        if (i2 != -1 && ln[i].age != nd[i2].age) return false;
        if (ln[i].child[0] == -1 && (i2 == -1 || nd[i2].child[0] != -1)) return false;
    }
    return spr.rn != spr.cn;
}

bool check_trees(const ArgTrees &tr)
{
    if (tr.blocks.empty()) return true;
    if (!spr_is_null(tr.blocks[0].spr) || !tr.blocks[0].mapping.empty()) return false;
    int seqlen = 0;
    const LTree *last = nullptr;
    for (const LBlock &B : tr.blocks) {
        seqlen += B.blocklen;
        if (B.blocklen < 0 || !check_tree(B.tree)) return false;
        if (last) {
            if (spr_is_null(B.spr)) {
                std::vector<char> seen(B.tree.nnodes, 0);
                for (int i = 0; i < B.tree.nnodes; i++) {
                    int m = B.mapping[i];
                    if (m == -1 || seen[m]) return false;
                    seen[m] = 1;
                }
            } else if (!check_spr(*last, B.tree, B.spr, B.mapping.data())) {
                return false;
            }
        }
        last = &B.tree;
    }
    return seqlen == tr.length();
}

void flatten(const ArgTrees &tr, FlatTrees &f)
{
    const int B = (int)tr.blocks.size();
    const int N = tr.nnodes;
    f.nblocks = B;
    f.nnodes = N;
    f.start_coord = tr.start_coord;
    f.seqlen = tr.length();
    f.parent.resize((size_t)B * N);
    f.child0.resize((size_t)B * N);
    f.child1.resize((size_t)B * N);
    f.age.resize((size_t)B * N);
    f.spr.resize((size_t)B * 4);
    f.mapping.resize((size_t)B * N);
    f.start.resize(B);
    f.len.resize(B);
    int pos = 0;
    for (int b = 0; b < B; b++) {
        f.start[b] = pos;
        f.len[b] = tr.blocks[b].blocklen;
        pos += f.len[b];
    }
    parallel_ranges(B, [&](int lo, int hi) {
        for (int b = lo; b < hi; b++) {
            const LBlock &blk = tr.blocks[b];
            for (int i = 0; i < N; i++) {
                const LNode &nd = blk.tree.nodes[i];
                f.parent[(size_t)b * N + i] = nd.parent;
                f.child0[(size_t)b * N + i] = nd.child[0];
                f.child1[(size_t)b * N + i] = nd.child[1];
                f.age[(size_t)b * N + i] = nd.age;
                f.mapping[(size_t)b * N + i] = blk.mapping.empty() ? -1 : blk.mapping[i];
            }
            f.spr[4 * b + 0] = blk.spr.rn;
            f.spr[4 * b + 1] = blk.spr.rt;
            f.spr[4 * b + 2] = blk.spr.cn;
            f.spr[4 * b + 3] = blk.spr.ct;
        }
    });
}

}  // namespace aw
