// thread_path.cpp -- re-threading an internal branch of the ARG, on the
// host: removal paths (a branch followed from tree to tree through the
// SPRs), removing a branch along such a path, which leaves every local tree
// in subtree/maintree form (a virtual root above the removed subtree and the
// rest of the tree), and adding a sampled thread of that subtree back.
//   removal paths             thread.cpp:562-1020
//   remove along a path       thread.cpp:1481-1679
//   add a thread path         thread.cpp:1204-1449
//   internal state spaces     states.cpp:109-166
// Every rand() draw is made in the reference's order, so the same seed
// gives the same path and the same ARG.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/argweaver_amd.h"
#include "local_trees.h"
#include "thread_path.h"

namespace aw {

namespace {

inline double frand_() { return rand() / double(RAND_MAX); }

// sample() (common.h:272-290)
int sample_w(const double *w, int n)
{
    double total = 0.0;
    for (int i = 0; i < n; i++) total += w[i];
    const double pick = frand_() * total;
    double x = 0.0;
    for (int i = 0; i < n; i++) {
        x += w[i];
        if (x >= pick) return i;
    }
    return n - 1;
}

double logadd_(double a, double b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    return std::max(a, b) + log1p(exp(-fabs(a - b)));
}

// logsum with the reference's cut-off (common.h:182-200)
double logsum_(const double *v, int n)
{
    if (n == 0) return 1.0;
    double mx = v[0];
    for (int i = 1; i < n; i++) mx = std::max(mx, v[i]);
    double s = 0.0;
    for (int i = 0; i < n; i++)
        if (v[i] - mx > -15) s += exp(v[i] - mx);
    return mx + log(s);
}

int sibling_of(const LTree &t, int node)
{
    const int p = t.nodes[node].parent;
    if (p == -1) return -1;
    const int *c = t.nodes[p].child;
    return c[0] == node ? c[1] : c[0];
}

}  // namespace

int recoal_node(const LTree &t, const int *mapping)
{
    std::vector<char> mapped(t.nnodes, 0);
    for (int i = 0; i < t.nnodes; i++)
        if (mapping[i] != -1) mapped[mapping[i]] = 1;
    for (int i = 0; i < t.nnodes; i++)
        if (!mapped[i]) return i;
    return -1;
}

// get_next_removal_nodes (thread.cpp:562-591): where the branch `node` of
// tree t1 continues in the next tree (out[0]) and, when the SPR re-coalesces
// onto it, the second continuation (out[1], else -1)
void next_removal_nodes(const LTree &t1, const Spr &spr2, const int *mapping2, int node,
                        int out[2], int recoal)
{
    if (recoal == -1) recoal = recoal_node(t1, mapping2);
    out[0] = mapping2[node];
    if (out[0] == -1) {
        const int sib = sibling_of(t1, spr2.rn);
        out[0] = (spr2.cn == sib) ? recoal : mapping2[sib];
    }
    out[1] = (spr2.cn == node) ? recoal : -1;
}

// get_prev_removal_nodes (thread.cpp:609-647)
void prev_removal_nodes(const LTree &t1, const Spr &spr2, const int *mapping2,
                        const int *inv_mapping, int node, int out[2])
{
    out[0] = inv_mapping[node];
    const int sib = sibling_of(t1, spr2.rn);
    if (out[0] == -1) {
        out[0] = spr2.cn;
        out[1] = (sib == spr2.cn) ? t1.nodes[sib].parent : -1;
    } else {
        out[1] = (mapping2[sib] == node && sib != spr2.cn) ? t1.nodes[sib].parent : -1;
    }
}

static void inverse_mapping(const std::vector<int> &m, int n, std::vector<int> &inv)
{
    inv.assign(n, -1);
    for (int i = 0; i < n; i++)
        if (m[i] != -1) inv[m[i]] = i;
}

void removal_leaf_path(const ArgTrees &tr, int node, int *path)
{
    int i = 0;
    path[i++] = node;
    for (size_t b = 1; b < tr.blocks.size(); b++) {
        const LBlock &B = tr.blocks[b];
        int nx[2];
        next_removal_nodes(tr.blocks[b - 1].tree, B.spr, B.mapping.data(), path[i - 1], nx, -1);
        path[i++] = nx[0];
    }
}

// sample_arg_removal_path_forward / _backward (thread.cpp:703-769).  The
// choice at a fork compares the integer rand() with prob_switch, as the
// reference does (so a fork keeps its first branch unless rand() returns 0),
// and draws only at forks.
static void removal_path_forward(const ArgTrees &tr, size_t b0, int node, int *path, int i,
                                 double prob_switch)
{
    path[i++] = node;
    for (size_t b = b0 + 1; b < tr.blocks.size(); b++) {
        const LBlock &B = tr.blocks[b];
        int nx[2];
        next_removal_nodes(tr.blocks[b - 1].tree, B.spr, B.mapping.data(), path[i - 1], nx, -1);
        const int j = (nx[1] == -1) ? 0 : int(rand() < prob_switch);
        path[i++] = nx[j];
    }
}

static void removal_path_backward(const ArgTrees &tr, size_t b0, int node, int *path, int i,
                                  double prob_switch)
{
    path[i--] = node;
    std::vector<int> inv;
    for (size_t b = b0; b >= 1; b--) {
        const LBlock &B2 = tr.blocks[b];
        const LTree &t1 = tr.blocks[b - 1].tree;
        inverse_mapping(B2.mapping, t1.nnodes, inv);
        int pv[2];
        prev_removal_nodes(t1, B2.spr, B2.mapping.data(), inv.data(), path[i + 1], pv);
        const int j = (pv[1] == -1) ? 0 : int(rand() < prob_switch);
        path[i--] = pv[j];
    }
}

void removal_path_at(const ArgTrees &tr, int node, int pos, int *path, double prob_switch)
{
    size_t b = 0;
    int end = tr.start_coord;
    for (; b < tr.blocks.size(); b++) {
        const int start = end;
        end += tr.blocks[b].blocklen;
        if (start <= pos && pos < end) break;
    }
    removal_path_forward(tr, b, node, path, (int)b, prob_switch);
    removal_path_backward(tr, b, node, path, (int)b, prob_switch);
}

void removal_path_from_start(const ArgTrees &tr, int node, int *path)
{
    removal_path_forward(tr, 0, node, path, 0, .5);
}

// sample_arg_removal_path_recomb (thread.cpp:802-912): a forward pass over
// the branches of every tree, weighting the branch that carries the next
// recombination by recomb_preference, then a stochastic traceback
void removal_path_recomb(const ArgTrees &tr, double pref, int *path)
{
    const int ntrees = (int)tr.blocks.size(), N = tr.nnodes;
    std::vector<double> fw((size_t)ntrees * N), trans((size_t)ntrees * N, 0.0);
    std::vector<int> back((size_t)ntrees * N * 2, -1);
    for (int j = 0; j < N; j++) fw[j] = 1.0 / N;
    std::vector<int> inv;
    for (int i = 1; i < ntrees; i++) {
        const LBlock &B = tr.blocks[i];
        const LTree &last = tr.blocks[i - 1].tree;
        const int recoal = recoal_node(last, B.mapping.data());
        inverse_mapping(B.mapping, N, inv);
        for (int j = 0; j < N; j++) {
            int nx[2];
            next_removal_nodes(last, B.spr, B.mapping.data(), j, nx, recoal);
            trans[(size_t)(i - 1) * N + j] = (nx[1] != -1) ? .5 : 1.0;
            prev_removal_nodes(last, B.spr, B.mapping.data(), inv.data(), j,
                               &back[((size_t)i * N + j) * 2]);
        }
        double norm = 0.0;
        for (int j = 0; j < N; j++) {
            double sum = 0.0;
            for (int ki = 0; ki < 2; ki++) {
                const int k = back[((size_t)i * N + j) * 2 + ki];
                if (k == -1) continue;
                sum += trans[(size_t)(i - 1) * N + j] * fw[(size_t)(i - 1) * N + k];
            }
            const double emit = (!spr_is_null(B.spr) && B.spr.rn == j) ? pref : 1.0 - pref;
            fw[(size_t)i * N + j] = sum * emit;
            norm += fw[(size_t)i * N + j];
        }
        for (int j = 0; j < N; j++) fw[(size_t)i * N + j] /= norm;
    }
    int i = ntrees - 1;
    path[i] = sample_w(&fw[(size_t)i * N], N);
    int j = path[i];
    for (i--; i >= 0; i--) {
        const int *bp = &back[((size_t)(i + 1) * N + j) * 2];
        if (bp[1] == -1) {
            j = path[i] = bp[0];
        } else {
            double probs[2];
            probs[0] = fw[(size_t)i * N + bp[0]] * trans[(size_t)i * N + j];
            probs[1] = fw[(size_t)i * N + bp[1]] * trans[(size_t)i * N + j];
            j = path[i] = bp[sample_w(probs, 2)];
        }
    }
}

// count_arg_removal_paths (thread.cpp:914-949): log number of removal
// paths that end at each branch of each tree, with back pointers
static void count_paths(const ArgTrees &tr, std::vector<double> &counts, std::vector<int> &back)
{
    const int ntrees = (int)tr.blocks.size(), N = tr.nnodes;
    counts.assign((size_t)ntrees * N, 0.0);
    back.assign((size_t)ntrees * N * 2, -1);
    std::vector<int> inv;
    for (int i = 1; i < ntrees; i++) {
        const LBlock &B = tr.blocks[i];
        const LTree &last = tr.blocks[i - 1].tree;
        inverse_mapping(B.mapping, N, inv);
        for (int j = 0; j < N; j++) {
            int *p = &back[((size_t)i * N + j) * 2];
            prev_removal_nodes(last, B.spr, B.mapping.data(), inv.data(), j, p);
            const double c0 = counts[(size_t)(i - 1) * N + p[0]];
            counts[(size_t)i * N + j] =
                (p[1] == -1) ? c0 : logadd_(c0, counts[(size_t)(i - 1) * N + p[1]]);
        }
    }
}

double removal_path_uniform(const ArgTrees &tr, int *path)
{
    const int ntrees = (int)tr.blocks.size(), N = tr.nnodes;
    std::vector<double> counts;
    std::vector<int> back;
    count_paths(tr, counts, back);
    const double *last = &counts[(size_t)(ntrees - 1) * N];
    const double norm = logsum_(last, N);
    std::vector<double> w(N);
    for (int j = 0; j < N; j++) w[j] = exp(last[j] - norm);
    path[ntrees - 1] = sample_w(w.data(), N);
    for (int i = ntrees - 1; i > 0; i--) {
        const int *p = &back[((size_t)i * N + path[i]) * 2];
        if (p[1] == -1) {
            path[i - 1] = p[0];
        } else {
            const double p1 = counts[(size_t)(i - 1) * N + p[0]];
            const double p2 = counts[(size_t)(i - 1) * N + p[1]];
            path[i - 1] = (log(frand_()) < (p1 - logadd_(p1, p2))) ? p[0] : p[1];
        }
    }
    return norm;
}

double count_removal_paths(const ArgTrees &tr)
{
    std::vector<double> counts;
    std::vector<int> back;
    count_paths(tr, counts, back);
    const int N = tr.nnodes;
    return logsum_(&counts[(tr.blocks.size() - 1) * N], N);
}

// get_coal_states_internal (states.cpp:109-166): every (node, time) of the
// main tree with time >= max(minage, age of the subtree root); the branch
// under the virtual root may rise to ntimes - 2.  A tree that is not in
// subtree/maintree form (its root below ntimes) has no states.
void internal_states(const LTree &t, int ntimes, int minage, std::vector<int> &node,
                     std::vector<int> &time)
{
    node.clear();
    time.clear();
    const LNode *nd = t.nodes.data();
    if (nd[t.root].age < ntimes) return;
    const int sub = nd[t.root].child[0];
    minage = std::max(minage, nd[sub].age);
    std::vector<char> skip(t.nnodes, 0);
    skip[t.root] = 1;
    std::vector<int> stack{sub};
    while (!stack.empty()) {
        const int x = stack.back();
        stack.pop_back();
        skip[x] = 1;
        if (nd[x].child[0] != -1) {
            stack.push_back(nd[x].child[0]);
            stack.push_back(nd[x].child[1]);
        }
    }
    for (int i = 0; i < t.nnodes; i++) {
        if (skip[i]) continue;
        const int p = nd[i].parent;
        const int top = (p == t.root) ? ntimes - 2 : nd[p].age;
        for (int a = std::max(nd[i].age, minage); a <= top; a++) {
            node.push_back(i);
            time.push_back(a);
        }
    }
}

// get_actual_nodes (thread.cpp:1452-1478): names of this block's nodes in
// the last tree of the run of null-SPR blocks that starts here
static size_t actual_nodes(const ArgTrees &tr, size_t b, std::vector<int> &names)
{
    const int N = tr.blocks[b].tree.nnodes;
    names.resize(N);
    for (int i = 0; i < N; i++) names[i] = i;
    size_t b2 = b + 1;
    for (; b2 < tr.blocks.size() && spr_is_null(tr.blocks[b2].spr); b2++) {
        const std::vector<int> &m = tr.blocks[b2].mapping;
        for (int i = 0; i < N; i++) names[i] = (m[i] != -1) ? m[names[i]] : -1;
    }
    return b2 - 1;
}

void remove_thread_path(ArgTrees &tr, const int *path, int maxtime, int *original_thread)
{
    const int L = tr.length();
    std::vector<int> onode, otime;
    if (original_thread) {
        onode.assign(L, -1);
        otime.assign(L, -1);
    }
    int end = tr.start_coord;
    LTree *last = nullptr;
    for (size_t i = 0; i < tr.blocks.size(); i++) {
        LBlock &B = tr.blocks[i];
        LTree &t = B.tree;
        LNode *nd = t.nodes.data();
        const int start = end;
        end += B.blocklen;
        const int rnode = path[i];
        if (rnode == t.root) {
            // the path has left the tree through its root: nothing to edit
            if (!B.mapping.empty() && path[i - 1] != last->root) B.mapping[last->root] = -1;
            last = &t;
            continue;
        }
        // subtree/maintree form: the removed branch hangs from a new root
        const int broken = nd[rnode].parent;
        const int coal_time = nd[broken].age;
        const int broken_child = sibling_of(t, rnode);
        apply_spr(t, Spr{rnode, nd[rnode].age, t.root, maxtime});
        const int sub = rnode;
        const int main = sibling_of(t, sub);
        int *c = nd[t.root].child;
        if (c[0] == main) {
            c[0] = sub;
            c[1] = main;
        }
        if (!B.mapping.empty() && path[i - 1] != last->root) B.mapping[last->root] = t.root;
        if (original_thread)
            for (int j = start; j < end; j++) {
                onode[j - tr.start_coord] = broken_child;
                otime[j - tr.start_coord] = coal_time;
            }
        last = &t;
        if (i + 1 == tr.blocks.size()) continue;

        // the SPR into the next tree
        LBlock &N = tr.blocks[i + 1];
        Spr &spr = N.spr;
        std::vector<int> &mapping = N.mapping;
        if (spr.rn == rnode) {
            spr = Spr{-1, -1, -1, -1};   // recombination on the removed branch
            continue;
        }
        if (spr.rn == broken) spr.rn = broken_child;
        int nx[2];
        next_removal_nodes(t, spr, mapping.data(), path[i], nx, -1);
        if (spr.cn == rnode) {
            if (path[i + 1] == nx[0]) {
                // the path takes the lower branch
                if (spr.rn == broken_child) {
                    spr = Spr{-1, -1, -1, -1};   // the SPR became a bubble
                    continue;
                }
                spr.cn = broken_child;           // mediated coalescence
                spr.ct = coal_time;
            }
        } else if (spr.cn == broken) {
            spr.cn = broken_child;
        }
        if (spr.rn == spr.cn) {
            spr = Spr{-1, -1, -1, -1};
            continue;
        }
        mapping[nd[spr.rn].parent] = -1;
    }

    if (original_thread) {
        const int ntimes = maxtime - 1;
        std::vector<int> names, sn, st;
        int e = tr.start_coord;
        for (size_t b = 0; b < tr.blocks.size(); b++) {
            const int s0 = e;
            e += tr.blocks[b].blocklen;
            const size_t b2 = actual_nodes(tr, b, names);
            const LTree &t2 = tr.blocks[b2].tree;
            internal_states(t2, ntimes, 0, sn, st);
            // node-major state index of (node, time)
            std::vector<int> first(t2.nnodes, -1), lo(t2.nnodes, 0);
            for (int k = (int)sn.size() - 1; k >= 0; k--) {
                first[sn[k]] = k;
                lo[sn[k]] = st[k];
            }
            for (int j = s0; j < e; j++) {
                const int x = j - tr.start_coord;
                if (sn.empty() || onode[x] < 0) {
                    original_thread[x] = 0;
                } else {
                    const int n2 = names[onode[x]];
                    original_thread[x] = first[n2] + otime[x] - lo[n2];
                }
            }
        }
    }
    remove_null_sprs(tr);
}

void add_spr_branch_internal(const LTree &t, const LTree &last, int st_node, int st_time,
                                    int last_node, Spr &spr, std::vector<int> &mapping,
                                    int sub, int last_sub)
{
    const LNode *nd = t.nodes.data();
    const LNode *ln = last.nodes.data();
    const int last_newcoal = ln[last_sub].parent;
    int newcoal;
    if (st_node != -1) {
        newcoal = nd[sub].parent;
    } else if (mapping[last_sub] != -1) {
        newcoal = nd[mapping[last_sub]].parent;
    } else {
        newcoal = nd[mapping[sibling_of(last, spr.rn)]].parent;
    }
    mapping[last_newcoal] = newcoal;
    const int recoal = nd[mapping[spr.rn]].parent;
    if (recoal == newcoal) {
        if (mapping[last_node] == st_node) {
            spr.rn = last_newcoal;   // the recombination sits above the thread
        } else if (spr.ct == ln[last_newcoal].age && st_node == mapping[spr.rn]) {
            // mediated coalescence
            if (st_time < ln[last_sub].age) {
                spr.cn = sibling_of(last, spr.rn);
            } else {
                spr.cn = last_sub;
            }
            spr.ct = st_time;
        }
    } else {
        const int *c = nd[recoal].child;
        if ((c[0] == newcoal || c[1] == newcoal) && ln[spr.cn].parent == last_newcoal)
            spr.cn = last_newcoal;
    }
    if (ln[spr.rn].parent == last_newcoal) {
        mapping[last_newcoal] = -1;
        const int p = ln[last_newcoal].parent;
        if (p != -1) mapping[p] = newcoal;
    } else {
        const int x = last_newcoal;
        int y = ln[x].child[0];
        if (y == spr.cn) y = ln[x].child[1];
        if (mapping[y] == -1) y = sibling_of(last, spr.rn);
        if (y == spr.cn) y = ln[x].child[1];
        mapping[last_newcoal] = nd[mapping[y]].parent;
    }
}

int add_thread_path(ArgTrees &tr, int ntimes, int minage, const int *thread_path,
                    const std::vector<int> &recomb_pos, const std::vector<int> &recomb_node,
                    const std::vector<int> &recomb_time)
{
    std::vector<int> sn, st;
    int end = tr.start_coord;
    size_t ir = 0;
    const int nr = (int)recomb_pos.size();
    int last_node = -1, last_time = -1, last_sub = -1;
    size_t last_b = 0;
    bool have_last = false;
    std::vector<LBlock> out;
    out.reserve(tr.blocks.size() + recomb_pos.size());
    for (size_t bi = 0; bi < tr.blocks.size(); bi++) {
        out.push_back(std::move(tr.blocks[bi]));
        LBlock *B = &out.back();
        int start = end;
        end += B->blocklen;
        const int sub = B->tree.nodes[B->tree.root].child[0];
        internal_states(B->tree, ntimes, minage, sn, st);
        const int nstates = (int)sn.size();
        int s_node = -1, s_time = -1;
        if (B->tree.nodes[B->tree.root].age > ntimes) {
            if (nstates == 0) return 1;
            const int x = thread_path[start - tr.start_coord];
            s_node = sn[x];
            s_time = st[x];
            apply_spr(B->tree, Spr{sub, B->tree.nodes[sub].age, s_node, s_time});
        }
        if (have_last && !B->mapping.empty() && last_node != -1)
            add_spr_branch_internal(B->tree, out[last_b].tree, s_node, s_time, last_node, B->spr,
                                    B->mapping, sub, last_sub);
        // split the block at each new recombination inside it
        for (; ir < (size_t)nr && recomb_pos[ir] < end; ir++) {
            const int pos = recomb_pos[ir];
            if (nstates == 0) return 1;
            const LTree &t = out.back().tree;
            const int x = thread_path[pos - tr.start_coord];
            const int xl = thread_path[pos - 1 - tr.start_coord];
            const int cur_node = sn[x], cur_time = st[x];
            const int prev_node = sn[xl], prev_time = st[xl];
            const int newcoal = t.nodes[sub].parent;
            Spr s2;
            s2.rn = recomb_node[ir];
            s2.rt = recomb_time[ir];
            if (s2.rn == sub) {
                s2.cn = cur_node;
                if (cur_node == prev_node && cur_time > prev_time) s2.cn = newcoal;
            } else {
                s2.cn = (cur_time > prev_time) ? t.nodes[sub].parent : sub;
            }
            s2.ct = cur_time;
            LBlock NB;
            NB.tree = t;
            NB.mapping.resize(t.nnodes);
            for (int j = 0; j < t.nnodes; j++) NB.mapping[j] = j;
            NB.mapping[t.nodes[s2.rn].parent] = -1;
            apply_spr(NB.tree, s2);
            NB.spr = s2;
            const int block_end = (ir + 1 < (size_t)nr) ? std::min(recomb_pos[ir + 1], end) : end;
            out.back().blocklen = pos - start;
            NB.blocklen = block_end - pos;
            out.push_back(std::move(NB));
            start = pos;
            s_node = cur_node;
            s_time = cur_time;
        }
        last_b = out.size() - 1;
        last_node = s_node;
        last_time = s_time;
        last_sub = sub;
        have_last = true;
    }
    (void)last_time;
    tr.blocks = std::move(out);
    return check_trees(tr) ? 0 : 2;
}

}  // namespace aw

using namespace aw;

extern "C" {

void arghmm_sample_arg_removal_path(LocalTrees *trees, int node, int *path)
{
    removal_path_from_start(*reinterpret_cast<ArgTrees *>(trees), node, path);
}

void arghmm_sample_arg_removal_path2(LocalTrees *trees, int node, int pos, int *path)
{
    // the reference's default switch probability for this form (thread.h:114-116)
    removal_path_at(*reinterpret_cast<ArgTrees *>(trees), node, pos, path, .1);
}

void arghmm_sample_arg_removal_leaf_path(LocalTrees *trees, int node, int *path)
{
    removal_leaf_path(*reinterpret_cast<ArgTrees *>(trees), node, path);
}

void arghmm_sample_arg_removal_path_recomb(LocalTrees *trees, double recomb_preference, int *path)
{
    removal_path_recomb(*reinterpret_cast<ArgTrees *>(trees), recomb_preference, path);
}

void arghmm_remove_arg_thread_path(LocalTrees *trees, int *removal_path, int maxtime)
{
    remove_thread_path(*reinterpret_cast<ArgTrees *>(trees), removal_path, maxtime, nullptr);
}

void arghmm_remove_arg_thread_path2(LocalTrees *trees, int *removal_path, int maxtime,
                                    int *original_thread)
{
    remove_thread_path(*reinterpret_cast<ArgTrees *>(trees), removal_path, maxtime,
                       original_thread);
}

void arghmm_get_thread_times(LocalTrees *trees, int ntimes, int *path, int *path_times)
{
    const ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    std::vector<int> sn, st;
    int end = tr.start_coord;
    for (const LBlock &B : tr.blocks) {
        const int start = end;
        end += B.blocklen;
        internal_states(B.tree, ntimes, 0, sn, st);
        for (int i = start; i < end; i++) path_times[i] = st[path[i]];
    }
}

}  // extern "C"
