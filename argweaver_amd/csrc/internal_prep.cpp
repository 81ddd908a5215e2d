// internal_prep.cpp -- the per-block inputs of the threading HMM for an
// internal branch, prepared on the host: every local tree is in
// subtree/maintree form (a virtual root over the removed subtree, child 0,
// and the rest of the tree, child 1), and the subtree is threaded back onto
// the main tree.  The output is the lane-run layout the device forward and
// traceback kernels read (the same arrays the device preparation kernel
// writes for an external thread), so the per-column recursion runs on the
// GPU unchanged.
//   state spaces      states.cpp:109-166      lineages  local_tree.cpp:82-131
//   TransMatrix       trans.cpp:26-115 (internal), trans.h:89-128 (minage)
//   switch            trans.cpp:156-275, 316-534, 538-739 (internal)
//   priors            trans.cpp:785-831       emissions emit.cpp:650-869
// A tree that is not in subtree/maintree form (the removal path left it
// through its root) has no states; it gets one placeholder state whose
// value is carried from column to column (sample_thread.cpp:186-205).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/argweaver_amd.h"
#include "engine.h"
#include "host_util.h"
#include "internal_prep.h"
#include "thread_path.h"

namespace aw {
namespace {

struct SwitchOut {
    std::vector<int> determ;
    std::vector<double> determprob, recombrow, recoalrow;
    int recombsrc = -1, recoalsrc = -1;
    double get(int i, int j) const
    {
        if (i == recoalsrc) return recoalrow[j];
        if (i == recombsrc) return recombrow[j];
        return determ[i] == j ? determprob[i] : 0.0;
    }
};

struct BlockStates {
    std::vector<int> node, time;
    std::vector<int> first, lo, hi;   // per node: first state index, time range (-1: none)
    int nstates() const { return (int)node.size(); }
    int lookup(int n, int t) const
    {
        if (n < 0 || first[n] < 0 || t < lo[n] || t > hi[n]) return -1;
        return first[n] + t - lo[n];
    }
};

void make_states(const LTree &t, int T, int minage, BlockStates &bs)
{
    internal_states(t, T, minage, bs.node, bs.time);
    bs.first.assign(t.nnodes, -1);
    bs.lo.assign(t.nnodes, 0);
    bs.hi.assign(t.nnodes, -1);
    for (int k = 0; k < bs.nstates(); k++) {
        const int n = bs.node[k];
        if (bs.first[n] < 0) {
            bs.first[n] = k;
            bs.lo[n] = bs.time[k];
        }
        bs.hi[n] = bs.time[k];
    }
}

inline int sub_root(const LTree &t) { return t.nodes[t.root].child[0]; }
inline int main_root(const LTree &t) { return t.nodes[t.root].child[1]; }
inline bool partial(const LTree &t, int T) { return t.nodes[t.root].age >= T; }

// count_lineages_internal (local_tree.cpp:82-131)
void lineages_internal(const LTree &t, int T, Lineages &L)
{
    for (int i = 0; i < T; i++) L.nb[i] = L.nr[i] = L.nc[i] = 0;
    const LNode *nd = t.nodes.data();
    const int sub = sub_root(t);
    const int minage = nd[sub].age;
    for (int i = 0; i < t.nnodes; i++) {
        if (i == sub || i == t.root) continue;
        const int p = nd[i].parent;
        const int page = (p == t.root) ? T - 2 : nd[p].age;
        for (int j = nd[i].age; j < page; j++) {
            L.nb[j]++;
            L.nr[j]++;
            L.nc[j]++;
        }
        L.nr[page]++;
        L.nc[page]++;
        if (p == t.root) L.nb[page]++;
    }
    for (int i = 0; i < minage; i++) {
        L.nb[i]--;
        L.nc[i]--;
        L.nr[i]--;
    }
    L.nb[T - 1] = 1;
}

// get_treelen_internal (local_tree.cpp:158-176): the branches under the
// virtual root, its own two branches excluded
double treelen_internal(const LTree &t, const double *times)
{
    double len = 0.0;
    for (int i = 0; i < t.nnodes; i++) {
        const int p = t.nodes[i].parent;
        if (p == t.root || p == -1) continue;
        len += times[t.nodes[p].age] - times[t.nodes[i].age];
    }
    return len;
}

// calc_transition_probs with internal (trans.cpp:26-115)
void trans_internal(const LTree &t, const Model &m, const Lineages &L, TransVec &tv)
{
    const int T = m.T;
    double rates_a[2 * MAXT + 1], C_a[2 * MAXT + 2];
    double *rates = rates_a + 1, *C = C_a + 2;
    rates[-1] = 0.0;
    for (int i = 0; i < 2 * T - 1; i++) rates[i] = m.csteps[i] * L.nb[i / 2] / (2.0 * m.pop[i / 2]);
    C[-2] = C[-1] = 0.0;
    for (int b = 0; b < 2 * T - 1; b++) C[b] = C[b - 1] + rates[b];
    const int sub = sub_root(t), mr = main_root(t);
    const int ri = t.nodes[mr].age;
    const double root_age = m.times[ri];
    const double len = treelen_internal(t, m.times) - m.times[t.nodes[sub].age];
    for (int b = 0; b < T; b++) {
        tv.D[b] = tv.E[b] = tv.lnB[b] = tv.lnE2[b] = tv.lnNegG1[b] = 0.0;
        tv.G2[b] = tv.G3[b] = tv.lnG4[b] = tv.norecombs[b] = 0.0;
    }
    for (int b = 0; b < T - 1; b++) {
        const int nb = L.nb[b], nr = L.nr[b], nc = L.nc[b];
        double len2 = len + m.times[b], len2_b;
        if (b > ri) {
            len2 += m.times[b] - root_age;
            len2_b = len2 + m.steps[b];
        } else {
            len2_b = len2 + m.steps[ri];
        }
        const double term = C[2 * b - 1] + log(m.steps[b] * (nb + 1.0) / (nr + 1.0));
        tv.lnB[b] = (b == 0) ? term : logadd(tv.lnB[b - 1], term);
        const double e2 =
            -C[2 * b - 2] + (b < T - 2 ? log(1 - exp(-rates[2 * b] - rates[2 * b - 1])) : 0.0);
        tv.lnE2[b] = e2;
        tv.lnG4[b] = e2;
        tv.lnNegG1[b] = C[2 * b - 1] +
            log(-m.steps[b] * ((nb / (nr + 1.0 + int(b < ri))) - (nb + 1.0) / (nr + 1.0)));
        const double q = (b < T - 2 ? 1.0 - exp(-rates[2 * b]) : 1.0);
        tv.G2[b] = q * m.steps[b] * (nb + 1.0) / (nr + 1.0);
        tv.G3[b] = q * m.steps[b] * (nb / (nr + 1.0 + int(b < ri)));
        tv.D[b] = (1.0 - exp(-m.rho * len2)) / len2_b;
        tv.E[b] = 1.0 / nc;
        tv.norecombs[b] = exp(-fmax(m.rho * len2, m.rho));
    }
}

// TransMatrix::get_time with a minimum age (trans.h:89-128)
double get_time_min(const TransVec &tv, int a, int b, int c, int minage, bool same)
{
    if (a < minage || b < minage) return 0.0;
    const double term1 = tv.D[a] * tv.E[b];
    const double mt = minage > 0 ? exp(tv.lnG4[b] + tv.lnB[minage - 1]) : 0.0;
    const double low = b > 0 ? exp(tv.lnE2[b] + tv.lnB[b - 1]) : 0.0;
    if (!same) {
        if (a < b) return term1 * (exp(tv.lnE2[b] + tv.lnB[a]) - exp(tv.lnE2[b] + tv.lnNegG1[a]) - mt);
        if (a == b) return term1 * (low + tv.G3[b] - mt);
        return term1 * (low + tv.G2[b] - mt);
    }
    const double ct = c > 0 ? exp(tv.lnG4[b] + tv.lnB[c - 1]) : 0.0;
    if (a < b)
        return term1 * (2 * (exp(tv.lnE2[b] + tv.lnB[a]) - exp(tv.lnE2[b] + tv.lnNegG1[a])) - ct - mt);
    if (a == b) return term1 * ((2 * (low + tv.G3[b])) - ct - mt) + tv.norecombs[a];
    return term1 * ((2 * (low + tv.G2[b])) - ct - mt);
}

// calc_state_priors (trans.cpp:785-822) with the iterator's minimum age
double prior_internal(int b, const Lineages &L, const Model &m, int minage)
{
    if (b < minage) return 0.0;
    double sum = 0.0;
    for (int x = 2 * minage; x < 2 * b - 1; x++) sum += m.csteps[x] * L.nb[x / 2] / (2.0 * m.pop[x / 2]);
    double p = exp(-sum) / L.nc[b];
    if (b < m.T - 2) {
        double Z = 0.0;
        if (b > minage) Z = m.csteps[2 * b - 1] * L.nb[b - 1] / (2.0 * m.pop[b - 1]);
        p *= 1.0 - exp(-m.csteps[2 * b] * L.nb[b] / (2.0 * m.pop[b]) - Z);
    }
    return p;
}

int root_child_of(const LTree &t, int node)
{
    int ptr = node, prev = -1;
    while (ptr != t.root) {
        prev = ptr;
        ptr = t.nodes[ptr].parent;
    }
    return prev;
}

// calc_recomb with internal (trans.cpp:316-409)
double recomb_internal(const LTree &last, const Model &m, const Lineages &L, const Spr &spr,
                       int node1, int a, double last_len)
{
    const LNode *ln = last.nodes.data();
    const int sub = sub_root(last), mr = main_root(last);
    const int root_age = ln[mr].age;
    const int k = spr.rt;
    if (spr.cn == sub) {
        if (a < spr.ct) return 0.0;
        if (spr.rn == mr && node1 != mr) return 0.0;
    }
    if (root_child_of(last, spr.cn) == sub && root_child_of(last, spr.rn) == mr && a == spr.rt) {
        // the state must not lie topologically below the recombination
        for (int p = ln[node1].parent; p != last.root; p = ln[p].parent)
            if (p == spr.rn) return 0.0;
    }
    last_len += m.times[a] - m.times[ln[sub].age];
    double len_b;
    if (a > root_age) {
        last_len += m.times[a] - m.times[root_age];
        len_b = last_len + m.steps[a];
    } else {
        len_b = last_len + m.steps[root_age];
    }
    const int nbk = L.nb[k] + int(k < a);
    const int nrk = L.nr[k] + int(k <= a) + int(k == a) - int(k >= std::max(root_age, a));
    return nbk * m.steps[k] / (nrk * len_b) * (1.0 - exp(-fmax(m.rho * last_len, m.rho)));
}

// calc_recoal with internal (trans.cpp:437-494)
double recoal_internal(const LTree &last, const Model &m, const Lineages &L, const Spr &spr, int a,
                       int rpa)
{
    const int k = spr.rt, j = spr.ct;
    int nbj = L.nb[j] - int(j < rpa) + int(j < a);
    int ncj = L.nc[j] - int(j <= rpa) - int(j == rpa) + int(j <= a) + int(j == a);
    bool over = false;
    if (spr.rn == main_root(last) && spr.ct >= last.nodes[sub_root(last)].age) {
        over = true;
        nbj = 1;
        ncj++;
    }
    double p = 1.0 / ncj;
    if (j < m.T - 2) {
        double Z = 0.0;
        if (j > k) {
            int b1 = L.nb[j - 1] - int(j - 1 < rpa) + int(j - 1 < a);
            if (over) b1 = 1;
            Z = m.csteps[2 * j - 1] * b1 / (2.0 * m.pop[j - 1]);
        }
        p *= 1.0 - exp(-m.csteps[2 * j] * nbj / (2.0 * m.pop[j]) - Z);
    }
    return p;
}

// calc_recomb_recoal with internal (trans.cpp:497-528)
double recomb_recoal_internal(const LTree &last, const Model &m, const Lineages &L, const Spr &spr,
                              int node1, int a, int rpa, double last_len)
{
    const int k = spr.rt, j = spr.ct;
    double p = recomb_internal(last, m, L, spr, node1, a, last_len);
    double sum = 0.0;
    for (int x = 2 * k; x < 2 * j - 1; x++) {
        const int nbm = L.nb[x / 2] - int(x / 2 < rpa) + int(x / 2 < a);
        sum += m.csteps[x] * nbm / (2.0 * m.pop[x / 2]);
    }
    p *= exp(-sum);
    p *= recoal_internal(last, m, L, spr, a, rpa);
    return p;
}

// get_deterministic_transitions with internal (trans.cpp:156-275)
int determ_internal(const LTree &last, const LTree &t, const BlockStates &s2, const Spr &spr,
                    const int *mapping, int node1, int time1)
{
    const LNode *ln = last.nodes.data(), *nd = t.nodes.data();
    if ((node1 == spr.cn && time1 == spr.ct) || (node1 == spr.rn && time1 == spr.rt)) return -1;
    if (node1 != spr.rn) {
        int node2;
        bool disrupt = false;
        if (ln[node1].child[0] == -1) {
            node2 = node1;
        } else if (spr.rn == ln[node1].child[0]) {
            node2 = mapping[ln[node1].child[1]];
            disrupt = true;
        } else if (spr.rn == ln[node1].child[1]) {
            node2 = mapping[ln[node1].child[0]];
            disrupt = true;
        } else {
            node2 = mapping[node1];
        }
        if ((spr.cn == node1 && spr.ct < time1) || (mapping[spr.cn] == node2 && spr.ct < time1) ||
            (disrupt && mapping[spr.cn] == node2 && spr.ct <= time1))
            node2 = nd[node2].parent;
        if (nd[node2].age > time1) return -1;   // a dead end
        const int p = nd[node2].parent;
        if (p != -1 && time1 > nd[p].age) return -1;
        return s2.lookup(node2, time1);
    }
    if (spr.rt > time1) return s2.lookup(mapping[spr.rn], time1);
    const int parent = ln[spr.rn].parent;
    const int time2 = ln[parent].age;
    const int *c = ln[parent].child;
    const int other = (c[1] == spr.rn ? c[0] : c[1]);
    const int node2 = (other == spr.cn) ? nd[mapping[other]].parent : mapping[other];
    return s2.lookup(node2, time2);
}

// the pieces of the emissions of one tree: per-node mutation probabilities,
// the postorder, the main-tree and subtree lengths (emit.cpp:97-116, 745-775)
struct EmitTree {
    FlatTree ft;
    std::vector<double> muts, nomuts;
    std::vector<int> order, pre;
    double mainlen = 0.0, sublen = 0.0;
    EmitTree(const LTree &t, const Model &m) : ft(t)
    {
        const int N = t.nnodes;
        const double mintime = m.times[1] * .1;
        muts.assign(N, 0.0);
        nomuts.assign(N, 0.0);
        for (int i = 0; i < N; i++) {
            if (i == t.root) continue;
            const int page = t.nodes[t.nodes[i].parent].age;
            if (page == m.T + 1) continue;   // the virtual root's branches
            const double d = fmax(m.times[page] - m.times[t.nodes[i].age], mintime);
            muts[i] = prob_branch(d, m.mu, true);
            nomuts[i] = prob_branch(d, m.mu, false);
        }
        order.resize(N);
        postorder(ft.view, order.data());
        auto branch_sum = [&](int top) {
            double s = 0.0;
            std::vector<int> st{top};
            while (!st.empty()) {
                const int x = st.back();
                st.pop_back();
                if (x != top) {
                    const int p = t.nodes[x].parent;
                    s += fmax(m.times[t.nodes[p].age] - m.times[t.nodes[x].age], mintime);
                }
                if (t.nodes[x].child[0] != -1) {
                    st.push_back(t.nodes[x].child[0]);
                    st.push_back(t.nodes[x].child[1]);
                }
            }
            return s;
        };
        mainlen = branch_sum(main_root(t));
        sublen = branch_sum(sub_root(t));
    }
};

// inner partials of every node (postorder) and outer partials of the main
// tree from its root down (likelihood_site_inner/outer, emit.cpp:151-296)
void partials(const LTree &t, const EmitTree &et, const char *bases, std::vector<double> &inner,
              std::vector<double> &outer)
{
    const int N = t.nnodes;
    const LNode *nd = t.nodes.data();
    const double *mu = et.muts.data(), *nm = et.nomuts.data();
    for (int oi = 0; oi < N; oi++) {
        const int j = et.order[oi];
        if (nd[j].child[0] == -1) {
            leaf_partial(bases[j], &inner[4 * j]);
            continue;
        }
        const int c1 = nd[j].child[0], c2 = nd[j].child[1];
        for (int a = 0; a < 4; a++) {
            double p1 = 0.0, p2 = 0.0;
            for (int b = 0; b < 4; b++) {
                p1 += inner[4 * c1 + b] * (a == b ? nm[c1] : mu[c1]);
                p2 += inner[4 * c2 + b] * (a == b ? nm[c2] : mu[c2]);
            }
            inner[4 * j + a] = p1 * p2;
        }
    }
    const int mr = main_root(t);
    std::vector<int> st{mr};
    while (!st.empty()) {
        const int j = st.back();
        st.pop_back();
        if (j == mr) {
            for (int a = 0; a < 4; a++) outer[4 * j + a] = 1.0;
        } else {
            const int p = nd[j].parent;
            const int sib = nd[p].child[0] == j ? nd[p].child[1] : nd[p].child[0];
            for (int a = 0; a < 4; a++) {
                double p1 = 0.0, p2 = 0.0;
                for (int b = 0; b < 4; b++) {
                    p1 += inner[4 * sib + b] * (a == b ? nm[sib] : mu[sib]);
                    if (p != mr) p2 += outer[4 * p + b] * (a == b ? nm[p] : mu[p]);
                }
                outer[4 * j + a] = (p != mr) ? p1 * p2 : p1;
            }
        }
        if (nd[j].child[0] != -1) {
            st.push_back(nd[j].child[0]);
            st.push_back(nd[j].child[1]);
        }
    }
}

// The switch matrix from tree `last` to tree `t` in compressed form
// (calc_transition_probs_switch with internal, trans.cpp:538-739): rows of
// last's states (determ/determprob), the two dense rows over t's states.  A
// tree without states counts as one placeholder state.
int switch_internal(const LTree &last, const LTree &t, const BlockStates &s1,
                    const BlockStates &s, const Spr &spr, const int *mapping, const Model &m,
                    SwitchOut &o)
{
    const int T = m.T, S1 = s1.nstates(), S = s.nstates();
    o.determ.assign(std::max(S1, 1), -1);
    o.determprob.assign(std::max(S1, 1), 0.0);
    o.recombrow.assign(std::max(S, 1), 0.0);
    o.recoalrow.assign(std::max(S, 1), 0.0);
    o.recombsrc = o.recoalsrc = -1;
    if (S1 == 0) {
        if (S == 0) {
            o.determ[0] = 0;
        } else {
            const int k = s.lookup(main_root(t), spr.ct);
            if (k < 0) return 1;
            o.determ[0] = k;
        }
        o.determprob[0] = 1.0;
        return 0;
    }
    Lineages PL;
    lineages_internal(last, T, PL);
    const double last_len = treelen_internal(last, m.times);
    const int rpa0 = last.nodes[last.nodes[spr.rn].parent].age;
    if (S == 0) {
        for (int i = 0; i < S1; i++) {
            const int n1 = s1.node[i], t1 = s1.time[i];
            const int rpa = (n1 == spr.rn && t1 > spr.rt) ? t1 : rpa0;
            o.determ[i] = 0;
            o.determprob[i] = recomb_recoal_internal(last, m, PL, spr, n1, t1, rpa, last_len);
        }
        return 0;
    }
    // calc_recoal_sums (trans.cpp:417-440)
    double sums = 0.0, sums2[2 * MAXT + 1] = {0.0};
    for (int x = 2 * spr.rt; x < 2 * spr.ct - 1; x++) {
        const int nbm = PL.nb[x / 2] - int(x / 2 < rpa0);
        sums += m.csteps[x] * nbm / (2.0 * m.pop[x / 2]);
    }
    double acc = 0.0;
    sums2[2 * spr.rt] = 0.0;
    for (int x = 2 * spr.rt; x < 2 * spr.ct - 1; x++) {
        acc += m.csteps[x] / (2.0 * m.pop[x / 2]);
        sums2[x + 1] = acc;
    }
    double recoals[MAXT];
    for (int a = 0; a < T; a++) recoals[a] = recoal_internal(last, m, PL, spr, a, rpa0);
    for (int i = 0; i < S1; i++) {
        const int n1 = s1.node[i], t1 = s1.time[i];
        const int tgt = determ_internal(last, t, s, spr, mapping, n1, t1);
        o.determ[i] = tgt;
        if (tgt < 0) continue;
        if (n1 == spr.rn && t1 > spr.rt) {
            o.determprob[i] = recomb_recoal_internal(last, m, PL, spr, n1, t1, t1, last_len);
        } else {
            const int idx = std::max(std::min(2 * spr.ct - 1, 2 * t1), 2 * spr.rt);
            o.determprob[i] = recomb_internal(last, m, PL, spr, n1, t1, last_len) *
                              exp(-sums - sums2[idx]) * recoals[t1];
        }
    }
    int rs = -1, cs = -1;
    for (int i = 0; i < S1; i++) {
        if (s1.node[i] == spr.rn && s1.time[i] == spr.rt) rs = i;
        else if (s1.node[i] == spr.cn && s1.time[i] == spr.ct) cs = i;
    }
    o.recombsrc = rs;
    o.recoalsrc = cs;
    if (rs != -1) {
        // get_recomb_transition_switch (trans.cpp:278-313)
        const int parent = last.nodes[spr.rn].parent;
        const int time2 = last.nodes[parent].age;
        const int *c = last.nodes[parent].child;
        const int other = (c[1] == spr.rn ? c[0] : c[1]);
        const int node2 = (other == spr.cn) ? t.nodes[mapping[other]].parent : mapping[other];
        const int stay = s.lookup(mapping[spr.rn], spr.rt);
        const int esc = s.lookup(node2, time2);
        if (stay != -1)
            o.recombrow[stay] =
                recomb_recoal_internal(last, m, PL, spr, s1.node[rs], s1.time[rs], rpa0, last_len);
        if (esc != -1)
            o.recombrow[esc] = recomb_recoal_internal(last, m, PL, spr, s1.node[rs], s1.time[rs],
                                                      s1.time[rs], last_len);
    }
    if (cs != -1) {
        const int node1 = s1.node[cs], time1 = s1.time[cs];
        const int lp = last.nodes[spr.rn].parent;
        int node3;
        if (lp == node1) {
            const int *c = last.nodes[lp].child;
            node3 = mapping[c[1] == spr.rn ? c[0] : c[1]];
        } else {
            node3 = mapping[node1];
        }
        const int parent = t.nodes[mapping[spr.rn]].parent;
        for (int k = 0; k < S; k++) {
            const int n2 = s.node[k], t2 = s.time[k];
            if (!((n2 == mapping[spr.rn] && t2 >= spr.rt) || (n2 == node3 && t2 == time1) ||
                  (n2 == parent && t2 == time1)))
                continue;
            Spr s2 = spr;
            s2.ct = t2;
            o.recoalrow[k] = recomb_recoal_internal(last, m, PL, s2, node1, time1, rpa0, last_len);
        }
    }
    return 0;
}

// The emissions of one tree in subtree/maintree form for its states
// (calc_emissions with internal, emit.cpp:650-869): per-state branch terms
// and invariant-column likelihoods, then one variant column at a time.
struct EmitInternal {
    const LTree &t;
    const BlockStates &s;
    EmitTree et;
    int sub, mr;
    std::vector<double> inv, mut, nomut;   // [S], [3S], [3S]
    std::vector<double> inner, outer;
    EmitInternal(const LTree &t_, const Model &m, const BlockStates &s_)
        : t(t_), s(s_), et(t_, m), sub(sub_root(t_)), mr(main_root(t_))
    {
        const int S = s.nstates(), T = m.T;
        const double mintime = m.times[1] * .1;
        const double time1 = m.times[t.nodes[sub].age];
        inv.resize(S);
        mut.resize(3 * S);
        nomut.resize(3 * S);
        for (int k = 0; k < S; k++) {
            const int n = s.node[k], par = t.nodes[n].parent;
            const double time2 = m.times[t.nodes[n].age];
            const double pt = (par != -1) ? m.times[std::min(t.nodes[par].age, T - 1)] : 0.0;
            const double ctm = m.times[s.time[k]];
            const double dist[3] = {fmax(ctm - time1, mintime), fmax(ctm - time2, mintime),
                                    fmax(pt - ctm, mintime)};
            for (int x = 0; x < 3; x++) {
                mut[3 * k + x] = prob_branch(dist[x], m.mu, true);
                nomut[3 * k + x] = prob_branch(dist[x], m.mu, false);
            }
            double len = et.mainlen + et.sublen + fmax(ctm - time1, mintime);
            if (n == mr) len += fmax(ctm - m.times[t.nodes[mr].age], mintime);
            inv[k] = .25 * exp(-m.mu * fmax(len, mintime));
        }
        inner.assign(4 * t.nnodes, 0.0);
        outer.assign(4 * t.nnodes, 0.0);
    }
    // the states' emissions of a variant column (bases[i] of leaf i)
    void variant(const char *bases, double *row)
    {
        partials(t, et, bases, inner, outer);
        const double *in2 = &inner[4 * sub];
        for (int k = 0; k < s.nstates(); k++) {
            const int n = s.node[k];
            row[k] = emit_variant(in2, &inner[4 * n], &outer[4 * n], n == mr, &mut[3 * k],
                                  &nomut[3 * k]);
        }
    }
};

}  // namespace

int prepare_internal(const ArgTrees &tr, const Model &m, char **seqs, int minage, HostPrep &hp)
{
    const int T = m.T, T1 = T - 1, B = (int)tr.blocks.size(), N = tr.nnodes;
    const int nleaves = tr.num_leaves();
    const int L = tr.length();
    std::vector<BlockStates> bs(B);
    hp.S.assign(B, 1);
    hp.Sreal.assign(B, 0);
    int Smax = 1;
    for (int b = 0; b < B; b++) {
        if (partial(tr.blocks[b].tree, T)) make_states(tr.blocks[b].tree, T, minage, bs[b]);
        else make_states(tr.blocks[b].tree, T, 1 << 20, bs[b]);   // no states
        hp.Sreal[b] = bs[b].nstates();
        hp.S[b] = std::max(1, hp.Sreal[b]);
        Smax = std::max(Smax, hp.S[b]);
    }
    hp.R = -1;
    for (int R : {1, 2, 4, 6, 8, 12, 16, 32})
        if (64 * R >= Smax) {
            hp.R = R;
            break;
        }
    if (hp.R < 0) {
        fprintf(stderr, "argweaver_amd: %d internal states in a block (limit 2048)\n", Smax);
        return 1;
    }
    const int R = hp.R, RW = 64 * R;
    const int TC = tc_base(T1, T1);
    hp.planes.assign((size_t)B * NCOEF * RW, 0.0);
    hp.info.assign((size_t)B * RW, INFO_FIRST | INFO_LAST | 31);
    hp.det.assign((size_t)B * RW, -1);
    hp.tm.assign((size_t)B * T1 * T1, 0.0);
    hp.tm2s.assign((size_t)B * TC, 0.0);
    hp.srcs.assign(2 * (size_t)B, -1);
    hp.prior.assign(RW, 0.0);
    hp.cls.assign(L, -1);
    hp.vemit.clear();
    hp.nvar = 0;
    // the sequences of the leaves in leaf order
    std::vector<const char *> rows(nleaves);
    for (int i = 0; i < nleaves; i++) rows[i] = seqs[tr.seqids[i]] + tr.start_coord;

    std::vector<char> bases(N, 'N');
    int end = 0;
    for (int b = 0; b < B; b++) {
        const LBlock &BL = tr.blocks[b];
        const LTree &t = BL.tree;
        const int start = end;
        end += BL.blocklen;
        const BlockStates &s = bs[b];
        const int S = s.nstates();
        double *pl = &hp.planes[(size_t)b * NCOEF * RW];
        int *info = &hp.info[(size_t)b * RW];
        double *tmb = &hp.tm[(size_t)b * T1 * T1];
        double *tm2 = &hp.tm2s[(size_t)b * TC];
        Lineages Lc{};
        TransVec tv{};
        if (S == 0) {
            // the placeholder state: value carried from column to column
            const int p = lane_pos(0, R);
            pl[CJ * RW + p] = 1.0;
            pl[CE * RW + p] = 1.0;
            info[p] = 0 | INFO_FIRST | INFO_LAST | INFO_VALID;
            for (int x = 0; x < TC; x++) tm2[x] = 1.0;
            if (b == 0) hp.prior[p] = 1.0;
        } else {
            lineages_internal(t, T, Lc);
            trans_internal(t, m, Lc, tv);
            const int mina = std::max(minage, t.nodes[sub_root(t)].age);
            for (int a = 0; a < T1; a++)
                for (int bb = 0; bb < T1; bb++) tmb[a * T1 + bb] = get_time_min(tv, a, bb, 0, mina, false);
            for (int c = 0, x = 0; c < T1; c++)
                for (int a = c; a < T1; a++)
                    for (int bb = c; bb < T1; bb++, x++) tm2[x] = get_time_min(tv, a, bb, c, mina, true);
            EmitInternal em(t, m, s);
            for (int k = 0; k < S; k++) {
                const int n = s.node[k], a = s.time[k], c = t.nodes[n].age;
                const int p = lane_pos(k, R);
                const double Bc = c > 0 ? exp(tv.lnB[c - 1]) : 0.0;
                const double low = a > 0 ? exp(tv.lnE2[a] + tv.lnB[a - 1]) : 0.0;
                const double ct = c > 0 ? exp(tv.lnG4[a] + tv.lnB[c - 1]) : 0.0;
                pl[CF * RW + p] = tv.D[a] * (exp(tv.lnB[a]) - exp(tv.lnNegG1[a]) - Bc);
                pl[CD * RW + p] = tv.D[a];
                pl[CG * RW + p] = tv.E[a] * exp(tv.lnE2[a]);
                pl[CK * RW + p] = tv.E[a] * (low + tv.G2[a] - ct);
                const double term1 = tv.D[a] * tv.E[a];
                pl[CJ * RW + p] = (term1 * ((2 * (low + tv.G3[a])) - ct) + tv.norecombs[a]) -
                                  term1 * (low + tv.G3[a]);
                pl[CE * RW + p] = em.inv[k];
                info[p] = a | (k == s.first[n] ? INFO_FIRST : 0) | (a == s.hi[n] ? INFO_LAST : 0) |
                          INFO_VALID | (c << 11) | (n << 16);
                if (b == 0) hp.prior[p] = prior_internal(a, Lc, m, minage);
            }
            // columns: masked / invariant / variant over the leaves' rows
            std::vector<double> row(S);
            for (int col = start; col < end; col++) {
                const char c0 = rows[0][col];
                bool inv = true;
                for (int x = 1; x < nleaves; x++)
                    if (rows[x][col] != c0) {
                        inv = false;
                        break;
                    }
                if (inv) {
                    hp.cls[col] = (c0 == 'N') ? -2 : -1;
                    continue;
                }
                hp.cls[col] = hp.nvar++;
                const size_t off = hp.vemit.size();
                hp.vemit.resize(off + RW, 0.0);
                for (int x = 0; x < nleaves; x++) bases[x] = rows[x][col];
                em.variant(bases.data(), row.data());
                for (int k = 0; k < S; k++) hp.vemit[off + lane_pos(k, R)] = row[k];
            }
        }
        if (S == 0)
            for (int col = start; col < end; col++) hp.cls[col] = -2;   // emission 1

        if (b == 0) continue;
        // the switch from block b-1 (calc_transition_probs_switch, internal)
        SwitchOut sw;
        if (switch_internal(tr.blocks[b - 1].tree, t, bs[b - 1], s, BL.spr, BL.mapping.data(), m,
                            sw))
            return 2;
        double *cdp = &pl[CDP * RW], *crr = &pl[CRR * RW], *ccr = &pl[CCR * RW];
        int *det = &hp.det[(size_t)b * RW];
        for (size_t i = 0; i < sw.determ.size(); i++) {
            det[lane_pos((int)i, R)] = sw.determ[i];
            cdp[lane_pos((int)i, R)] = sw.determprob[i];
        }
        for (int k = 0; k < S; k++) {
            crr[lane_pos(k, R)] = sw.recombrow[k];
            ccr[lane_pos(k, R)] = sw.recoalrow[k];
        }
        hp.srcs[2 * b] = sw.recombsrc;
        hp.srcs[2 * b + 1] = sw.recoalsrc;
    }
    return 0;
}

// recombinations implied by an internal thread (sample_recombinations with
// internal, recomb.cpp:14-235); draws from rand() in the reference's order
void sample_recombinations_internal(const ArgTrees &tr, const Model &m, int minage,
                                    const int *path, std::vector<int> &rpos,
                                    std::vector<int> &rnode, std::vector<int> &rtime)
{
    const int T = m.T;
    BlockStates s;
    std::vector<int> cn, ct;
    std::vector<double> probs;
    int end = tr.start_coord;
    for (size_t bi = 0; bi < tr.blocks.size(); bi++) {
        const LBlock &B = tr.blocks[bi];
        const LTree &t = B.tree;
        int start = end;
        end += B.blocklen;
        if (!partial(t, T)) continue;
        make_states(t, T, minage, s);
        if (s.nstates() == 0) continue;
        Lineages L;
        lineages_internal(t, T, L);
        TransVec tv;
        trans_internal(t, m, L, tv);
        const int sub = sub_root(t), mr = main_root(t);
        const int mina = t.nodes[sub].age;   // TransMatrix::get's minimum age
        start++;   // no new recombination at a block's first column
        int next_recomb = -1;
        for (int i = start; i < end; i++) {
            const int pi = path[i - tr.start_coord], pl = path[i - 1 - tr.start_coord];
            if (pi == pl) {
                if (i > next_recomb) {
                    const int a = s.time[pl];
                    const double self =
                        get_time_min(tv, a, a, t.nodes[s.node[pl]].age, mina, true);
                    const double rate = 1.0 - (tv.norecombs[a] / self);
                    const double ev = -log(rand() / double(RAND_MAX)) / rate;
                    next_recomb = int(fmin(double(end), i + ev));
                }
                if (i < next_recomb) continue;
            }
            next_recomb = -1;
            const int sn = s.node[pi], st = s.time[pi], ln = s.node[pl], ltm = s.time[pl];
            cn.clear();
            ct.clear();
            const int end_time = std::min(st, ltm);
            if (sn == ln)
                for (int k = t.nodes[sn].age; k <= end_time; k++) {
                    cn.push_back(sn);
                    ct.push_back(k);
                }
            for (int k = t.nodes[sub].age; k <= end_time; k++) {
                cn.push_back(sub);
                ct.push_back(k);
            }
            probs.clear();
            const int root_time = std::max(t.nodes[mr].age, ltm);
            for (size_t x = 0; x < cn.size(); x++) {
                const int k = ct[x], j = st, node = cn[x];
                const int rpa = (node == sub || t.nodes[node].parent == -1 || node == ln)
                                    ? ltm : t.nodes[t.nodes[node].parent].age;
                const int nbk = L.nb[k] + int(k < ltm);
                const int nrk = L.nr[k] + int(k <= ltm) + int(k == ltm) - int(k == root_time);
                const double precomb = nbk * m.steps[k] / nrk;
                double sum = 0.0;
                for (int y = k; y < j - 1; y++) {
                    const int nbm = L.nb[y] + int(y < ltm) - int(y < rpa);
                    sum += m.steps[y] * nbm / (2.0 * m.pop[y]);
                }
                double pcoal = 1.0;
                const int nbj = L.nb[j] + int(j < ltm) - int(j < rpa);
                if (k == j) {
                    if (j < T - 2) pcoal = 1.0 - exp(-m.csteps[2 * j] * nbj / (2.0 * m.pop[j]));
                } else {
                    const int y = j - 1;
                    const int nbm = L.nb[y] + int(y < ltm) - int(y < rpa);
                    sum += m.csteps[2 * y] * nbm / (2.0 * m.pop[y]);
                    if (j < T - 2)
                        pcoal = 1.0 - exp(-m.csteps[2 * j - 1] * nbm / (2.0 * m.pop[y]) -
                                          m.csteps[2 * j] * nbj / (2.0 * m.pop[j]));
                }
                const int ncj = L.nc[j] - int(j <= rpa) - int(j == rpa) + int(j <= ltm) + int(j == ltm);
                pcoal /= ncj;
                probs.push_back(precomb * exp(-sum) * pcoal);
            }
            double total = 0.0;
            for (double p : probs) total += p;
            const double pick = rand() / double(RAND_MAX) * total;
            double acc = 0.0;
            int sel = (int)probs.size() - 1;
            for (size_t x = 0; x < probs.size(); x++) {
                acc += probs[x];
                if (acc >= pick) {
                    sel = (int)x;
                    break;
                }
            }
            rpos.push_back(i);
            rnode.push_back(cn[sel]);
            rtime.push_back(ct[sel]);
        }
    }
}

}  // namespace aw

// ---------------------------------------------------------------------------
// The reference's brute-force checks of the internal-branch matrices: every
// entry against the probability it stands for, computed on trees with the
// subtree attached at the states (the test suite's assert_* exports).

namespace aw {
namespace {

Model check_model(const double *times, int ntimes, const double *popsizes, double rho, double mu)
{
    std::vector<double> pops(ntimes, 1.0);
    if (popsizes)
        for (int i = 0; i < ntimes; i++) pops[i] = popsizes[i];
    Model m;
    model_init(m, ntimes, times, pops.data(), rho, mu);
    return m;
}

bool check_grid(const char *who, int ntimes)
{
    if (ntimes >= 3 && ntimes <= MAXT) return true;
    fprintf(stderr, "argweaver_amd: %s: ntimes %d outside 3..%d\n", who, ntimes, MAXT);
    return false;
}

// calc_recomb_prob (trans.cpp:849-853)
double recomb_prob(const LTree &t, const Model &m)
{
    FlatTree ft(t);
    return 1.0 - exp(-fmax(m.rho * tree_length(ft.view, m.times), m.rho));
}

// calc_spr_prob (total_prob.cpp:210-257) on tree t with its own length
double spr_prob_on(const LTree &t, const Model &m, const Spr &spr)
{
    FlatTree ft(t);
    return spr_logprob(ft.view, m, spr, tree_length(ft.view, m.times));
}

}  // namespace
}  // namespace aw

using namespace aw;

extern "C" {

// assert_transmat_internal (trans.cpp:1035-1121) behind
// arghmm_assert_transmat_internal (trans.cpp:1314-1343): a random branch
// lifted above the root, then every entry of the within-block matrix
// against the sum over the SPRs that lead from state i to state j
bool arghmm_assert_transmat_internal(int nnodes, int *ptree, int *ages, int ntimes,
                                     double *times, double *popsizes, double rho)
{
    if (!check_grid("arghmm_assert_transmat_internal", ntimes)) return false;
    const Model m = check_model(times, ntimes, popsizes, rho, 0.0);
    LTree tree;
    tree_from_parents(tree, ptree, ages, nnodes, nnodes + 2);
    const int node = irand(nnodes);
    apply_spr(tree, Spr{node, tree.nodes[node].age, tree.root, ntimes + 1});
    int *c = tree.nodes[tree.root].child;
    if (c[0] != node) std::swap(c[0], c[1]);

    BlockStates s;
    make_states(tree, ntimes, 0, s);
    const int S = s.nstates();
    if (S == 0) return true;
    Lineages Lc{};
    TransVec tv{};
    lineages_internal(tree, ntimes, Lc);
    trans_internal(tree, m, Lc, tv);
    const int sub = sub_root(tree), suba = tree.nodes[sub].age;
    LTree t2 = tree;
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++) {
            const int n1 = s.node[i], a1 = s.time[i], n2 = s.node[j], a2 = s.time[j];
            const int sister_age = t2.nodes[n1].age;
            apply_spr(t2, Spr{sub, suba, n1, a1});
            const double rp = recomb_prob(t2, m);
            double p = -INFINITY;
            // the recombination on the new branch, re-coalescing at state j
            Spr spr{sub, 0, n2, a2};
            for (int r = t2.nodes[sub].age; r <= std::min(a1, a2); r++) {
                spr.rt = r;
                p = logadd(p, spr_prob_on(t2, m, spr));
            }
            if (n1 == n2) {
                // the recombination on the sister branch, re-coalescing on the
                // new branch or on the node above it
                spr.rn = n1;
                spr.cn = (a2 < a1) ? sub : t2.nodes[sub].parent;
                spr.ct = a2;
                for (int r = sister_age; r <= std::min(a1, a2); r++) {
                    spr.rt = r;
                    p = logadd(p, spr_prob_on(t2, m, spr));
                }
            }
            p += log(rp);
            if (i == j) p = logadd(p, log(1.0 - rp));
            const double p2 = log(get_time_min(tv, a1, a2, tree.nodes[n2].age, suba, n1 == n2));
            if (!close_enough(p, p2, 1e-4, 1e-9)) return false;
            apply_spr(t2, Spr{sub, t2.nodes[sub].age, t2.root, ntimes + 1});
        }
    return true;
}

// assert_transmat_switch_internal (trans.cpp:1124-1186) behind
// arghmm_assert_transmat_switch_internal (trans.cpp:1346-1379): a random
// branch removed along a path from the first tree (the caller's ARG is left
// in subtree/maintree form, as the reference leaves it), then every nonzero
// entry of every switch matrix against the probability of the SPR it
// implies once both trees carry the subtree
bool arghmm_assert_transmat_switch_internal(LocalTrees *trees, int ntimes, double *times,
                                            double *popsizes, double rho)
{
    if (!trees || !check_grid("arghmm_assert_transmat_switch_internal", ntimes)) return false;
    const Model m = check_model(times, ntimes, popsizes, rho, 0.0);
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    std::vector<int> path(tr.blocks.size());
    const int node = irand(tr.nnodes);
    removal_path_from_start(tr, node, path.data());
    remove_thread_path(tr, path.data(), ntimes + 1, nullptr);

    const int B = (int)tr.blocks.size();
    std::vector<BlockStates> bs(B);
    for (int b = 0; b < B; b++) make_states(tr.blocks[b].tree, ntimes, 0, bs[b]);
    for (int b = 1; b < B; b++) {
        const LBlock &BL = tr.blocks[b];
        const LTree &last = tr.blocks[b - 1].tree, &t = BL.tree;
        const BlockStates &s1 = bs[b - 1], &s2 = bs[b];
        SwitchOut sw;
        if (switch_internal(last, t, s1, s2, BL.spr, BL.mapping.data(), m, sw)) return false;
        const int last_sub = sub_root(last), sub = sub_root(t);
        for (int i = 0; i < s1.nstates(); i++)
            for (int j = 0; j < s2.nstates(); j++) {
                const double p2 = log(sw.get(i, j));
                if (p2 == -INFINITY) continue;
                LTree last2 = last, t2 = t;
                Spr spr = BL.spr;
                std::vector<int> mapping(BL.mapping);
                apply_spr(last2, Spr{last_sub, last2.nodes[last_sub].age, s1.node[i], s1.time[i]});
                apply_spr(t2, Spr{sub, t2.nodes[sub].age, s2.node[j], s2.time[j]});
                add_spr_branch_internal(t2, last2, s2.node[j], s2.time[j], s1.node[i], spr, mapping,
                                        sub, last_sub);
                const double p = log(recomb_prob(last2, m)) + spr_prob_on(last2, m, spr);
                if (!close_enough(p, p2, 1e-4, 1e-9)) return false;
            }
    }
    return true;
}

// assert_emissions_internal (emit.cpp:1253-1280) behind
// arghmm_assert_emit_internal (emit.cpp:1340-1380): a uniformly drawn
// removal path applied to a copy of the ARG, then per block the emissions
// of the internal states against a full likelihood of the tree with the
// subtree attached at each state (calc_emissions_internal_slow,
// emit.cpp:1171-1214); seqs[i] is leaf i's sequence
bool arghmm_assert_emit_internal(LocalTrees *trees, int ntimes, double *times, double mu,
                                 char **seqs, int nseqs, int seqlen)
{
    (void)seqlen;
    if (!trees || !check_grid("arghmm_assert_emit_internal", ntimes)) return false;
    const Model m = check_model(times, ntimes, nullptr, 0.0, mu);
    ArgTrees tr = *reinterpret_cast<ArgTrees *>(trees);
    std::vector<int> path(tr.blocks.size());
    removal_path_uniform(tr, path.data());
    remove_thread_path(tr, path.data(), ntimes + 1, nullptr);

    std::vector<const char *> rows(nseqs);
    std::vector<char> bases(tr.nnodes, 'N');
    int end = tr.start_coord;
    for (const LBlock &BL : tr.blocks) {
        const int start = end, len = BL.blocklen;
        end = start + len;
        const LTree &t = BL.tree;
        BlockStates s;
        make_states(t, ntimes, 0, s);
        const int S = s.nstates();
        if (S == 0) continue;
        for (int j = 0; j < nseqs; j++) rows[j] = seqs[j] + start;
        std::vector<char> inv(len);
        for (int i = 0; i < len; i++) {
            bool same = true;
            for (int x = 1; x < nseqs; x++)
                if (rows[x][i] != rows[0][i]) {
                    same = false;
                    break;
                }
            inv[i] = same;
        }
        // the fast table (calc_emissions_internal)
        std::vector<double> fast((size_t)len * S);
        EmitInternal em(t, m, s);
        const int nleaves = t.num_leaves();
        for (int i = 0; i < len; i++) {
            double *row = &fast[(size_t)i * S];
            if (inv[i]) {
                for (int k = 0; k < S; k++) row[k] = (rows[0][i] == 'N') ? 1.0 : em.inv[k];
                continue;
            }
            for (int x = 0; x < nleaves; x++) bases[x] = rows[x][i];
            em.variant(bases.data(), row);
        }
        // the slow table: the subtree attached at each state
        std::vector<double> slow((size_t)len * S);
        std::vector<double *> slow_rows(len);
        for (int i = 0; i < len; i++) slow_rows[i] = &slow[(size_t)i * S];
        const int sub = sub_root(t), suba = t.nodes[sub].age;
        LTree t2 = t;
        for (int k = 0; k < S; k++) {
            apply_spr(t2, Spr{sub, suba, s.node[k], s.time[k]});
            emission_column_slow(t2, m, rows.data(), len, inv, k, slow_rows.data());
            apply_spr(t2, Spr{sub, suba, t2.root, ntimes + 1});
        }
        for (size_t x = 0; x < fast.size(); x++)
            if (!close_enough(fast[x], slow[x], .0001, 1e-12)) return false;
    }
    return true;
}

}  // extern "C"
