/*
 * oracle.cpp -- CPU restatement of ARGweaver's threading HMM.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain scalar C++, FP64,
 * single-threaded, following the reference's loops and summation order so
 * that it reproduces the compiled reference to rounding.  Reference
 * citations are relative to the reference repository root.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct Node {
    int parent, child[2], age;
};

struct Tree {
    int nnodes = 0, root = -1;
    std::vector<Node> nodes;
};

// set_ptree (src/argweaver/local_tree.h set_ptree): children are filled in
// node-index order, first child seen goes to child[0].
Tree make_tree(const int *ptree, const int *ages, int nnodes)
{
    Tree t;
    t.nnodes = nnodes;
    t.nodes.resize(nnodes);
    for (int i = 0; i < nnodes; i++) {
        t.nodes[i].parent = ptree[i];
        t.nodes[i].child[0] = t.nodes[i].child[1] = -1;
        t.nodes[i].age = ages[i];
    }
    for (int i = 0; i < nnodes; i++) {
        int p = ptree[i];
        if (p != -1) {
            if (t.nodes[p].child[0] == -1) t.nodes[p].child[0] = i;
            else t.nodes[p].child[1] = i;
        } else {
            t.root = i;
        }
    }
    return t;
}

int sibling(const Tree &t, int node)
{
    int p = t.nodes[node].parent;
    if (p == -1) return -1;
    const int *c = t.nodes[p].child;
    return c[0] == node ? c[1] : c[0];
}

struct Model {
    int T;
    std::vector<double> times, steps, csteps, pop;
    double rho, mu;
};

// setup_time_steps (src/argweaver/model.h:322-333) and get_coal_time_steps
// (src/argweaver/model.cpp:9-21).  The reference leaves times2[2T-2] and
// coal_time_steps[2T-1] uninitialised; neither is read on this path, the
// restatement fills them with times[T-1] / INFINITY.
Model make_model(const orc_instance *in)
{
    Model m;
    m.T = in->ntimes;
    m.times.assign(in->times, in->times + m.T);
    m.pop.assign(in->popsizes, in->popsizes + m.T);
    m.rho = in->rho;
    m.mu = in->mu;
    m.steps.resize(m.T);
    for (int i = 0; i < m.T - 1; i++) m.steps[i] = m.times[i + 1] - m.times[i];
    m.steps[m.T - 1] = INFINITY;
    m.csteps.resize(2 * m.T);
    orc_coal_time_steps(m.times.data(), m.T, m.csteps.data());
    return m;
}

struct State {
    int node, time;
};
typedef std::vector<State> States;

// get_coal_states_external (src/argweaver/states.cpp:53-75)
States coal_states(const Tree &t, int T)
{
    States s;
    for (int i = 0; i < t.nnodes; i++) {
        int p = t.nodes[i].parent;
        int top = (p == -1) ? T - 2 : t.nodes[p].age;
        for (int time = t.nodes[i].age; time <= top; time++) s.push_back({i, time});
    }
    return s;
}

// NodeStateLookup (src/argweaver/states.h:53-131): (node,time) -> index
struct Lookup {
    std::vector<int> first, lo, cnt;
    Lookup(const States &s, int nnodes) : first(nnodes, -1), lo(nnodes, 0), cnt(nnodes, 0)
    {
        for (int k = 0; k < (int)s.size(); k++) {
            int n = s[k].node;
            if (cnt[n] == 0) { first[n] = k; lo[n] = s[k].time; }
            cnt[n]++;
        }
    }
    int operator()(int node, int time) const
    {
        if (node < 0 || cnt[node] == 0) return -1;
        int d = time - lo[node];
        if (d < 0 || d >= cnt[node]) return -1;
        return first[node] + d;
    }
};

struct Lineages {
    std::vector<int> nb, nr, nc;
};

// count_lineages (src/argweaver/local_tree.cpp:34-69)
Lineages count_lineages(const Tree &t, int T)
{
    Lineages L{std::vector<int>(T, 0), std::vector<int>(T, 0), std::vector<int>(T, 0)};
    for (int i = 0; i < t.nnodes; i++) {
        int p = t.nodes[i].parent;
        int page = (p == -1) ? T - 2 : t.nodes[p].age;
        for (int j = t.nodes[i].age; j < page; j++) { L.nb[j]++; L.nr[j]++; L.nc[j]++; }
        L.nr[page]++;
        L.nc[page]++;
        if (p == -1) L.nb[page]++;
    }
    L.nb[T - 1] = 1;
    return L;
}

// get_treelen (src/argweaver/local_tree.cpp:136-155)
double treelen(const Tree &t, const double *times, bool use_basal)
{
    double len = 0.0;
    for (int i = 0; i < t.nnodes; i++) {
        int p = t.nodes[i].parent, age = t.nodes[i].age;
        if (p == -1) {
            if (use_basal) len += times[age + 1] - times[age];
        } else {
            len += times[t.nodes[p].age] - times[age];
        }
    }
    return len;
}

// get_treelen_branch (src/argweaver/local_tree.cpp:179-203), use_basal=false
double treelen_branch(const Tree &t, const double *times, int node, int time, double len)
{
    double blen = times[time];
    double len2 = len + blen;
    if (node == t.root) len2 += blen - times[t.nodes[t.root].age];
    return len2;
}

// get_basal_branch (src/argweaver/local_tree.cpp:206-219)
double basal_branch(const Tree &t, const double *times, int node, int time)
{
    if (node == t.root) return times[time + 1] - times[time];
    int r = t.nodes[t.root].age;
    return times[r + 1] - times[r];
}

// logadd (src/argweaver/common.h:157-164)
double logadd(double a, double b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    return std::max(a, b) + log1p(exp(-fabs(a - b)));
}

struct Trans {
    int T;
    std::vector<double> D, E, lnB, lnE2, lnNegG1, G2, G3, lnG4, norecombs;

    // TransMatrix::get_time (src/argweaver/trans.h:89-128), minage = 0
    double get_time(int a, int b, int c, bool same) const
    {
        double term1 = D[a] * E[b];
        if (!same) {
            if (a < b) return term1 * (exp(lnE2[b] + lnB[a]) - exp(lnE2[b] + lnNegG1[a]));
            double low = (b > 0 ? exp(lnE2[b] + lnB[b - 1]) : 0.0);
            if (a == b) return term1 * (low + G3[b]);
            return term1 * (low + G2[b]);
        }
        double c_term = (c > 0 ? exp(lnG4[b] + lnB[c - 1]) : 0.0);
        if (a < b)
            return term1 * (2 * (exp(lnE2[b] + lnB[a]) - exp(lnE2[b] + lnNegG1[a])) - c_term);
        double low = (b > 0 ? exp(lnE2[b] + lnB[b - 1]) : 0.0);
        if (a == b) return term1 * ((2 * (low + G3[b])) - c_term) + norecombs[a];
        return term1 * ((2 * (low + G2[b])) - c_term);
    }

    // TransMatrix::get (src/argweaver/trans.h:64-83)
    double get(const Tree &t, const States &s, int i, int j) const
    {
        return get_time(s[i].time, s[j].time, t.nodes[s[j].node].age, s[i].node == s[j].node);
    }
};

// calc_transition_probs (src/argweaver/trans.cpp:13-115), external threading
Trans calc_trans(const Tree &t, const Model &m, const Lineages &L)
{
    const int T = m.T;
    Trans tr;
    tr.T = T;
    for (auto *v : {&tr.D, &tr.E, &tr.lnB, &tr.lnE2, &tr.lnNegG1, &tr.G2, &tr.G3, &tr.lnG4,
                    &tr.norecombs})
        v->assign(T, 0.0);

    // coal_rates[-1..2T-1], C[-2..2T-2] (trans.cpp:13-21, 49-54)
    std::vector<double> rates_alloc(2 * T + 1, 0.0), C_alloc(2 * T + 2, 0.0);
    double *rates = &rates_alloc[1];
    double *C = &C_alloc[2];
    rates[-1] = 0.0;
    for (int i = 0; i < 2 * T; i++)
        rates[i] = m.csteps[i] * L.nb[i / 2] / (2.0 * m.pop[i / 2]);
    C[-2] = 0.0;
    C[-1] = 0.0;
    for (int b = 0; b < 2 * T - 1; b++) C[b] = C[b - 1] + rates[b];

    const int root_idx = t.nodes[t.root].age;
    const double root_age = m.times[root_idx];
    const double len = treelen(t, m.times.data(), false);
    const int *nb = L.nb.data(), *nr = L.nr.data(), *nc = L.nc.data();

    for (int b = 0; b < T - 1; b++) {
        double len2 = len + m.times[b];
        double len2_b;
        if (b > root_idx) {
            len2 += m.times[b] - root_age;
            len2_b = len2 + m.steps[b];
        } else {
            len2_b = len2 + m.steps[root_idx];
        }
        double term = C[2 * b - 1] + log(m.steps[b] * (nb[b] + 1.0) / (nr[b] + 1.0));
        tr.lnB[b] = (b == 0) ? term : logadd(tr.lnB[b - 1], term);
        tr.lnE2[b] = -C[2 * b - 2] +
            (b < T - 2 ? log(1 - exp(-rates[2 * b] - rates[2 * b - 1])) : 0.0);
        tr.lnNegG1[b] = C[2 * b - 1] +
            log(-m.steps[b] * ((nb[b] / (nr[b] + 1.0 + int(b < root_idx))) -
                               (nb[b] + 1.0) / (nr[b] + 1.0)));
        double q = (b < T - 2 ? 1.0 - exp(-rates[2 * b]) : 1.0);
        tr.G2[b] = q * m.steps[b] * (nb[b] + 1.0) / (nr[b] + 1.0);
        tr.G3[b] = q * m.steps[b] * (nb[b] / (nr[b] + 1.0 + int(b < root_idx)));
        tr.lnG4[b] = -C[2 * b - 2] +
            (b < T - 2 ? log(1.0 - exp(-rates[2 * b] - rates[2 * b - 1])) : 0.0);
        tr.D[b] = (1.0 - exp(-m.rho * len2)) / len2_b;
        tr.E[b] = 1.0 / nc[b];
        tr.norecombs[b] = exp(-std::max(m.rho * len2, m.rho));
    }
    tr.E[T - 2] = 1.0 / nc[T - 2];
    return tr;
}

struct Spr {
    int rn, rt, cn, ct;
};

// calc_recomb (src/argweaver/trans.cpp:317-413), external branch
double calc_recomb(const Tree &last, const Model &m, const Lineages &L, const Spr &spr,
                   const State &s1, double last_len)
{
    int a = s1.time;
    int k = spr.rt;
    int root_age = last.nodes[last.root].age;
    double len = treelen_branch(last, m.times.data(), s1.node, s1.time, last_len);
    double len_b = len + basal_branch(last, m.times.data(), s1.node, s1.time);
    int nbk = L.nb[k] + int(k < a);
    int nrk = L.nr[k] + int(k <= a) + int(k == a) - int(k >= std::max(root_age, a));
    return nbk * m.steps[k] / (nrk * len_b) * (1.0 - exp(-std::max(m.rho * len, m.rho)));
}

// calc_recoal_sums (src/argweaver/trans.cpp:417-440)
void recoal_sums(const Model &m, const Lineages &L, const Spr &spr, int rpa, double *sums,
                 double *sums2)
{
    int k = spr.rt, j = spr.ct;
    double sum = 0.0;
    for (int x = 2 * k; x < 2 * j - 1; x++) {
        int nbm = L.nb[x / 2] - int(x / 2 < rpa);
        sum += m.csteps[x] * nbm / (2.0 * m.pop[x / 2]);
    }
    *sums = sum;
    sum = 0.0;
    sums2[2 * k] = sum;
    for (int x = 2 * k; x < 2 * j - 1; x++) {
        sum += m.csteps[x] / (2.0 * m.pop[x / 2]);
        sums2[x + 1] = sum;
    }
}

// calc_recoal (src/argweaver/trans.cpp:443-502), external branch
double calc_recoal(const Model &m, const Lineages &L, const Spr &spr, int a, int rpa)
{
    int k = spr.rt, j = spr.ct;
    int nbj = L.nb[j] - int(j < rpa) + int(j < a);
    int ncj = L.nc[j] - int(j <= rpa) - int(j == rpa) + int(j <= a) + int(j == a);
    double p = 1.0 / ncj;
    if (j < m.T - 2) {
        double Z = 0.0;
        if (j > k) {
            int b1 = L.nb[j - 1] - int(j - 1 < rpa) + int(j - 1 < a);
            Z = m.csteps[2 * j - 1] * b1 / (2.0 * m.pop[j - 1]);
        }
        p *= 1.0 - exp(-m.csteps[2 * j] * nbj / (2.0 * m.pop[j]) - Z);
    }
    return p;
}

// calc_recomb_recoal (src/argweaver/trans.cpp:505-534)
double calc_recomb_recoal(const Tree &last, const Model &m, const Lineages &L, const Spr &spr,
                          const State &s1, int rpa, double last_len)
{
    int a = s1.time, k = spr.rt, j = spr.ct;
    double p = calc_recomb(last, m, L, spr, s1, last_len);
    double sum = 0.0;
    for (int x = 2 * k; x < 2 * j - 1; x++) {
        int nbm = L.nb[x / 2] - int(x / 2 < rpa) + int(x / 2 < a);
        sum += m.csteps[x] * nbm / (2.0 * m.pop[x / 2]);
    }
    p *= exp(-sum);
    p *= calc_recoal(m, L, spr, s1.time, rpa);
    return p;
}

struct Switch {
    std::vector<int> determ;
    std::vector<double> determprob, recombrow, recoalrow;
    int recombsrc = -1, recoalsrc = -1;

    // TransMatrixSwitch::get (src/argweaver/trans.h:209-222)
    double get(int i, int j) const
    {
        if (i == recoalsrc) return recoalrow[j];
        if (i == recombsrc) return recombrow[j];
        return determ[i] == j ? determprob[i] : 0.0;
    }
};

// get_deterministic_transitions (src/argweaver/trans.cpp:156-275), external
void deterministic(const Tree &last, const Tree &t, const Spr &spr, const int *mapping,
                   const States &s1, const Lookup &lk2, std::vector<int> &next)
{
    const int n1 = (int)s1.size();
    next.assign(n1, -1);
    for (int i = 0; i < n1; i++) {
        const int node1 = s1[i].node, time1 = s1[i].time;
        if ((node1 == spr.cn && time1 == spr.ct) || (node1 == spr.rn && time1 == spr.rt)) {
            next[i] = -1;
        } else if (node1 != spr.rn) {
            int node2;
            bool disrupt = false;
            const Node &nd = last.nodes[node1];
            if (nd.child[0] == -1) {
                node2 = node1;
            } else if (spr.rn == nd.child[0]) {
                node2 = mapping[nd.child[1]];
                disrupt = true;
            } else if (spr.rn == nd.child[1]) {
                node2 = mapping[nd.child[0]];
                disrupt = true;
            } else {
                node2 = mapping[node1];
            }
            if ((spr.cn == node1 && spr.ct < time1) ||
                (mapping[spr.cn] == node2 && spr.ct < time1) ||
                (disrupt && mapping[spr.cn] == node2 && spr.ct <= time1))
                node2 = t.nodes[node2].parent;
            next[i] = lk2(node2, time1);
        } else {
            if (spr.rt > time1) {
                next[i] = lk2(mapping[spr.rn], time1);
            } else {
                int parent = last.nodes[spr.rn].parent;
                int time2 = last.nodes[parent].age;
                const int *c = last.nodes[parent].child;
                int other = (c[1] == spr.rn ? c[0] : c[1]);
                int node2 = (other == spr.cn ? t.nodes[mapping[other]].parent : mapping[other]);
                next[i] = lk2(node2, time2);
            }
        }
    }
}

// calc_transition_probs_switch (src/argweaver/trans.cpp:538-739), external
Switch calc_switch(const Tree &t, const Tree &last, const Spr &spr, const int *mapping,
                   const States &s1, const States &s2, const Model &m, const Lineages &L)
{
    const int n1 = (int)s1.size(), n2 = (int)s2.size();
    Switch sw;
    sw.determprob.assign(std::max(n1, 1), 0.0);
    sw.recombrow.assign(std::max(n2, 1), 0.0);
    sw.recoalrow.assign(std::max(n2, 1), 0.0);
    const double last_len = treelen(last, m.times.data(), false);
    Lookup lk2(s2, t.nnodes);
    deterministic(last, t, spr, mapping, s1, lk2, sw.determ);

    int rpa = last.nodes[last.nodes[spr.rn].parent].age;
    double sums;
    std::vector<double> sums2(2 * m.T + 1, 0.0);
    recoal_sums(m, L, spr, rpa, &sums, sums2.data());
    std::vector<double> recoals(m.T);
    for (int a = 0; a < m.T; a++) recoals[a] = calc_recoal(m, L, spr, a, rpa);

    for (int i = 0; i < n1; i++) {
        if (sw.determ[i] < 0) continue;
        if (s1[i].node == spr.rn && s1[i].time > spr.rt) {
            sw.determprob[i] = calc_recomb_recoal(last, m, L, spr, s1[i], s1[i].time, last_len);
        } else {
            int idx = std::max(std::min(2 * spr.ct - 1, 2 * s1[i].time), 2 * spr.rt);
            sw.determprob[i] = calc_recomb(last, m, L, spr, s1[i], last_len) *
                               exp(-sums - sums2[idx]) * recoals[s1[i].time];
        }
    }

    for (int i = 0; i < n1; i++) {
        if (s1[i].node == spr.rn && s1[i].time == spr.rt) sw.recombsrc = i;
        else if (s1[i].node == spr.cn && s1[i].time == spr.ct) sw.recoalsrc = i;
    }

    if (sw.recombsrc != -1) {
        // get_recomb_transition_switch (trans.cpp:278-313)
        int parent = last.nodes[spr.rn].parent;
        int time2 = last.nodes[parent].age;
        int other = sibling(last, spr.rn);
        int node2 = (other == spr.cn ? t.nodes[mapping[other]].parent : mapping[other]);
        int stay = lk2(mapping[spr.rn], spr.rt);
        int escape = lk2(node2, time2);
        if (stay != -1)
            sw.recombrow[stay] = calc_recomb_recoal(last, m, L, spr, s1[sw.recombsrc], rpa, last_len);
        if (escape != -1)
            sw.recombrow[escape] = calc_recomb_recoal(last, m, L, spr, s1[sw.recombsrc],
                                                      s1[sw.recombsrc].time, last_len);
    }

    if (sw.recoalsrc == -1) return sw;

    int node1 = s1[sw.recoalsrc].node, time1 = s1[sw.recoalsrc].time;
    int node3;
    int last_parent = last.nodes[spr.rn].parent;
    if (last_parent == node1) {
        const int *c = last.nodes[last_parent].child;
        node3 = mapping[c[1] == spr.rn ? c[0] : c[1]];
    } else {
        node3 = mapping[node1];
    }
    int parent = t.nodes[mapping[spr.rn]].parent;
    for (int j = 0; j < n2; j++) {
        int node2 = s2[j].node, time2 = s2[j].time;
        if (!((node2 == mapping[spr.rn] && time2 >= spr.rt) ||
              (node2 == node3 && time2 == time1) || (node2 == parent && time2 == time1)))
            continue;
        Spr spr2 = spr;
        spr2.ct = time2;
        sw.recoalrow[j] = calc_recomb_recoal(last, m, L, spr2, s1[sw.recoalsrc], rpa, last_len);
    }
    return sw;
}

// calc_state_priors (src/argweaver/trans.cpp:785-831), minage = 0
double prior_time(int b, const Lineages &L, const Model &m)
{
    double sum = 0.0;
    for (int x = 0; x < 2 * b - 1; x++)
        sum += m.csteps[x] * L.nb[x / 2] / (2.0 * m.pop[x / 2]);
    double p = exp(-sum) / L.nc[b];
    if (b < m.T - 2) {
        double Z = 0.0;
        if (b > 0) Z = m.csteps[2 * b - 1] * L.nb[b - 1] / (2.0 * m.pop[b - 1]);
        p *= 1.0 - exp(-m.csteps[2 * b] * L.nb[b] / (2.0 * m.pop[b]) - Z);
    }
    return p;
}

// JC69 prob_branch (src/argweaver/emit.cpp:86-93)
double prob_branch(double t, double mu, bool mut)
{
    const double f = 4. / 3.;
    if (!mut) return .25 * (1.0 + 3. * exp(-f * mu * t));
    return .25 * (1.0 - exp(-f * mu * t));
}

int base_index(char c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    }
    return -1;
}

void leaf_vec(char c, double *v)
{
    if (c == 'N') { v[0] = v[1] = v[2] = v[3] = 1.0; return; }
    v[0] = v[1] = v[2] = v[3] = 0.0;
    v[base_index(c)] = 1.0;
}

// calc_emissions (src/argweaver/emit.cpp:650-869) for external threading:
// seqs[0..nleaves-1] are the tree leaves' rows, seqs[nleaves] the new leaf.
void emissions(const States &s, const Tree &t, const char *const *seqs, int nseqs, int start,
               int blocklen, const Model &m, double *emit)
{
    const int S = (int)s.size();
    const int N = t.nnodes;
    const double mintime = m.times[1] * .1;   // get_mintime (model.h:211-213)
    const int newleaf = (N + 1) / 2;

    // prob_tree_mutation (emit.cpp:97-116)
    std::vector<double> muts(N, 0.0), nomuts(N, 0.0);
    for (int i = 0; i < N; i++) {
        if (i == t.root) continue;
        int page = t.nodes[t.nodes[i].parent].age;
        if (page == m.T + 1) continue;
        double tt = std::max(m.times[page] - m.times[t.nodes[i].age], mintime);
        muts[i] = prob_branch(tt, m.mu, true);
        nomuts[i] = prob_branch(tt, m.mu, false);
    }

    // postorder (local_tree.h get_postorder)
    std::vector<int> order(N);
    {
        std::vector<char> visit(N, 0);
        int i;
        for (i = 0; i < N; i++) {
            if (t.nodes[i].child[0] != -1) break;
            order[i] = i;
        }
        int end = i;
        for (i = 0; i < N; i++) {
            int p = t.nodes[order[i]].parent;
            if (p != -1 && ++visit[p] == 2) order[end++] = p;
        }
    }

    // tree length of the main tree with floored branches (emit.cpp:745-759)
    double mainlen = 0.0;
    for (int i = 0; i < N; i++)
        if (i != t.root) {
            int p = t.nodes[i].parent;
            mainlen += std::max(m.times[t.nodes[p].age] - m.times[t.nodes[i].age], mintime);
        }

    // per-state branch terms (emit.cpp:778-819)
    std::vector<double> mut(3 * S), nomut(3 * S), inv(S);
    for (int j = 0; j < S; j++) {
        int node2 = s[j].node;
        int parent = t.nodes[node2].parent;
        double time2 = m.times[t.nodes[node2].age];
        double parent_time = (parent != -1) ? m.times[std::min(t.nodes[parent].age, m.T - 1)] : 0.0;
        double coal_time = m.times[s[j].time];
        double dist[3];
        dist[0] = std::max(coal_time - 0.0, mintime);
        dist[1] = std::max(coal_time - time2, mintime);
        dist[2] = std::max(parent_time - coal_time, mintime);
        for (int x = 0; x < 3; x++) {
            mut[3 * j + x] = prob_branch(dist[x], m.mu, true);
            nomut[3 * j + x] = prob_branch(dist[x], m.mu, false);
        }
        double len;
        if (node2 == t.root)
            len = mainlen + std::max(coal_time - 0.0, mintime) +
                  std::max(coal_time - m.times[t.nodes[t.root].age], mintime);
        else
            len = mainlen + std::max(coal_time - 0.0, mintime);
        inv[j] = .25 * exp(-m.mu * std::max(len, mintime));
    }

    std::vector<double> inner(4 * N), outer(4 * N);
    std::vector<int> stack(N);
    for (int i = 0; i < blocklen; i++) {
        const int col = start + i;
        // find_invariant_sites / find_masked_sites (emit.cpp:30-58)
        char c0 = seqs[0][col];
        bool invariant = true;
        for (int x = 1; x < nseqs; x++)
            if (seqs[x][col] != c0) { invariant = false; break; }
        double *row = emit + (long long)i * S;
        if (invariant) {
            bool masked = (c0 == 'N');
            for (int j = 0; j < S; j++) row[j] = masked ? 1.0 : inv[j];
            continue;
        }
        // inner table, postorder (emit.cpp:151-196, 252-269)
        for (int oi = 0; oi < N; oi++) {
            int j = order[oi];
            const Node &nd = t.nodes[j];
            if (nd.child[0] == -1) {
                leaf_vec(seqs[j][col], &inner[4 * j]);
            } else {
                int c1 = nd.child[0], c2 = nd.child[1];
                for (int a = 0; a < 4; a++) {
                    double p1 = 0.0, p2 = 0.0;
                    for (int b = 0; b < 4; b++) {
                        if (a == b) {
                            p1 += inner[4 * c1 + b] * nomuts[c1];
                            p2 += inner[4 * c2 + b] * nomuts[c2];
                        } else {
                            p1 += inner[4 * c1 + b] * muts[c1];
                            p2 += inner[4 * c2 + b] * muts[c2];
                        }
                    }
                    inner[4 * j + a] = p1 * p2;
                }
            }
        }
        // outer table, preorder from the root (emit.cpp:200-248, 273-296)
        int top = 0;
        stack[top++] = t.root;
        while (top > 0) {
            int j = stack[--top];
            if (j == t.root) {
                for (int a = 0; a < 4; a++) outer[4 * j + a] = 1.0;
            } else {
                int sib = sibling(t, j);
                int parent = t.nodes[j].parent;
                for (int a = 0; a < 4; a++) {
                    double p1 = 0.0, p2 = 0.0;
                    for (int b = 0; b < 4; b++) {
                        if (a == b) {
                            p1 += inner[4 * sib + b] * nomuts[sib];
                            if (parent != t.root) p2 += outer[4 * parent + b] * nomuts[parent];
                        } else {
                            p1 += inner[4 * sib + b] * muts[sib];
                            if (parent != t.root) p2 += outer[4 * parent + b] * muts[parent];
                        }
                    }
                    outer[4 * j + a] = (parent != t.root) ? p1 * p2 : p1;
                }
            }
            if (t.nodes[j].child[0] != -1) {
                stack[top++] = t.nodes[j].child[0];
                stack[top++] = t.nodes[j].child[1];
            }
        }
        // new leaf inner vector and calc_emit (emit.cpp:686-703, 620-645)
        double in2[4];
        leaf_vec(seqs[newleaf][col], in2);
        for (int j = 0; j < S; j++) {
            int node2 = s[j].node;
            const double *mu3 = &mut[3 * j], *nm3 = &nomut[3 * j];
            double e = 0.0;
            for (int a = 0; a < 4; a++) {
                double p1 = 0.0, p2 = 0.0, p3 = 0.0;
                for (int b = 0; b < 4; b++) {
                    if (a == b) {
                        p1 += in2[b] * nm3[0];
                        p2 += inner[4 * node2 + b] * nm3[1];
                        p3 += outer[4 * node2 + b] * nm3[2];
                    } else {
                        p1 += in2[b] * mu3[0];
                        p2 += inner[4 * node2 + b] * mu3[1];
                        p3 += outer[4 * node2 + b] * mu3[2];
                    }
                }
                if (node2 != t.root) e += p1 * p2 * p3 * .25;
                else e += p1 * p2 * .25;
            }
            row[j] = e;
        }
    }
}

// sample (src/argweaver/common.h:272-290) with pick = u * total
int sample_u(const double *w, int n, double u)
{
    double total = 0.0;
    for (int i = 0; i < n; i++) total += w[i];
    double pick = u * total;
    double x = 0.0;
    for (int i = 0; i < n; i++) {
        x += w[i];
        if (x >= pick) return i;
    }
    return n - 1;
}

double frand() { return rand() / double(RAND_MAX); }

// One block of the instance with everything the HMM needs.
struct Block {
    Tree tree;
    States states;
    Lineages lin;
    Trans trans;
    Spr spr;
    std::vector<int> mapping;
    int start, len;
};

std::vector<Block> blocks_of(const orc_instance *in, const Model &m)
{
    std::vector<Block> bl(in->ntrees);
    int pos = in->start_coord;
    const int N = in->nnodes;
    for (int b = 0; b < in->ntrees; b++) {
        Block &B = bl[b];
        B.tree = make_tree(in->ptrees + (long long)b * N, in->ages + (long long)b * N, N);
        B.states = coal_states(B.tree, m.T);
        B.lin = count_lineages(B.tree, m.T);
        B.trans = calc_trans(B.tree, m, B.lin);
        const int *sp = in->sprs + 4 * b;
        B.spr = {sp[0], sp[1], sp[2], sp[3]};
        B.mapping.assign(N, 0);
        if (b > 0) {
            if (in->mappings) {
                std::copy(in->mappings + (long long)b * N, in->mappings + (long long)(b + 1) * N,
                          B.mapping.begin());
            } else {
                // make_node_mapping (src/argweaver/local_tree.h:767-776)
                for (int j = 0; j < N; j++) B.mapping[j] = j;
                B.mapping[in->ptrees[(long long)(b - 1) * N + sp[0]]] = -1;
            }
        }
        B.start = pos;
        B.len = in->blocklens[b];
        pos += B.len;
    }
    return bl;
}

Switch block_switch(const std::vector<Block> &bl, int b, const Model &m)
{
    const Block &B = bl[b], &P = bl[b - 1];
    return calc_switch(B.tree, P.tree, B.spr, B.mapping.data(), P.states, B.states, m, P.lin);
}

std::vector<const char *> block_seqs(const orc_instance *in)
{
    const int nleaves = (in->nnodes + 1) / 2;
    std::vector<const char *> rows(nleaves + 1);
    for (int i = 0; i < nleaves; i++) rows[i] = in->seqs[in->seqids[i]];
    rows[nleaves] = in->seqs[in->new_chrom];
    return rows;
}

}  // namespace

extern "C" {

void orc_time_points(int ntimes, double maxtime, double delta, double *times)
{
    for (int i = 0; i < ntimes; i++)
        times[i] = (exp(i / double(ntimes - 1) * log(1.0 + delta * maxtime)) - 1) / delta;
}

void orc_coal_time_steps(const double *times, int ntimes, double *cs)
{
    std::vector<double> t2(2 * ntimes + 1, 0.0);
    for (int i = 0; i < ntimes - 1; i++) {
        t2[2 * i] = times[i];
        t2[2 * i + 1] = sqrt((times[i + 1] + 1.0) * (times[i] + 1.0));
    }
    t2[2 * ntimes - 2] = times[ntimes - 1];
    t2[2 * ntimes] = times[ntimes - 1];
    for (int i = 0; i < 2 * ntimes - 2; i++) cs[i] = t2[std::min(i + 1, 2 * ntimes)] - t2[i];
    cs[2 * ntimes - 2] = INFINITY;
    cs[2 * ntimes - 1] = INFINITY;
}

void orc_nstates(const orc_instance *in, int *nstates)
{
    Model m = make_model(in);
    auto bl = blocks_of(in, m);
    for (int b = 0; b < in->ntrees; b++) nstates[b] = (int)bl[b].states.size();
}

int orc_states(const orc_instance *in, int tree, int *out)
{
    const int N = in->nnodes;
    Tree t = make_tree(in->ptrees + (long long)tree * N, in->ages + (long long)tree * N, N);
    States s = coal_states(t, in->ntimes);
    for (size_t k = 0; k < s.size(); k++) {
        out[2 * k] = s[k].node;
        out[2 * k + 1] = s[k].time;
    }
    return (int)s.size();
}

void orc_lineages(const orc_instance *in, int tree, int *nb, int *nr, int *nc)
{
    const int N = in->nnodes;
    Tree t = make_tree(in->ptrees + (long long)tree * N, in->ages + (long long)tree * N, N);
    Lineages L = count_lineages(t, in->ntimes);
    std::copy(L.nb.begin(), L.nb.end(), nb);
    std::copy(L.nr.begin(), L.nr.end(), nr);
    std::copy(L.nc.begin(), L.nc.end(), nc);
}

double orc_treelen(const orc_instance *in, int tree, int use_basal)
{
    const int N = in->nnodes;
    Tree t = make_tree(in->ptrees + (long long)tree * N, in->ages + (long long)tree * N, N);
    return treelen(t, in->times, use_basal != 0);
}

void orc_trans_vectors(const orc_instance *in, int tree, double *out)
{
    Model m = make_model(in);
    const int N = in->nnodes;
    Tree t = make_tree(in->ptrees + (long long)tree * N, in->ages + (long long)tree * N, N);
    Trans tr = calc_trans(t, m, count_lineages(t, m.T));
    const std::vector<double> *v[9] = {&tr.D, &tr.E, &tr.lnB, &tr.lnE2, &tr.lnNegG1,
                                       &tr.G2, &tr.G3, &tr.lnG4, &tr.norecombs};
    for (int x = 0; x < 9; x++) std::copy(v[x]->begin(), v[x]->end(), out + x * m.T);
}

void orc_trans_dense(const orc_instance *in, int tree, double *out)
{
    Model m = make_model(in);
    const int N = in->nnodes;
    Tree t = make_tree(in->ptrees + (long long)tree * N, in->ages + (long long)tree * N, N);
    Trans tr = calc_trans(t, m, count_lineages(t, m.T));
    States s = coal_states(t, m.T);
    const int S = (int)s.size();
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++) out[(long long)i * S + j] = tr.get(t, s, i, j);
}

void orc_switch(const orc_instance *in, int tree, int *determ, double *determprob,
                double *recombrow, double *recoalrow, int *srcs)
{
    Model m = make_model(in);
    auto bl = blocks_of(in, m);
    Switch sw = block_switch(bl, tree, m);
    const int n1 = (int)bl[tree - 1].states.size(), n2 = (int)bl[tree].states.size();
    std::copy(sw.determ.begin(), sw.determ.begin() + n1, determ);
    std::copy(sw.determprob.begin(), sw.determprob.begin() + n1, determprob);
    std::copy(sw.recombrow.begin(), sw.recombrow.begin() + n2, recombrow);
    std::copy(sw.recoalrow.begin(), sw.recoalrow.begin() + n2, recoalrow);
    srcs[0] = sw.recombsrc;
    srcs[1] = sw.recoalsrc;
}

void orc_prior(const orc_instance *in, double *prior)
{
    Model m = make_model(in);
    const int N = in->nnodes;
    Tree t = make_tree(in->ptrees, in->ages, N);
    Lineages L = count_lineages(t, m.T);
    States s = coal_states(t, m.T);
    for (size_t k = 0; k < s.size(); k++) prior[k] = prior_time(s[k].time, L, m);
}

void orc_emissions(const orc_instance *in, int tree, double *emit)
{
    Model m = make_model(in);
    auto bl = blocks_of(in, m);
    auto rows = block_seqs(in);
    const Block &B = bl[tree];
    emissions(B.states, B.tree, rows.data(), (int)rows.size(), B.start - in->start_coord, B.len,
              m, emit);
}

long long orc_forward_size(const orc_instance *in)
{
    Model m = make_model(in);
    auto bl = blocks_of(in, m);
    long long n = 0;
    for (auto &B : bl) n += (long long)B.len * std::max((int)B.states.size(), 1);
    return n;
}

// arghmm_forward_alg (src/argweaver/sample_thread.cpp:394-460)
void orc_forward(const orc_instance *in, int slow, double *fw, double *lnl)
{
    Model m = make_model(in);
    auto bl = blocks_of(in, m);
    auto rows = block_seqs(in);
    const int T = m.T;
    double logl = 0.0;
    long long off = 0;
    const double *prev = nullptr;   // last column of the previous block
    for (int b = 0; b < in->ntrees; b++) {
        const Block &B = bl[b];
        const int S = (int)B.states.size();
        double *blk = fw + off;
        std::vector<double> emit((size_t)B.len * S);
        emissions(B.states, B.tree, rows.data(), (int)rows.size(), B.start - in->start_coord,
                  B.len, m, emit.data());

        if (b == 0) {
            // calc_state_priors (trans.cpp:815-831): column 0 is the prior,
            // with no emission factor and no normalisation.
            double tot = 0.0;
            for (int k = 0; k < S; k++) {
                blk[k] = prior_time(B.states[k].time, B.lin, m);
                tot += blk[k];
            }
            logl += log(tot);
            for (int k = 0; k < S; k++) blk[k] /= 1.0;   // unnormalised, as in the reference
            logl -= log(tot);   // mass is carried into column 1's normaliser
        } else {
            // arghmm_forward_switch (sample_thread.cpp:345-389)
            Switch sw = block_switch(bl, b, m);
            const int S1 = (int)bl[b - 1].states.size();
            for (int k = 0; k < S; k++) blk[k] = 0.0;
            for (int j = 0; j < S1; j++) {
                int k = sw.determ[j];
                if (j != sw.recombsrc && j != sw.recoalsrc && k != -1)
                    blk[k] += prev[j] * sw.determprob[j];
            }
            double norm = 0.0;
            for (int k = 0; k < S; k++) {
                if (sw.recombsrc != -1 && sw.recombrow[k] > 0.0)
                    blk[k] += prev[sw.recombsrc] * sw.recombrow[k];
                if (sw.recoalsrc != -1 && sw.recoalrow[k] > 0.0)
                    blk[k] += prev[sw.recoalsrc] * sw.recoalrow[k];
                blk[k] *= emit[k];
                norm += blk[k];
            }
            for (int k = 0; k < S; k++) blk[k] /= norm;
            logl += log(norm);
        }

        if (slow) {
            // arghmm_forward_block_slow (sample_thread.cpp:303-338)
            std::vector<double> tm((size_t)S * S);
            for (int k = 0; k < S; k++)
                for (int j = 0; j < S; j++) tm[(size_t)j * S + k] = B.trans.get(B.tree, B.states, j, k);
            for (int i = 1; i < B.len; i++) {
                const double *c1 = blk + (long long)(i - 1) * S;
                double *c2 = blk + (long long)i * S;
                double norm = 0.0;
                for (int k = 0; k < S; k++) {
                    double sum = 0.0;
                    for (int j = 0; j < S; j++) sum += c1[j] * tm[(size_t)j * S + k];
                    c2[k] = sum * emit[(size_t)i * S + k];
                    norm += c2[k];
                }
                for (int k = 0; k < S; k++) c2[k] /= norm;
                logl += log(norm);
            }
        } else {
            // arghmm_forward_block (sample_thread.cpp:186-296)
            std::vector<double> tmat((size_t)T * T, 0.0), tmat2((size_t)T * S, 0.0);
            for (int a = 0; a < T - 1; a++) {
                for (int bb = 0; bb < T - 1; bb++) tmat[a * T + bb] = B.trans.get_time(a, bb, 0, false);
                for (int k = 0; k < S; k++) {
                    int bb = B.states[k].time;
                    int c = B.tree.nodes[B.states[k].node].age;
                    tmat2[(size_t)a * S + k] =
                        B.trans.get_time(a, bb, c, true) - B.trans.get_time(a, bb, 0, false);
                }
            }
            int maxtime = 0;
            for (int k = 0; k < S; k++) maxtime = std::max(maxtime, B.states[k].time);
            Lookup lk(B.states, B.tree.nnodes);
            const int N = B.tree.nnodes;
            std::vector<int> ages1(N), ages2(N), idx(N);
            for (int n = 0; n < N; n++) {
                ages1[n] = B.tree.nodes[n].age;
                idx[n] = lk(n, ages1[n]);
                ages2[n] = (n == B.tree.root) ? maxtime : B.tree.nodes[B.tree.nodes[n].parent].age;
            }
            std::vector<double> fg(T), tf(T);
            for (int i = 1; i < B.len; i++) {
                const double *c1 = blk + (long long)(i - 1) * S;
                double *c2 = blk + (long long)i * S;
                std::fill(fg.begin(), fg.end(), 0.0);
                for (int j = 0; j < S; j++) fg[B.states[j].time] += c1[j];
                for (int bb = 0; bb < T - 1; bb++) {
                    double sum = 0.0;
                    for (int a = 0; a < T - 1; a++) sum += tmat[a * T + bb] * fg[a];
                    tf[bb] = sum;
                }
                double norm = 0.0;
                for (int k = 0; k < S; k++) {
                    int bb = B.states[k].time, n2 = B.states[k].node;
                    double sum = tf[bb];
                    for (int j = idx[n2], a = ages1[n2]; a <= ages2[n2]; j++, a++)
                        sum += tmat2[(size_t)a * S + k] * c1[j];
                    c2[k] = sum * emit[(size_t)i * S + k];
                    norm += c2[k];
                }
                for (int k = 0; k < S; k++) c2[k] /= norm;
                logl += log(norm);
            }
        }
        prev = blk + (long long)(B.len - 1) * S;
        off += (long long)B.len * S;
    }
    if (lnl) *lnl = logl;
}

// stochastic_traceback (src/argweaver/sample_thread.cpp:470-569)
void orc_traceback(const orc_instance *in, const double *fw, const double *u, int *path)
{
    Model m = make_model(in);
    auto bl = blocks_of(in, m);
    const int nb = in->ntrees;
    std::vector<long long> off(nb + 1, 0);
    for (int b = 0; b < nb; b++)
        off[b + 1] = off[b] + (long long)bl[b].len * std::max((int)bl[b].states.size(), 1);
    const int L = in->seqlen;
    int draw = 0;
    // column c (relative to start_coord) of block b: fw + off[b] + (c - start)*S
    auto col = [&](int b, int c) {
        return fw + off[b] + (long long)(c - (bl[b].start - in->start_coord)) *
                                 (long long)bl[b].states.size();
    };
    {
        const Block &B = bl[nb - 1];
        path[L - 1] = sample_u(col(nb - 1, L - 1), (int)B.states.size(), u[draw++]);
    }
    std::vector<double> A, tr;
    for (int b = nb - 1; b >= 0; b--) {
        const Block &B = bl[b];
        const int S = (int)B.states.size();
        const int start = B.start - in->start_coord;
        A.resize(S);
        tr.resize(S);
        int last_k = -1;
        for (int i = start + B.len - 2; i >= start; i--) {
            int k = path[i + 1];
            if (k != last_k) {
                for (int j = 0; j < S; j++) tr[j] = B.trans.get(B.tree, B.states, j, k);
                last_k = k;
            }
            const double *c = col(b, i);
            for (int j = 0; j < S; j++) A[j] = c[j] * tr[j];
            path[i] = sample_u(A.data(), S, u[draw++]);
        }
        if (b > 0) {
            Switch sw = block_switch(bl, b, m);
            const int S1 = (int)bl[b - 1].states.size();
            const double *c = col(b - 1, start - 1);
            std::vector<double> A1(S1);
            for (int j = 0; j < S1; j++) A1[j] = c[j] * sw.get(j, path[start]);
            path[start - 1] = sample_u(A1.data(), S1, u[draw++]);
        }
    }
}

// sample_recombinations (src/argweaver/recomb.cpp:14-235), external threading
int orc_sample_recombinations(const orc_instance *in, const int *path, int *recomb_pos,
                              int *recomb_node, int *recomb_time, int maxrecomb)
{
    Model m = make_model(in);
    auto bl = blocks_of(in, m);
    int nrec = 0;
    std::vector<State> cand;
    std::vector<double> probs;
    for (int b = 0; b < in->ntrees; b++) {
        const Block &B = bl[b];
        const Tree &t = B.tree;
        const States &s = B.states;
        int next_recomb = -1;
        int start = B.start - in->start_coord;
        int end = start + B.len;
        start++;   // switch column or very first column
        for (int i = start; i < end; i++) {
            if (path[i] == path[i - 1]) {
                if (i > next_recomb) {
                    int ls = path[i - 1];
                    int a = s[ls].time;
                    double self = B.trans.get(t, s, ls, ls);
                    double rate = 1.0 - (B.trans.norecombs[a] / self);
                    double ev = -log(frand()) / rate;   // expovariate (common.h:124-125)
                    next_recomb = int(std::min(double(end), i + ev));
                }
                if (i < next_recomb) continue;
            }
            next_recomb = -1;
            State st = s[path[i]], last = s[path[i - 1]];
            // get_possible_recomb (recomb.cpp:124-148)
            cand.clear();
            int end_time = std::min(st.time, last.time);
            if (st.node == last.node)
                for (int k = t.nodes[st.node].age; k <= end_time; k++) cand.push_back({st.node, k});
            for (int k = 0; k <= end_time; k++) cand.push_back({-1, k});
            // recomb_prob_unnormalized (recomb.cpp:14-119)
            probs.clear();
            for (const State &rc : cand) {
                int k = rc.time, j = st.time;
                int root_time = std::max(t.nodes[t.root].age, last.time);
                int rpa = (rc.node == -1 || t.nodes[rc.node].parent == -1 || rc.node == last.node)
                              ? last.time
                              : t.nodes[t.nodes[rc.node].parent].age;
                int nbk = B.lin.nb[k] + int(k < last.time);
                int nrk = B.lin.nr[k] + int(k <= last.time) + int(k == last.time) - int(k == root_time);
                double precomb = nbk * m.steps[k] / nrk;
                double sum = 0.0;
                for (int x = k; x < j - 1; x++) {
                    int nbm = B.lin.nb[x] + int(x < last.time) - int(x < rpa);
                    sum += (m.steps[x] * nbm / (2.0 * m.pop[x]));
                }
                double pcoal = 1.0;
                int nbj = B.lin.nb[j] + int(j < last.time) - int(j < rpa);
                if (k == j) {
                    if (j < m.T - 2)
                        pcoal = 1.0 - exp(-m.csteps[2 * j] * nbj / (2.0 * m.pop[j]));
                } else {
                    int x = j - 1;
                    int nbm = B.lin.nb[x] + int(x < last.time) - int(x < rpa);
                    sum += (m.csteps[2 * x] * nbm / (2.0 * m.pop[x]));
                    if (j < m.T - 2)
                        pcoal = 1.0 - exp(-m.csteps[2 * j - 1] * nbm / (2.0 * m.pop[x]) -
                                          m.csteps[2 * j] * nbj / (2.0 * m.pop[j]));
                }
                int ncj = B.lin.nc[j] - int(j <= rpa) - int(j == rpa) + int(j <= last.time) +
                          int(j == last.time);
                pcoal /= ncj;
                probs.push_back(precomb * exp(-sum) * pcoal);
            }
            int pick = sample_u(probs.data(), (int)probs.size(), frand());
            if (nrec < maxrecomb) {
                recomb_pos[nrec] = i + in->start_coord;
                recomb_node[nrec] = cand[pick].node;
                recomb_time[nrec] = cand[pick].time;
            }
            nrec++;
        }
    }
    return nrec;
}

}  // extern "C"
