// aw_math.h -- per-block threading-HMM math shared by the host drivers and
// the gfx950 prep kernels.
//
// Everything here works on one local tree stored as flat arrays
// (parent/child0/child1/age, leaves first) with fixed upper bounds so the
// same code runs inside a kernel without allocation.  Formulas follow the
// reference (citations are file:line in the reference repository):
//   lineage counts      src/argweaver/local_tree.cpp:34-69
//   tree lengths        src/argweaver/local_tree.cpp:136-219
//   TransMatrix         src/argweaver/trans.cpp:13-115, trans.h:89-128
//   TransMatrixSwitch   src/argweaver/trans.cpp:156-739
//   state priors        src/argweaver/trans.cpp:785-831
//   emission terms      src/argweaver/emit.cpp:86-116, 620-869
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace aw {

constexpr int MAXT = 32;      // ntimes upper bound
constexpr int MAXN = 256;     // nodes per local tree (128 leaves)

struct Model {
    int T;
    double rho, mu;
    double times[MAXT];
    double steps[MAXT];       // times[i+1]-times[i], steps[T-1] = inf
    double csteps[2 * MAXT];  // half-interval steps (model.cpp:9-21)
    double pop[MAXT];
};

// Build the model tables from the caller's time points and population sizes
// (ArgModel::set_times / setup_time_steps, model.h:322-333).
__host__ __device__ inline void model_init(Model &m, int T, const double *times,
                                           const double *pop, double rho, double mu)
{
    m.T = T;
    m.rho = rho;
    m.mu = mu;
    for (int i = 0; i < T; i++) {
        m.times[i] = times[i];
        m.pop[i] = pop[i];
    }
    for (int i = 0; i < T - 1; i++) m.steps[i] = times[i + 1] - times[i];
    m.steps[T - 1] = INFINITY;
    // geometric midpoints; entries 2T-3.. are never read on this path
    double t2[2 * MAXT + 1];
    for (int i = 0; i < T - 1; i++) {
        t2[2 * i] = times[i];
        t2[2 * i + 1] = sqrt((times[i + 1] + 1.0) * (times[i] + 1.0));
    }
    t2[2 * T - 2] = times[T - 1];
    for (int i = 0; i < 2 * T - 2; i++) m.csteps[i] = t2[i + 1] - t2[i];
    m.csteps[2 * T - 2] = INFINITY;
    m.csteps[2 * T - 1] = INFINITY;
}

// A local tree viewed through flat arrays.
struct TreeView {
    int n;          // nodes
    int root;
    const int *parent, *c0, *c1, *age;

    __host__ __device__ int sibling(int node) const
    {
        int p = parent[node];
        if (p == -1) return -1;
        return c0[p] == node ? c1[p] : c0[p];
    }
    // top time of the branch above node: parent age, or T-2 for the root
    __host__ __device__ int top(int node, int T) const
    {
        int p = parent[node];
        return p == -1 ? T - 2 : age[p];
    }
};

struct Lineages {
    int nb[MAXT], nr[MAXT], nc[MAXT];
};

__host__ __device__ inline void count_lineages(const TreeView &t, int T, Lineages &L)
{
    for (int i = 0; i < T; i++) L.nb[i] = L.nr[i] = L.nc[i] = 0;
    for (int i = 0; i < t.n; i++) {
        int p = t.parent[i];
        int page = (p == -1) ? T - 2 : t.age[p];
        for (int j = t.age[i]; j < page; j++) {
            L.nb[j]++;
            L.nr[j]++;
            L.nc[j]++;
        }
        L.nr[page]++;
        L.nc[page]++;
        if (p == -1) L.nb[page]++;
    }
    L.nb[T - 1] = 1;
}

// tree length without the basal stub
__host__ __device__ inline double tree_length(const TreeView &t, const double *times)
{
    double len = 0.0;
    for (int i = 0; i < t.n; i++) {
        int p = t.parent[i];
        if (p != -1) len += times[t.age[p]] - times[t.age[i]];
    }
    return len;
}

__host__ __device__ inline double logadd(double a, double b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    return fmax(a, b) + log1p(exp(-fabs(a - b)));
}

// The nine per-time vectors of the compressed within-block transition
// matrix (TransMatrix).
struct TransVec {
    double D[MAXT], E[MAXT], lnB[MAXT], lnE2[MAXT], lnNegG1[MAXT];
    double G2[MAXT], G3[MAXT], lnG4[MAXT], norecombs[MAXT];

    // P((node_i,a) -> (node_j,b)) with c = age(node_j); external threading, minage = 0
    __host__ __device__ double get_time(int a, int b, int c, bool same) const
    {
        double term1 = D[a] * E[b];
        double low = (b > 0 ? exp(lnE2[b] + lnB[b - 1]) : 0.0);
        if (!same) {
            if (a < b) return term1 * (exp(lnE2[b] + lnB[a]) - exp(lnE2[b] + lnNegG1[a]));
            if (a == b) return term1 * (low + G3[b]);
            return term1 * (low + G2[b]);
        }
        double c_term = (c > 0 ? exp(lnG4[b] + lnB[c - 1]) : 0.0);
        if (a < b)
            return term1 * (2 * (exp(lnE2[b] + lnB[a]) - exp(lnE2[b] + lnNegG1[a])) - c_term);
        if (a == b) return term1 * ((2 * (low + G3[b])) - c_term) + norecombs[a];
        return term1 * ((2 * (low + G2[b])) - c_term);
    }
};

__host__ __device__ inline void calc_trans(const TreeView &t, const Model &m, const Lineages &L,
                                           TransVec &tv)
{
    const int T = m.T;
    double rates_a[2 * MAXT + 1];
    double C_a[2 * MAXT + 2];
    double *rates = rates_a + 1;
    double *C = C_a + 2;
    rates[-1] = 0.0;
    for (int i = 0; i < 2 * T - 1; i++) rates[i] = m.csteps[i] * L.nb[i / 2] / (2.0 * m.pop[i / 2]);
    C[-2] = 0.0;
    C[-1] = 0.0;
    for (int b = 0; b < 2 * T - 1; b++) C[b] = C[b - 1] + rates[b];

    const int ri = t.age[t.root];
    const double root_age = m.times[ri];
    const double len = tree_length(t, m.times);
    for (int b = 0; b < T; b++) {
        tv.D[b] = tv.E[b] = tv.lnB[b] = tv.lnE2[b] = tv.lnNegG1[b] = 0.0;
        tv.G2[b] = tv.G3[b] = tv.lnG4[b] = tv.norecombs[b] = 0.0;
    }
    for (int b = 0; b < T - 1; b++) {
        const int nb = L.nb[b], nr = L.nr[b], nc = L.nc[b];
        double len2 = len + m.times[b];
        double len2_b;
        if (b > ri) {
            len2 += m.times[b] - root_age;
            len2_b = len2 + m.steps[b];
        } else {
            len2_b = len2 + m.steps[ri];
        }
        double term = C[2 * b - 1] + log(m.steps[b] * (nb + 1.0) / (nr + 1.0));
        tv.lnB[b] = (b == 0) ? term : logadd(tv.lnB[b - 1], term);
        double e2 = -C[2 * b - 2] + (b < T - 2 ? log(1 - exp(-rates[2 * b] - rates[2 * b - 1])) : 0.0);
        tv.lnE2[b] = e2;
        tv.lnG4[b] = e2;
        tv.lnNegG1[b] = C[2 * b - 1] +
            log(-m.steps[b] * ((nb / (nr + 1.0 + int(b < ri))) - (nb + 1.0) / (nr + 1.0)));
        double q = (b < T - 2 ? 1.0 - exp(-rates[2 * b]) : 1.0);
        tv.G2[b] = q * m.steps[b] * (nb + 1.0) / (nr + 1.0);
        tv.G3[b] = q * m.steps[b] * (nb / (nr + 1.0 + int(b < ri)));
        tv.D[b] = (1.0 - exp(-m.rho * len2)) / len2_b;
        tv.E[b] = 1.0 / nc;
        tv.norecombs[b] = exp(-fmax(m.rho * len2, m.rho));
    }
}

// P(first-column state at time b) (calc_state_priors, minage = 0)
__host__ __device__ inline double prior_time(int b, const Lineages &L, const Model &m)
{
    double sum = 0.0;
    for (int x = 0; x < 2 * b - 1; x++) sum += m.csteps[x] * L.nb[x / 2] / (2.0 * m.pop[x / 2]);
    double p = exp(-sum) / L.nc[b];
    if (b < m.T - 2) {
        double Z = 0.0;
        if (b > 0) Z = m.csteps[2 * b - 1] * L.nb[b - 1] / (2.0 * m.pop[b - 1]);
        p *= 1.0 - exp(-m.csteps[2 * b] * L.nb[b] / (2.0 * m.pop[b]) - Z);
    }
    return p;
}

//---------------------------------------------------------------------------
// switch matrix pieces (block b-1 -> block b)

struct Spr {
    int rn, rt, cn, ct;
};

// per-node state indexing of a node-major state list
struct NodeIndex {
    int first[MAXN];   // index of (node, age(node))
    int hi[MAXN];      // top time of the node's branch
};

__host__ __device__ inline int lookup(const TreeView &t, const NodeIndex &ni, int node, int time)
{
    if (node < 0) return -1;
    if (time < t.age[node] || time > ni.hi[node]) return -1;
    return ni.first[node] + time - t.age[node];
}

// recombination part of a switch transition (calc_recomb, external)
__host__ __device__ inline double sw_recomb(const TreeView &last, const Model &m,
                                            const Lineages &L, const Spr &spr, int node1,
                                            int a, double last_len)
{
    const int k = spr.rt;
    const int root_age = last.age[last.root];
    double len = last_len + m.times[a];
    double basal;
    if (node1 == last.root) {
        len += m.times[a] - m.times[last.age[last.root]];
        basal = m.times[a + 1] - m.times[a];
    } else {
        basal = m.times[root_age + 1] - m.times[root_age];
    }
    double len_b = len + basal;
    int nbk = L.nb[k] + int(k < a);
    int nrk = L.nr[k] + int(k <= a) + int(k == a) - int(k >= (root_age > a ? root_age : a));
    return nbk * m.steps[k] / (nrk * len_b) * (1.0 - exp(-fmax(m.rho * len, m.rho)));
}

// re-coalescence part (calc_recoal, external)
__host__ __device__ inline double sw_recoal(const Model &m, const Lineages &L, int k, int j,
                                            int a, int rpa)
{
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

// calc_recomb_recoal with coal time j (spr.ct replaced by j)
__host__ __device__ inline double sw_recomb_recoal(const TreeView &last, const Model &m,
                                                   const Lineages &L, const Spr &spr, int j,
                                                   int node1, int a, int rpa, double last_len)
{
    const int k = spr.rt;
    double p = sw_recomb(last, m, L, spr, node1, a, last_len);
    double sum = 0.0;
    for (int x = 2 * k; x < 2 * j - 1; x++) {
        int nbm = L.nb[x / 2] - int(x / 2 < rpa) + int(x / 2 < a);
        sum += m.csteps[x] * nbm / (2.0 * m.pop[x / 2]);
    }
    p *= exp(-sum);
    p *= sw_recoal(m, L, k, j, a, rpa);
    return p;
}

// Shared per-switch scalars (calc_recoal_sums and friends).
struct SwitchCtx {
    double last_len;
    int rpa;                    // age of the recomb node's parent in the last tree
    double sums;
    double sums2[2 * MAXT + 1];
    double recoals[MAXT];
};

// everything of the switch context but the per-time recoalescence terms
// (sw_recoal(m, L, spr.rt, spr.ct, a, rpa) for every time a), given the
// last tree's length
__host__ __device__ inline void switch_ctx_sums(const TreeView &last, const Model &m,
                                                const Lineages &L, const Spr &spr, double last_len,
                                                SwitchCtx &cx)
{
    cx.last_len = last_len;
    cx.rpa = last.age[last.parent[spr.rn]];
    const int k = spr.rt, j = spr.ct;
    double sum = 0.0;
    for (int x = 2 * k; x < 2 * j - 1; x++) {
        int nbm = L.nb[x / 2] - int(x / 2 < cx.rpa);
        sum += m.csteps[x] * nbm / (2.0 * m.pop[x / 2]);
    }
    cx.sums = sum;
    for (int x = 0; x < 2 * MAXT + 1; x++) cx.sums2[x] = 0.0;
    sum = 0.0;
    cx.sums2[2 * k] = sum;
    for (int x = 2 * k; x < 2 * j - 1; x++) {
        sum += m.csteps[x] / (2.0 * m.pop[x / 2]);
        cx.sums2[x + 1] = sum;
    }
}

// Deterministic target of last-tree state (node1,time1) (get_deterministic_transitions)
__host__ __device__ inline int sw_determ(const TreeView &last, const TreeView &t,
                                         const NodeIndex &ni2, const Spr &spr,
                                         const int *mapping, int node1, int time1)
{
    if ((node1 == spr.cn && time1 == spr.ct) || (node1 == spr.rn && time1 == spr.rt)) return -1;
    if (node1 != spr.rn) {
        int node2;
        bool disrupt = false;
        if (last.c0[node1] == -1) {
            node2 = node1;
        } else if (spr.rn == last.c0[node1]) {
            node2 = mapping[last.c1[node1]];
            disrupt = true;
        } else if (spr.rn == last.c1[node1]) {
            node2 = mapping[last.c0[node1]];
            disrupt = true;
        } else {
            node2 = mapping[node1];
        }
        if ((spr.cn == node1 && spr.ct < time1) || (mapping[spr.cn] == node2 && spr.ct < time1) ||
            (disrupt && mapping[spr.cn] == node2 && spr.ct <= time1))
            node2 = t.parent[node2];
        return lookup(t, ni2, node2, time1);
    }
    if (spr.rt > time1) return lookup(t, ni2, mapping[spr.rn], time1);
    int parent = last.parent[spr.rn];
    int time2 = last.age[parent];
    int other = (last.c1[parent] == spr.rn ? last.c0[parent] : last.c1[parent]);
    int node2 = (other == spr.cn ? t.parent[mapping[other]] : mapping[other]);
    return lookup(t, ni2, node2, time2);
}

// weight of a deterministic transition from (node1,time1) (calc_transition_probs_switch)
__host__ __device__ inline double sw_determprob(const TreeView &last, const Model &m,
                                                const Lineages &L, const Spr &spr,
                                                const SwitchCtx &cx, int node1, int time1)
{
    if (node1 == spr.rn && time1 > spr.rt)
        return sw_recomb_recoal(last, m, L, spr, spr.ct, node1, time1, time1, cx.last_len);
    int idx2 = 2 * time1 < 2 * spr.ct - 1 ? 2 * time1 : 2 * spr.ct - 1;
    int idx = idx2 > 2 * spr.rt ? idx2 : 2 * spr.rt;
    return sw_recomb(last, m, L, spr, node1, time1, cx.last_len) *
           exp(-cx.sums - cx.sums2[idx]) * cx.recoals[time1];
}

//---------------------------------------------------------------------------
// emissions (JC69)

__host__ __device__ inline double prob_branch(double t, double mu, bool mut)
{
    const double f = 4. / 3.;
    if (!mut) return .25 * (1.0 + 3. * exp(-f * mu * t));
    return .25 * (1.0 - exp(-f * mu * t));
}

__host__ __device__ inline int base_code(char c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    }
    return -1;   // 'N' (or anything else): all-ones partial likelihood
}

// per-state emission terms: 3 branch mutation probabilities and the
// invariant-column likelihood (emit.cpp:778-819)
struct EmitTerms {
    double mut[3], nomut[3], inv;
};

// the terms that depend on the state's time only: the new leaf's branch
// (dist[0]) and the invariant-column likelihood, for a state off and on the
// root branch
struct EmitTime {
    double mut0, nomut0, inv, inv_root;
};

__host__ __device__ inline void emit_time_terms(const TreeView &t, const Model &m, double mainlen,
                                                int time, EmitTime &et)
{
    const double mintime = m.times[1] * .1;
    const double coal_time = m.times[time];
    const double d0 = fmax(coal_time - 0.0, mintime);
    et.mut0 = prob_branch(d0, m.mu, true);
    et.nomut0 = prob_branch(d0, m.mu, false);
    const double len = mainlen + fmax(coal_time - 0.0, mintime);
    et.inv = .25 * exp(-m.mu * fmax(len, mintime));
    const double len_r = mainlen + fmax(coal_time - 0.0, mintime) +
                         fmax(coal_time - m.times[t.age[t.root]], mintime);
    et.inv_root = .25 * exp(-m.mu * fmax(len_r, mintime));
}

__host__ __device__ inline void emit_terms(const TreeView &t, const Model &m, const EmitTime &et,
                                           int node2, int time, EmitTerms &e)
{
    const double mintime = m.times[1] * .1;
    int parent = t.parent[node2];
    double time2 = m.times[t.age[node2]];
    double parent_time = (parent != -1) ? m.times[min(t.age[parent], m.T - 1)] : 0.0;
    double coal_time = m.times[time];
    double dist[3];
    dist[1] = fmax(coal_time - time2, mintime);
    dist[2] = fmax(parent_time - coal_time, mintime);
    e.mut[0] = et.mut0;
    e.nomut[0] = et.nomut0;
    for (int x = 1; x < 3; x++) {
        e.mut[x] = prob_branch(dist[x], m.mu, true);
        e.nomut[x] = prob_branch(dist[x], m.mu, false);
    }
    e.inv = (node2 == t.root) ? et.inv_root : et.inv;
}

// calc_emit for one state given the node's inner/outer partials and the
// new leaf's vector (emit.cpp:620-645)
__host__ __device__ inline double emit_variant(const double *in2, const double *inner,
                                               const double *outer, bool is_root,
                                               const double *mut, const double *nomut)
{
    double e = 0.0;
    for (int a = 0; a < 4; a++) {
        double p1 = 0.0, p2 = 0.0, p3 = 0.0;
        for (int b = 0; b < 4; b++) {
            if (a == b) {
                p1 += in2[b] * nomut[0];
                p2 += inner[b] * nomut[1];
                p3 += outer[b] * nomut[2];
            } else {
                p1 += in2[b] * mut[0];
                p2 += inner[b] * mut[1];
                p3 += outer[b] * mut[2];
            }
        }
        if (!is_root) e += p1 * p2 * p3 * .25;
        else e += p1 * p2 * .25;
    }
    return e;
}

__host__ __device__ inline void leaf_partial(char c, double *v)
{
    int x = base_code(c);
    for (int a = 0; a < 4; a++) v[a] = (x < 0) ? 1.0 : (a == x ? 1.0 : 0.0);
}

// Felsenstein inner (postorder) and outer (preorder) partials of one column
// (emit.cpp:151-324).  `order` is a postorder of the nodes, `bases[i]` the
// character of leaf i; muts/nomuts per node.  Serial: run by one lane.
__host__ __device__ inline void pruning(const TreeView &t, const int *order, const char *bases,
                                        const double *muts, const double *nomuts, double *inner,
                                        double *outer)
{
    for (int oi = 0; oi < t.n; oi++) {
        int j = order[oi];
        if (t.c0[j] == -1) {
            leaf_partial(bases[j], inner + 4 * j);
            continue;
        }
        int c1 = t.c0[j], c2 = t.c1[j];
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
    // reverse postorder visits every parent before its children
    for (int oi = t.n - 1; oi >= 0; oi--) {
        int j = order[oi];
        if (j == t.root) {
            for (int a = 0; a < 4; a++) outer[4 * j + a] = 1.0;
            continue;
        }
        int sib = t.sibling(j);
        int parent = t.parent[j];
        for (int a = 0; a < 4; a++) {
            double p1 = 0.0, p2 = 0.0;
            for (int b = 0; b < 4; b++) {
                double ms = (a == b) ? nomuts[sib] : muts[sib];
                p1 += inner[4 * sib + b] * ms;
                if (parent != t.root) {
                    double mp = (a == b) ? nomuts[parent] : muts[parent];
                    p2 += outer[4 * parent + b] * mp;
                }
            }
            outer[4 * j + a] = (parent != t.root) ? p1 * p2 : p1;
        }
    }
}

// postorder with leaves first (LocalTree::get_postorder)
__host__ __device__ inline void postorder(const TreeView &t, int *order)
{
    unsigned char visit[MAXN];
    for (int i = 0; i < t.n; i++) visit[i] = 0;
    int i;
    for (i = 0; i < t.n; i++) {
        if (t.c0[i] != -1) break;
        order[i] = i;
    }
    int end = i;
    for (i = 0; i < t.n; i++) {
        int p = t.parent[order[i]];
        if (p != -1 && ++visit[p] == 2) order[end++] = p;
    }
}

}  // namespace aw
