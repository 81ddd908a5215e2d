// arg_stats.cpp -- probability of an ARG: the data likelihood given the
// local trees and the prior of the trees and their SPRs (total_prob.cpp),
// on the host.  These are the statistics a resample_arg chain reports per
// iteration; they read the ARG handle and the sequences once.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/argweaver_amd.h"
#include "aw_math.h"
#include "host_util.h"
#include "local_trees.h"

namespace aw {
namespace {

// children before parents (a node's partials depend on its children's
// only, so every such order gives the same values)
void child_first_order(const TreeView &t, std::vector<int> &order)
{
    const int N = t.n;
    order.clear();
    std::vector<int> seen(N, 0);
    for (int i = 0; i < N; i++)
        if (t.c0[i] == -1) order.push_back(i);
    for (size_t x = 0; x < order.size(); x++) {
        const int p = t.parent[order[x]];
        if (p != -1 && ++seen[p] == 2) order.push_back(p);
    }
}

// likelihood_tree (emit.cpp:399-449): sum over columns [start, end) of the
// log site likelihood under JC69 with a uniform root.  rows[j] is leaf j's
// sequence.  The first invariant column's likelihood is reused for every
// later invariant column of the tree, as the reference does.
double tree_loglik(const TreeView &t, const Model &m, const char *const *rows, int nrows, int start,
                   int end)
{
    const int N = t.n;
    const double mintime = m.times[1] * .1;
    std::vector<double> muts(N, 0.0), nomuts(N, 0.0);
    for (int i = 0; i < N; i++) {
        if (i == t.root) continue;
        const double tt = fmax(m.times[t.age[t.parent[i]]] - m.times[t.age[i]], mintime);
        muts[i] = prob_branch(tt, m.mu, true);
        nomuts[i] = prob_branch(tt, m.mu, false);
    }
    std::vector<int> order;
    child_first_order(t, order);
    std::vector<std::array<double, 4>> inner(N);
    double invariant_lk = -1.0;
    double lnl = 0.0;
    for (int col = start; col < end; col++) {
        bool invariant = true;
        for (int j = 1; j < nrows && invariant; j++) invariant = rows[j][col] == rows[0][col];
        double lk;
        if (invariant && invariant_lk > 0) {
            lk = invariant_lk;
        } else {
            for (int j : order) {
                if (t.c0[j] == -1) {
                    const int b = base_code(rows[j][col]);   // -1: 'N', all ones
                    for (int a = 0; a < 4; a++) inner[j][a] = (b < 0 || a == b) ? 1.0 : 0.0;
                    continue;
                }
                const int c1 = t.c0[j], c2 = t.c1[j];
                for (int a = 0; a < 4; a++) {
                    double p1 = 0.0, p2 = 0.0;
                    for (int b = 0; b < 4; b++) {
                        p1 += inner[c1][b] * (a == b ? nomuts[c1] : muts[c1]);
                        p2 += inner[c2][b] * (a == b ? nomuts[c2] : muts[c2]);
                    }
                    inner[j][a] = p1 * p2;
                }
            }
            lk = 0.0;
            for (int a = 0; a < 4; a++) lk += inner[t.root][a] * .25;
            if (invariant) invariant_lk = lk;
        }
        lnl += log(lk);
    }
    return lnl;
}

// prob_coal_counts (total_prob.cpp:106-125): P(a lineages become b in time
// t at population size n)
double prob_coal_counts(int a, int b, double t, double n)
{
    double C = 1.0;
    for (int y = 0; y < b; y++) C *= (b + y) * (a - y) / double(a + y);
    double s = exp(-b * (b - 1) * t / 2.0 / n) * C;
    for (int k = b + 1; k < a + 1; k++) {
        const double k1 = double(k - 1);
        C *= double(b + k1) * (a - k1) / (a + k1) / (b - k);
        s += exp(-k * k1 * t / 2.0 / n) * (2 * k - 1) / double(k1 + b) * C;
    }
    for (int i = 1; i <= b; i++) s /= i;
    return s;
}

}  // namespace

// calc_spr_prob (total_prob.cpp:213-257): log probability of the SPR that
// leads out of tree t (recombination point, then re-coalescence)
double spr_logprob(const TreeView &t, const Model &m, const Spr &spr, double treelen)
{
    const int T = m.T;
    const int root_age = t.age[t.root];
    const double treelen_b = treelen + m.steps[root_age];
    Lineages L;
    count_lineages(t, T, L);
    L.nr[root_age]--;
    double lnl = 0.0;
    const int k = spr.rt;
    lnl += log(L.nb[k] * m.steps[k] / (L.nr[k] * treelen_b));
    const int broken_age = t.age[t.parent[spr.rn]];
    double rates[2 * MAXT];
    for (int i = 0; i < 2 * T; i++) {
        const int nb = L.nb[i / 2] - int(i / 2 < broken_age);
        rates[i] = m.csteps[i] * nb / (2.0 * m.pop[i / 2]);
    }
    const int j = spr.ct;
    const int ncoals_j = L.nc[j] - int(j <= broken_age) - int(j == broken_age);
    lnl -= log(ncoals_j);
    if (j < T - 2) lnl += log(1.0 - exp(-rates[2 * j] - (j > k ? rates[2 * j - 1] : 0.0)));
    for (int x = 2 * k; x < 2 * j - 1; x++) lnl -= rates[x];
    return lnl;
}

namespace {

// calc_arg_prior (total_prob.cpp:262-299)
double arg_logprior(const ArgTrees &tr, const Model &m)
{
    double lnl = 0.0;
    int end = tr.start_coord;
    for (size_t bi = 0; bi < tr.blocks.size(); bi++) {
        const LBlock &B = tr.blocks[bi];
        end += B.blocklen;
        FlatTree ft(B.tree);
        const double treelen = tree_length(ft.view, m.times);
        const double recomb_rate = fmax(m.rho * treelen, m.rho);
        if (end < tr.end_coord) {
            lnl += log(recomb_rate) - recomb_rate * B.blocklen;
            lnl += spr_logprob(ft.view, m, tr.blocks[bi + 1].spr, treelen);
        } else {
            lnl += -recomb_rate * B.blocklen;
        }
    }
    return lnl;
}

// calc_tree_prior (total_prob.cpp:128-163) of one tree: coalescent counts
// per time interval with harmonic-mean population sizes
double tree_logprior(const TreeView &t, const Model &m)
{
    const int T = m.T;
    Lineages L;
    count_lineages(t, T, L);
    const int nleaves = (t.n + 1) / 2;
    std::vector<double> times(T), pops(T);
    times[0] = m.csteps[0];
    pops[0] = m.pop[0];
    for (int i = 1; i < T - 1; i++) {
        const double t1 = m.csteps[2 * i - 1], t2 = m.csteps[2 * i];
        const double n1 = m.pop[i - 1], n2 = m.pop[i];
        times[i] = t1 + t2;
        pops[i] = (t1 + t2) / (t1 / n1 + t2 / n2);
    }
    double lnl = 0.0;
    for (int i = 0; i < T - 1; i++) {
        const int a = (i == 0 ? nleaves : L.nb[i - 1]);
        const int b = L.nb[i];
        lnl += log(prob_coal_counts(a, b, times[i], 2.0 * pops[i]));
    }
    return lnl;
}

// calc_arg_likelihood (total_prob.cpp:19-45): leaf j reads sequence row
// seqids[j]; a column is invariant when all nseqs rows agree
double arg_loglik(const ArgTrees &tr, const Model &m, char **seqs, int nseqs, int seqlen)
{
    if (tr.nnodes < 3) return log(.25) * seqlen;
    std::vector<const char *> rows(nseqs);
    for (int j = 0; j < nseqs; j++) rows[j] = seqs[j < (int)tr.seqids.size() ? tr.seqids[j] : j];
    double lnl = 0.0;
    int end = tr.start_coord;
    for (const LBlock &B : tr.blocks) {
        const int start = end;
        end = start + B.blocklen;
        FlatTree ft(B.tree);
        lnl += tree_loglik(ft.view, m, rows.data(), nseqs, start, end);
    }
    return lnl;
}

Model make_model(const double *times, int ntimes, const double *popsizes, double rho, double mu)
{
    std::vector<double> pops(ntimes, 1.0);
    if (popsizes)
        for (int i = 0; i < ntimes; i++) pops[i] = popsizes[i];
    Model m;
    model_init(m, ntimes, times, pops.data(), rho, mu);
    return m;
}

bool check_model(const char *who, int ntimes)
{
    if (ntimes < 3 || ntimes > MAXT) {
        fprintf(stderr, "argweaver_amd: %s: ntimes %d outside 3..%d\n", who, ntimes, MAXT);
        return false;
    }
    return true;
}

}  // namespace
}  // namespace aw

using namespace aw;

extern "C" {

double arghmm_likelihood(LocalTrees *trees, double *times, int ntimes, double mu, char **seqs,
                         int nseqs, int seqlen)
{
    if (!check_model("arghmm_likelihood", ntimes)) return NAN;
    const Model m = make_model(times, ntimes, nullptr, 0.0, mu);
    return arg_loglik(*reinterpret_cast<ArgTrees *>(trees), m, seqs, nseqs, seqlen);
}

double arghmm_likelihood_parsimony(LocalTrees *, double *, int, double, char **, int, int)
{
    // the reference's export aborts (total_prob.cpp:329-342)
    fprintf(stderr, "argweaver_amd: arghmm_likelihood_parsimony is not available\n");
    abort();
}

double arghmm_prior_prob(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                         double rho)
{
    if (!check_model("arghmm_prior_prob", ntimes)) return NAN;
    const Model m = make_model(times, ntimes, popsizes, rho, 0.0);
    return arg_logprior(*reinterpret_cast<ArgTrees *>(trees), m);
}

double arghmm_tree_prior_prob(LocalTrees *trees, double *times, int ntimes, double *popsizes)
{
    if (!check_model("arghmm_tree_prior_prob", ntimes)) return NAN;
    const Model m = make_model(times, ntimes, popsizes, 0.0, 0.0);
    const ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    FlatTree ft(tr.blocks.front().tree);
    return tree_logprior(ft.view, m);
}

double arghmm_joint_prob(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                         double mu, double rho, char **seqs, int nseqs, int seqlen)
{
    if (!check_model("arghmm_joint_prob", ntimes)) return NAN;
    const Model m = make_model(times, ntimes, popsizes, rho, mu);
    const ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    return arg_loglik(tr, m, seqs, nseqs, seqlen) + arg_logprior(tr, m);
}

}  // extern "C"
