// host_hmm.cpp -- the reference's dense-matrix and generic-HMM exports, on
// the host: within-block and switch transition matrices, emission tables,
// the brute-force checks the reference runs against them, the generic
// log-space forward / backward / posterior sampler, and the coalescent
// population-size estimate from the local trees.
//
// None of these is on the threading path: the device kernels never form a
// dense S x S matrix or a full emission table.  They keep the C ABI complete
// for callers that use them (the reference's test_hmm.py, for one), and they
// are built from the same per-time pieces as the device preparation kernel
// (aw_math.h), so a dense matrix here holds the entries the kernels use.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/argweaver_amd.h"
#include "aw_math.h"
#include "host_util.h"
#include "local_trees.h"

namespace aw {

// relative-or-absolute closeness used by the reference's checks (fequal,
// common.h)
bool close_enough(double a, double b, double rel, double eabs)
{
    if (a == b) return true;
    const double d = fabs(a - b);
    if (d <= eabs) return true;
    return d / fmax(fabs(a), fabs(b)) <= rel;
}

// likelihood_sites (emit.cpp:328-397) for one state; the callers attach the
// new branch first (calc_emissions_external_slow, emit.cpp:1138-1168, and
// calc_emissions_internal_slow, emit.cpp:1171-1214)
void emission_column_slow(const LTree &t2, const Model &m, const char *const *rows, int seqlen,
                          const std::vector<char> &invariant, int statei, double **emit)
{
    FlatTree ft(t2);
    const TreeView &t = ft.view;
    const int N = t.n;
    const double mintime = m.times[1] * .1;
    std::vector<double> muts(N, 0.0), nomuts(N, 0.0);
    double treelen = 0.0;
    for (int i = 0; i < N; i++) {
        if (i == t.root) continue;
        const double d = fmax(m.times[t.age[t.parent[i]]] - m.times[t.age[i]], mintime);
        muts[i] = prob_branch(d, m.mu, true);
        nomuts[i] = prob_branch(d, m.mu, false);
        treelen += d;
    }
    const double inv_lk = .25 * exp(-m.mu * fmax(treelen, mintime));
    std::vector<int> order(N);
    postorder(t, order.data());
    std::vector<double> inner(4 * N), outer(4 * N);
    std::vector<char> bases(N);
    const int nleaves = (N + 1) / 2;
    for (int i = 0; i < seqlen; i++) {
        if (invariant[i]) {
            emit[i][statei] = (rows[0][i] == 'N') ? 1.0 : inv_lk;
            continue;
        }
        for (int x = 0; x < nleaves; x++) bases[x] = rows[x][i];
        pruning(t, order.data(), bases.data(), muts.data(), nomuts.data(), inner.data(),
                outer.data());
        double lk = 0.0;
        for (int a = 0; a < 4; a++) lk += .25 * inner[4 * t.root + a];
        emit[i][statei] = lk;
    }
}

namespace {

// rows allocated one by one (new_matrix / delete_matrix, common.h:229-247)
double **new_rows(int nrows, int ncols)
{
    double **m = new double *[nrows > 0 ? nrows : 1];
    for (int i = 0; i < nrows; i++) m[i] = new double[ncols > 0 ? ncols : 1];
    return m;
}

void delete_rows(double **m, int nrows)
{
    if (!m) return;
    for (int i = 0; i < nrows; i++) delete[] m[i];
    delete[] m;
}

// A local tree built from a parent array (LocalTree(ptree, nnodes, ages),
// local_tree.cpp:321-350) with room for a threaded branch.
struct HostTree {
    LTree t;
    HostTree(const int *ptree, const int *ages, int nnodes)
    {
        tree_from_parents(t, ptree, ages, nnodes, nnodes + 2);
    }
};

// per-node index of a node-major state list (NodeStateLookup, states.h:53-131)
void node_index(const LTree &t, int T, NodeIndex &ni)
{
    int s = 0;
    for (int i = 0; i < t.nnodes; i++) {
        const int p = t.nodes[i].parent;
        const int top = (p == -1) ? T - 2 : t.nodes[p].age;
        ni.first[i] = s;
        ni.hi[i] = top;
        if (top >= t.nodes[i].age) s += top - t.nodes[i].age + 1;
    }
}

Model host_model(const double *times, int ntimes, const double *popsizes, double rho, double mu)
{
    std::vector<double> pops(ntimes, 1.0);
    if (popsizes)
        for (int i = 0; i < ntimes; i++) pops[i] = popsizes[i];
    Model m;
    model_init(m, ntimes, times, pops.data(), rho, mu);
    return m;
}

bool grid_ok(const char *who, int ntimes)
{
    if (ntimes >= 3 && ntimes <= MAXT) return true;
    fprintf(stderr, "argweaver_amd: %s: ntimes %d outside 3..%d\n", who, ntimes, MAXT);
    return false;
}

// The switch matrix between two neighbouring trees in compressed form
// (TransMatrixSwitch, trans.h:162-235; calc_transition_probs_switch,
// trans.cpp:538-739, external threading).
struct SwitchMat {
    std::vector<int> determ;
    std::vector<double> determprob, recoalrow, recombrow;
    int recoalsrc = -1, recombsrc = -1;

    double get(int i, int j) const
    {
        if (i == recoalsrc) return recoalrow[j];
        if (i == recombsrc) return recombrow[j];
        return determ[i] == j ? determprob[i] : 0.0;
    }
};

void switch_matrix(const LTree &last_t, const LTree &t, const Spr &spr, const int *mapping,
                   const Model &m, SwitchMat &sw)
{
    const int T = m.T;
    FlatTree lf(last_t), nf(t);
    const TreeView &pt = lf.view, &tv = nf.view;
    Lineages PL;
    count_lineages(pt, T, PL);
    NodeIndex pni, ni;
    node_index(last_t, T, pni);
    node_index(t, T, ni);
    std::vector<int> n1s, t1s, n2s, t2s;
    tree_states(last_t, T, n1s, t1s);
    tree_states(t, T, n2s, t2s);
    const int S1 = (int)n1s.size(), S2 = (int)n2s.size();

    SwitchCtx cx;
    switch_ctx_sums(pt, m, PL, spr, tree_length(pt, m.times), cx);
    for (int a = 0; a < T; a++) cx.recoals[a] = sw_recoal(m, PL, spr.rt, spr.ct, a, cx.rpa);

    sw.determ.assign(S1 > 0 ? S1 : 1, -1);
    sw.determprob.assign(S1 > 0 ? S1 : 1, 0.0);
    sw.recoalrow.assign(S2 > 0 ? S2 : 1, 0.0);
    sw.recombrow.assign(S2 > 0 ? S2 : 1, 0.0);
    for (int i = 0; i < S1; i++) {
        const int tgt = sw_determ(pt, tv, ni, spr, mapping, n1s[i], t1s[i]);
        sw.determ[i] = tgt;
        if (tgt >= 0) sw.determprob[i] = sw_determprob(pt, m, PL, spr, cx, n1s[i], t1s[i]);
    }
    sw.recombsrc = lookup(pt, pni, spr.rn, spr.rt);
    sw.recoalsrc = lookup(pt, pni, spr.cn, spr.ct);

    if (sw.recombsrc != -1) {
        // get_recomb_transition_switch (trans.cpp:278-313): stay and escape
        const int parent = pt.parent[spr.rn];
        const int time2 = pt.age[parent];
        const int other = pt.sibling(spr.rn);
        const int node2 = (other == spr.cn) ? tv.parent[mapping[other]] : mapping[other];
        const int stay = lookup(tv, ni, mapping[spr.rn], spr.rt);
        const int esc = lookup(tv, ni, node2, time2);
        if (stay != -1)
            sw.recombrow[stay] =
                sw_recomb_recoal(pt, m, PL, spr, spr.ct, spr.rn, spr.rt, cx.rpa, cx.last_len);
        if (esc != -1)
            sw.recombrow[esc] =
                sw_recomb_recoal(pt, m, PL, spr, spr.ct, spr.rn, spr.rt, spr.rt, cx.last_len);
    }
    if (sw.recoalsrc != -1) {
        const int node1 = spr.cn, time1 = spr.ct;
        const int last_parent = pt.parent[spr.rn];
        int node3;
        if (last_parent == node1) {
            const int other = (pt.c1[last_parent] == spr.rn) ? pt.c0[last_parent] : pt.c1[last_parent];
            node3 = mapping[other];
        } else {
            node3 = mapping[node1];
        }
        const int mrn = mapping[spr.rn];
        const int rparent = tv.parent[mrn];
        for (int k = 0; k < S2; k++) {
            const int n2 = n2s[k], t2 = t2s[k];
            if (!((n2 == mrn && t2 >= spr.rt) || (n2 == node3 && t2 == time1) ||
                  (n2 == rparent && t2 == time1)))
                continue;
            sw.recoalrow[k] =
                sw_recomb_recoal(pt, m, PL, spr, t2, spr.cn, spr.ct, cx.rpa, cx.last_len);
        }
    }
}

// JC69 emission table of one tree for the new sequence (calc_emissions,
// emit.cpp:650-869, external threading): rows[0..nleaves-1] are the leaves'
// sequences, rows[nleaves] the new one, rows[nleaves+1..] only take part in
// the invariant test.  emit[i][j] for column i and state j.
void emission_table(const LTree &lt, const Model &m, const char *const *rows, int nrows,
                    int seqlen, const std::vector<int> &snode, const std::vector<int> &stime,
                    double **emit)
{
    FlatTree ft(lt);
    const TreeView &t = ft.view;
    const int N = t.n, S = (int)snode.size(), newleaf = (N + 1) / 2;
    const double mintime = m.times[1] * .1;
    std::vector<double> muts(N, 0.0), nomuts(N, 0.0);
    double mainlen = 0.0;
    for (int i = 0; i < N; i++) {
        if (i == t.root) continue;
        const double d = fmax(m.times[t.age[t.parent[i]]] - m.times[t.age[i]], mintime);
        muts[i] = prob_branch(d, m.mu, true);
        nomuts[i] = prob_branch(d, m.mu, false);
        mainlen += d;
    }
    std::vector<EmitTerms> terms(S);
    for (int j = 0; j < S; j++) {
        EmitTime et;
        emit_time_terms(t, m, mainlen, stime[j], et);
        emit_terms(t, m, et, snode[j], stime[j], terms[j]);
    }
    std::vector<int> order(N);
    postorder(t, order.data());
    std::vector<double> inner(4 * N), outer(4 * N);
    std::vector<char> bases(N);
    for (int i = 0; i < seqlen; i++) {
        bool invariant = true;
        const char c0 = rows[0][i];
        for (int x = 1; x < nrows; x++)
            if (rows[x][i] != c0) {
                invariant = false;
                break;
            }
        if (invariant) {
            for (int j = 0; j < S; j++) emit[i][j] = (c0 == 'N') ? 1.0 : terms[j].inv;
            continue;
        }
        for (int x = 0; x < newleaf && x < N; x++) bases[x] = rows[x][i];
        pruning(t, order.data(), bases.data(), muts.data(), nomuts.data(), inner.data(),
                outer.data());
        double in2[4];
        leaf_partial(rows[newleaf][i], in2);
        for (int j = 0; j < S; j++) {
            const int n = snode[j];
            emit[i][j] = emit_variant(in2, &inner[4 * n], &outer[4 * n], n == t.root,
                                      terms[j].mut, terms[j].nomut);
        }
    }
}

// assert_emissions (emit.cpp:1222-1250) on one block
bool check_emissions(const LTree &lt, const Model &m, const char *const *rows, int nrows,
                     int seqlen)
{
    std::vector<int> sn, st;
    tree_states(lt, m.T, sn, st);
    const int S = (int)sn.size();
    double **e1 = new_rows(seqlen, S), **e2 = new_rows(seqlen, S);
    emission_table(lt, m, rows, nrows, seqlen, sn, st, e1);
    std::vector<char> inv(seqlen);
    for (int i = 0; i < seqlen; i++) {
        bool same = true;
        for (int x = 1; x < nrows; x++)
            if (rows[x][i] != rows[0][i]) {
                same = false;
                break;
            }
        inv[i] = same;
    }
    LTree t2 = lt;
    t2.ensure_capacity(lt.nnodes + 2);
    const int newleaf = lt.num_leaves();
    bool ok = true;
    for (int j = 0; j < S && ok; j++) {
        add_tree_branch(t2, sn[j], st[j]);
        emission_column_slow(t2, m, rows, seqlen, inv, j, e2);
        remove_tree_branch(t2, newleaf, nullptr);
    }
    for (int j = 0; j < S && ok; j++)
        for (int i = 0; i < seqlen && ok; i++)
            if (!close_enough(e1[i][j], e2[i][j], .0001, 1e-12)) ok = false;
    delete_rows(e1, seqlen);
    delete_rows(e2, seqlen);
    return ok;
}

// 1 - exp(-max(rho * treelen, rho)) (calc_recomb_prob, trans.cpp:846-853)
double recomb_prob(const LTree &t, const Model &m)
{
    FlatTree ft(t);
    const double len = tree_length(ft.view, m.times);
    return 1.0 - exp(-fmax(m.rho * len, m.rho));
}

// log-sum-exp with the reference's cut-off (logsum, common.h:182-200)
double log_sum(const double *v, int n)
{
    if (n == 0) return 1.0;
    double mx = v[0];
    for (int i = 1; i < n; i++)
        if (v[i] > mx) mx = v[i];
    double s = 0.0;
    for (int i = 0; i < n; i++)
        if (v[i] - mx > -15) s += exp(v[i] - mx);
    return mx + log(s);
}

int sample_weights(const double *w, int n)
{
    double total = 0.0;
    for (int i = 0; i < n; i++) total += w[i];
    const double pick = rand() / double(RAND_MAX) * total;
    double x = 0.0;
    for (int i = 0; i < n; i++) {
        x += w[i];
        if (x >= pick) return i;
    }
    return n - 1;
}

}  // namespace
}  // namespace aw

using namespace aw;

extern "C" {

double **new_transition_probs(int nnodes, int *ptree, int *ages, double treelen, intstate *istates,
                              int nstates, int ntimes, double *times, double *time_steps,
                              int *nbranches, int *nrecombs, int *ncoals, double *popsizes,
                              double rho)
{
    // the tree length, time steps and lineage counts are recomputed from the
    // tree and the time points, as the reference does (trans.cpp:1190-1208)
    (void)treelen; (void)time_steps; (void)nbranches; (void)nrecombs; (void)ncoals;
    if (!grid_ok("new_transition_probs", ntimes)) return nullptr;
    const Model m = host_model(times, ntimes, popsizes, rho, 0.0);
    HostTree h(ptree, ages, nnodes);
    FlatTree ft(h.t);
    Lineages L;
    count_lineages(ft.view, ntimes, L);
    TransVec tv;
    calc_trans(ft.view, m, L, tv);
    double **out = new_rows(nstates, nstates);
    for (int i = 0; i < nstates; i++)
        for (int j = 0; j < nstates; j++) {
            const int a = istates[i][1], b = istates[j][1];
            const int c = ages[istates[j][0]];
            out[i][j] = log(tv.get_time(a, b, c, istates[i][0] == istates[j][0]));
        }
    return out;
}

double **new_transition_probs_switch(int *ptree, int *last_ptree, int nnodes, int recomb_name,
                                     int recomb_time, int coal_name, int coal_time,
                                     int *ages_index, int *last_ages_index, double treelen,
                                     double last_treelen, intstate *istates1, int nstates1,
                                     intstate *istates2, int nstates2, int ntimes, double *times,
                                     double *time_steps, int *nbranches, int *nrecombs,
                                     int *ncoals, double *popsizes, double rho)
{
    (void)treelen; (void)last_treelen; (void)time_steps; (void)nbranches; (void)nrecombs;
    (void)ncoals; (void)istates1; (void)istates2;
    if (!grid_ok("new_transition_probs_switch", ntimes)) return nullptr;
    const Model m = host_model(times, ntimes, popsizes, rho, 0.0);
    HostTree h(ptree, ages_index, nnodes), hl(last_ptree, last_ages_index, nnodes);
    // make_node_mapping (local_tree.h:767-776): the recombination's broken
    // parent has no image
    std::vector<int> mapping(nnodes);
    for (int j = 0; j < nnodes; j++) mapping[j] = j;
    mapping[last_ptree[recomb_name]] = -1;
    const Spr spr{recomb_name, recomb_time, coal_name, coal_time};
    SwitchMat sw;
    switch_matrix(hl.t, h.t, spr, mapping.data(), m, sw);
    double **out = new_rows(nstates1, nstates2);
    for (int i = 0; i < nstates1; i++)
        for (int j = 0; j < nstates2; j++) out[i][j] = log(sw.get(i, j));
    return out;
}

void delete_transition_probs(double **transmat, int nstates) { delete_rows(transmat, nstates); }

bool arghmm_assert_transmat(int nnodes, int *ptree, int *ages, int ntimes, double *times,
                            double *popsizes, double rho)
{
    // assert_transmat (trans.cpp:856-943): every entry against the sum over
    // the SPRs that lead from state i to state j, computed on the tree with
    // the new branch attached
    if (!grid_ok("arghmm_assert_transmat", ntimes)) return false;
    const Model m = host_model(times, ntimes, popsizes, rho, 0.0);
    HostTree h(ptree, ages, nnodes);
    const LTree &tree = h.t;
    FlatTree ft(tree);
    Lineages L;
    count_lineages(ft.view, ntimes, L);
    TransVec tv;
    calc_trans(ft.view, m, L, tv);
    std::vector<int> sn, st;
    tree_states(tree, ntimes, sn, st);
    const int S = (int)sn.size();
    const int newleaf = tree.num_leaves();
    LTree t2 = tree;
    t2.ensure_capacity(nnodes + 2);
    for (int i = 0; i < S; i++) {
        for (int j = 0; j < S; j++) {
            const int sister_age = t2.nodes[sn[i]].age;
            add_tree_branch(t2, sn[i], st[i]);
            FlatTree f2(t2);
            const double len2 = tree_length(f2.view, m.times);
            double p = -INFINITY;
            Spr spr{newleaf, 0, sn[j], st[j]};
            for (int r = 0; r <= std::min(st[i], st[j]); r++) {
                spr.rt = r;
                p = logadd(p, spr_logprob(f2.view, m, spr, len2));
            }
            if (sn[i] == sn[j]) {
                spr.rn = sn[i];
                spr.cn = (st[j] < st[i]) ? newleaf : t2.nodes[newleaf].parent;
                spr.ct = st[j];
                for (int r = sister_age; r <= std::min(st[i], st[j]); r++) {
                    spr.rt = r;
                    p = logadd(p, spr_logprob(f2.view, m, spr, len2));
                }
            }
            const double rp = recomb_prob(t2, m);
            p += log(rp);
            if (i == j) p = logadd(p, log(1.0 - rp));
            const double p2 = log(tv.get_time(st[i], st[j], ages[sn[j]], sn[i] == sn[j]));
            remove_tree_branch(t2, newleaf, nullptr);
            if (!close_enough(p, p2, 1e-4, 1e-9)) return false;
        }
    }
    return true;
}

bool arghmm_assert_transmat_switch(int nnodes, int *ptree, int *ages, int recomb_node,
                                   int recomb_time, int coal_node, int coal_time, int ntimes,
                                   double *times, double *popsizes, double rho)
{
    // assert_transmat_switch (trans.cpp:947-1032): every nonzero entry
    // against the probability of the SPR it implies on the grown trees
    if (!grid_ok("arghmm_assert_transmat_switch", ntimes)) return false;
    const Model m = host_model(times, ntimes, popsizes, rho, 0.0);
    HostTree hl(ptree, ages, nnodes);
    LTree t1 = hl.t;
    LTree t2 = hl.t;
    const Spr spr0{recomb_node, recomb_time, coal_node, coal_time};
    apply_spr(t2, spr0);
    std::vector<int> mapping(nnodes + 2);
    for (int i = 0; i < nnodes; i++) mapping[i] = i;
    mapping[t1.nodes[recomb_node].parent] = -1;
    SwitchMat sw;
    switch_matrix(t1, t2, spr0, mapping.data(), m, sw);
    std::vector<int> n1, a1, n2, a2;
    tree_states(t1, ntimes, n1, a1);
    tree_states(t2, ntimes, n2, a2);
    const int newleaf = t1.num_leaves(), displaced = nnodes, newcoal = nnodes + 1;
    t1.ensure_capacity(nnodes + 2);
    t2.ensure_capacity(nnodes + 2);
    for (int i = 0; i < (int)n1.size(); i++) {
        for (int j = 0; j < (int)n2.size(); j++) {
            const double p2 = log(sw.get(i, j));
            if (p2 == -INFINITY) continue;
            Spr spr = spr0;
            std::vector<int> map2(mapping);
            add_tree_branch(t1, n1[i], a1[i]);
            add_tree_branch(t2, n2[j], a2[j]);
            add_spr_branch(t2, t1, n2[j], a2[j], n1[i], spr, map2.data(), newleaf, displaced,
                           newcoal);
            FlatTree f1(t1);
            double p = log(recomb_prob(t1, m));
            p += spr_logprob(f1.view, m, spr, tree_length(f1.view, m.times));
            remove_tree_branch(t1, newleaf, nullptr);
            remove_tree_branch(t2, newleaf, nullptr);
            if (!close_enough(p, p2, 1e-4, 1e-9)) return false;
        }
    }
    return true;
}

double **new_emissions(intstate *istates, int nstates, int *ptree, int nnodes, int *ages_index,
                       char **seqs, int nseqs, int seqlen, double *times, int ntimes, double mu)
{
    if (!grid_ok("new_emissions", ntimes)) return nullptr;
    const Model m = host_model(times, ntimes, nullptr, 0.0, mu);
    HostTree h(ptree, ages_index, nnodes);
    std::vector<int> sn(nstates), st(nstates);
    for (int j = 0; j < nstates; j++) {
        sn[j] = istates[j][0];
        st[j] = istates[j][1];
    }
    double **emit = new_rows(seqlen, nstates);
    emission_table(h.t, m, seqs, nseqs, seqlen, sn, st, emit);
    return emit;
}

void delete_emissions(double **emit, int seqlen) { delete_rows(emit, seqlen); }

bool arghmm_assert_emit(LocalTrees *trees, int ntimes, double *times, double mu, char **seqs,
                        int nseqs, int seqlen)
{
    (void)seqlen;
    if (!grid_ok("arghmm_assert_emit", ntimes)) return false;
    const Model m = host_model(times, ntimes, nullptr, 0.0, mu);
    const ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    std::vector<const char *> rows(nseqs);
    int end = tr.start_coord;
    for (const LBlock &B : tr.blocks) {
        const int start = end;
        end = start + B.blocklen;
        for (int j = 0; j < nseqs; j++) rows[j] = seqs[j] + start;
        if (!check_emissions(B.tree, m, rows.data(), nseqs, B.blocklen)) return false;
    }
    return true;
}

// hmm.cpp:27-99: log-space forward, backward and posterior sampling over a
// dense transition matrix; the first forward column and the last backward
// column (and the last path state) are the caller's
void forward_alg(int n, int nstates, double **trans, double **emit, double **fw)
{
    std::vector<double> v(nstates);
    for (int i = 1; i < n; i++)
        for (int k = 0; k < nstates; k++) {
            for (int j = 0; j < nstates; j++) v[j] = fw[i - 1][j] + trans[j][k];
            fw[i][k] = log_sum(v.data(), nstates) + emit[i][k];
        }
}

void backward_alg(int n, int nstates, double **trans, double **emit, double **bw)
{
    std::vector<double> v(nstates);
    for (int i = n - 2; i > -1; i--)
        for (int j = 0; j < nstates; j++) {
            for (int k = 0; k < nstates; k++) v[k] = trans[j][k] + bw[i + 1][k] + emit[i + 1][k];
            bw[i][j] = log_sum(v.data(), nstates);
        }
}

void sample_hmm_posterior(int n, int nstates, double **trans, double **fw, int *path)
{
    std::vector<double> A(nstates);
    for (int i = n - 2; i >= 0; i--) {
        const int k = path[i + 1];
        for (int j = 0; j < nstates; j++) A[j] = fw[i][j] + trans[j][k];
        const double total = log_sum(A.data(), nstates);
        for (int j = 0; j < nstates; j++) A[j] = exp(A[j] - total);
        path[i] = sample_weights(A.data(), nstates);
    }
}

// est_popsize_trees (est_popsize.cpp:68-135): one tree every `step` bases,
// then per time interval 0.5 * step_j * pairs / coalescences
void arghmm_est_popsizes_trees(LocalTrees *trees, double *times, int ntimes, int step,
                               double *popsizes)
{
    if (!grid_ok("arghmm_est_popsizes_trees", ntimes) || step <= 0) return;
    const Model m = host_model(times, ntimes, nullptr, 0.0, 0.0);
    const ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    std::vector<long long> ncoal(ntimes, 0), npairs(ntimes, 0);
    const int nleaves = tr.num_leaves();
    int end = tr.start_coord, pos = end;
    for (const LBlock &B : tr.blocks) {
        const int start = end;
        end += B.blocklen;
        while (start <= pos && pos < end) {
            FlatTree ft(B.tree);
            Lineages L;
            count_lineages(ft.view, ntimes, L);
            for (int j = 0; j < ntimes - 1; j++) {
                const int a = (j == 0) ? nleaves : L.nb[j - 1];
                const int b = L.nb[j];
                ncoal[j] += a - b;
                npairs[j] += a * (a - 1) / 2;
            }
            pos += step;
        }
    }
    for (int j = 0; j < ntimes - 1; j++)
        popsizes[j] = ncoal[j] == 0 ? 0.0 : .5 * m.steps[j] * double(npairs[j]) / double(ncoal[j]);
}

}  // extern "C"
