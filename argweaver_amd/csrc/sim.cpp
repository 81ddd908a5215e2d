// sim.cpp -- seeded synthetic threading workloads.
//
// A discrete sequentially-Markov coalescent (DSMC) simulator following the
// reference's Python simulator (argweaver/sim.py:31-294): an initial
// coalescent tree under piecewise-constant population sizes, discretised to
// the time grid; then recombinations along the sequence at rate
// max(treelen*rho, rho); recombination time ~ nbranches*time_step, branch
// uniform; re-coalescence walks up the half-interval grid; coalescing branch
// uniform among the states at that time (excluding the recombining lineage
// and the broken node).  Mutations follow sim.py:297-439 (Poisson on branches
// in proportion to length, random ancestral and derived bases).
//
// The simulator works directly in compressed columns: rho and mu are per
// column (the caller passes rho*c and mu*c, as arg-sample does after
// compression, arg-sample.cpp:428-439).  Trees are produced in the
// LocalTrees convention (leaves first, SPR applied by reusing the broken
// node), so they can be handed straight to the threading path.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "local_trees.h"

namespace aw {
namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        for (int i = 0; i < 4; i++) {   // splitmix64
            seed += 0x9e3779b97f4a7c15ULL;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()   // xoshiro256**
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
    double expo(double rate) { return -log(1.0 - uniform()) / rate; }
    int below(int n) { return (int)(uniform() * n); }
};

// nearest grid point among times[0..T-2] ("closer" rounding of
// discretize_arg with ignore_top, argweaver/__init__.py:210-240)
int discretize(double age, const double *times, int T)
{
    int last = T - 2;
    if (age <= times[0]) return 0;
    if (age >= times[last]) return last;
    int j = 1;
    while (times[j] < age) j++;
    int i = j - 1;
    if (times[j] == age) return j;
    return (age - times[i] < times[j] - age) ? i : j;
}

// sample_tree (sim.py:31-64) then a random topology with those coal times
void initial_tree(int k, const double *times, const double *pop2, int T, Rng &rng, LTree &t)
{
    std::vector<double> coal_times;
    int timei = 0;
    double n = pop2[0];
    double tt = 0.0;
    int k2 = k;
    while (k2 > 1) {
        double rate = (k2 * (k2 - 1) / 2) / n;
        double t2 = rng.expo(rate);
        if (timei < T - 2 && tt + t2 > times[timei + 1]) {
            timei++;
            tt = times[timei];
            n = pop2[timei];
            continue;
        }
        tt += t2;
        coal_times.push_back(tt);
        k2--;
    }
    const int nnodes = 2 * k - 1;
    std::vector<int> parent(nnodes, -1), age(nnodes, 0);
    std::vector<int> lineages(k);
    for (int i = 0; i < k; i++) lineages[i] = i;
    int next = k;
    for (double ct : coal_times) {
        int a = rng.below((int)lineages.size());
        int x = lineages[a];
        lineages.erase(lineages.begin() + a);
        int b = rng.below((int)lineages.size());
        int y = lineages[b];
        lineages.erase(lineages.begin() + b);
        parent[x] = parent[y] = next;
        age[next] = discretize(ct, times, T);
        lineages.push_back(next);
        next++;
    }
    tree_from_parents(t, parent.data(), age.data(), nnodes, nnodes + 2);
}

double branch_length_sum(const LTree &t, const double *times)
{
    double len = 0.0;
    for (int i = 0; i < t.nnodes; i++) {
        int p = t.nodes[i].parent;
        if (p != -1) len += times[t.nodes[p].age] - times[t.nodes[i].age];
    }
    return len;
}

}  // namespace

// Simulate k haplotypes over seqlen columns.  Returns the ARG (k leaves) and
// fills seqs[k*seqlen] row-major.
ArgTrees *simulate_arg(int k, int seqlen, const double *times, int T, const double *popsizes,
                       double rho, double mu, uint64_t seed, char *seqs)
{
    Rng rng(seed);
    // sim.py's popsize is the haploid size 2N that the C model writes as 2.0*popsizes
    std::vector<double> pop2(T);
    for (int i = 0; i < T; i++) pop2[i] = 2.0 * popsizes[i];
    std::vector<double> tsteps(T);
    for (int i = 0; i < T - 1; i++) tsteps[i] = times[i + 1] - times[i];
    // midpoints of the half-interval grid (sim.py get_coal_times)
    std::vector<double> times2(2 * T);
    for (int i = 0; i < T - 1; i++) {
        times2[2 * i] = times[i];
        times2[2 * i + 1] = sqrt((times[i + 1] + 1) * (times[i] + 1));
    }
    times2[2 * T - 2] = times[T - 1];

    ArgTrees *tr = new ArgTrees();
    tr->start_coord = 0;
    tr->end_coord = seqlen;
    tr->nnodes = 2 * k - 1;
    tr->seqids.resize(k);
    for (int i = 0; i < k; i++) tr->seqids[i] = i;

    LBlock first;
    initial_tree(k, times, pop2.data(), T, rng, first.tree);
    first.spr = Spr{-1, -1, -1, -1};
    tr->blocks.push_back(first);

    const int N = tr->nnodes;
    int pos = 0, block_start = 0;
    std::vector<int> nb(T), snode, stime;
    std::vector<char> excl(N);
    std::vector<int> cand;
    while (true) {
        LTree &t = tr->blocks.back().tree;
        double len = branch_length_sum(t, times);
        int blocklen = 0;
        while (blocklen < 1) blocklen = (int)rng.expo(std::max(len * rho, rho));
        if (pos + blocklen >= seqlen - 1) {
            tr->blocks.back().blocklen = seqlen - block_start;
            break;
        }
        pos += blocklen;

        // lineage counts over the states (get_nlineages_recomb_coal)
        const int ri = t.nodes[t.root].age;
        tree_states(t, T, snode, stime);
        std::fill(nb.begin(), nb.end(), 0);
        for (size_t s = 0; s < snode.size(); s++) {
            int p = t.nodes[snode[s]].parent;
            if (p == -1 || stime[s] < t.nodes[p].age) nb[stime[s]]++;
        }
        nb[T - 1] = 1;

        // recombination time and branch
        double tot = 0.0;
        for (int i = 0; i <= ri; i++) tot += nb[i] * tsteps[i];
        double pick = rng.uniform() * tot, acc = 0.0;
        int rt = ri;
        for (int i = 0; i <= ri; i++) {
            acc += nb[i] * tsteps[i];
            if (pick < acc) { rt = i; break; }
        }
        cand.clear();
        for (size_t s = 0; s < snode.size(); s++)
            if (stime[s] == rt && snode[s] != t.root) cand.push_back(snode[s]);
        if (cand.empty()) continue;   // only the root branch at rt: no recombination
        int rn = cand[rng.below((int)cand.size())];

        // re-coalescence time
        int j = rt;
        int last_kj = nb[std::max(j - 1, 0)];
        const int rparent_age = t.nodes[t.nodes[rn].parent].age;
        while (j < T - 2) {
            int kj = nb[j];
            if (t.nodes[rn].age <= j && j <= rparent_age && rparent_age > j) kj -= 1;
            double A = (times2[2 * j + 1] - times2[2 * j]) * kj;
            if (j > rt) A += (times2[2 * j] - times2[2 * j - 1]) * last_kj;
            double coal_prob = 1.0 - exp(-A / pop2[j]);
            if (rng.uniform() < coal_prob) break;
            j++;
            last_kj = kj;
        }
        int ct = j;

        // coalescing branch: states at ct outside the recombining subtree's
        // points at ct and not the broken node's own point
        std::fill(excl.begin(), excl.end(), 0);
        {
            std::vector<int> stack{rn};
            while (!stack.empty()) {
                int x = stack.back();
                stack.pop_back();
                excl[x] = 1;
                if (t.nodes[x].age == ct && t.nodes[x].child[0] != -1) {
                    stack.push_back(t.nodes[x].child[0]);
                    stack.push_back(t.nodes[x].child[1]);
                }
            }
        }
        int broken = t.nodes[rn].parent;
        cand.clear();
        for (size_t s = 0; s < snode.size(); s++) {
            if (stime[s] != ct || excl[snode[s]]) continue;
            if (snode[s] == broken && stime[s] == t.nodes[broken].age) continue;
            cand.push_back(snode[s]);
        }
        if (cand.empty()) continue;
        int cn = cand[rng.below((int)cand.size())];

        tr->blocks.back().blocklen = pos - block_start;
        block_start = pos;
        LBlock nbk;
        nbk.tree = t;
        Spr spr{rn, rt, cn, ct};
        apply_spr(nbk.tree, spr);
        nbk.spr = spr;
        nbk.mapping.resize(N);
        for (int x = 0; x < N; x++) nbk.mapping[x] = x;
        nbk.mapping[broken] = -1;
        tr->blocks.push_back(std::move(nbk));
    }

    // mutations (sim.py:297-323) and alignment (sim.py:376-439)
    memset(seqs, 'A', (size_t)k * seqlen);
    std::vector<char> leafset;
    int start = 0;
    for (const LBlock &B : tr->blocks) {
        const LTree &t = B.tree;
        for (int i = 0; i < t.nnodes; i++) {
            int p = t.nodes[i].parent;
            if (p == -1) continue;
            double blen = times[t.nodes[p].age] - times[t.nodes[i].age];
            if (blen <= 0.0) continue;
            double rate = mu * blen;
            double x = start + rng.expo(rate);
            while (x < start + B.blocklen) {
                int col = (int)x;
                // derived base on the leaves below node i
                char anc = "ACGT"[rng.below(4)];
                bool fresh = true;
                for (int l = 0; l < k; l++)
                    if (seqs[(size_t)l * seqlen + col] != 'A') fresh = false;
                if (fresh)
                    for (int l = 0; l < k; l++) seqs[(size_t)l * seqlen + col] = anc;
                std::vector<int> stack{i};
                leafset.assign(k, 0);
                while (!stack.empty()) {
                    int y = stack.back();
                    stack.pop_back();
                    if (t.nodes[y].child[0] == -1) leafset[y] = 1;
                    else {
                        stack.push_back(t.nodes[y].child[0]);
                        stack.push_back(t.nodes[y].child[1]);
                    }
                }
                int leaf0 = -1;
                for (int l = 0; l < k; l++) if (leafset[l]) { leaf0 = l; break; }
                char cur = seqs[(size_t)leaf0 * seqlen + col];
                char der;
                do { der = "ACGT"[rng.below(4)]; } while (der == cur);
                for (int l = 0; l < k; l++)
                    if (leafset[l]) seqs[(size_t)l * seqlen + col] = der;
                x += rng.expo(rate);
            }
        }
        start += B.blocklen;
    }
    if (!check_trees(*tr)) {
        delete tr;
        return nullptr;
    }
    return tr;
}

}  // namespace aw
