// capi.cpp -- the C ABI: the reference's arghmm_* entry points for the
// threading path, backed by the gfx950 engine, plus the handle/support
// symbols callers use around them.  Host work here is what the reference
// also does on the host between HMM passes: flattening the ARG, drawing
// the traceback's uniforms from the process-global glibc rand() stream,
// sampling recombination points and splicing the thread into the ARG.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/argweaver_amd.h"
#include "engine.h"
#include "serve.h"
#include "host_util.h"
#include "local_trees.h"
#include "thread_path.h"
#include "internal_prep.h"
#include "region.h"

namespace aw {

static int g_log_level = 0;

// frand / expovariate (common.h:103-125) on the process-global glibc stream
static inline double frand() { return rand() / double(RAND_MAX); }

// sample() (common.h:272-290)
static int sample_weights(const double *w, int n)
{
    double total = 0.0;
    for (int i = 0; i < n; i++) total += w[i];
    double pick = frand() * total;
    double x = 0.0;
    for (int i = 0; i < n; i++) {
        x += w[i];
        if (x >= pick) return i;
    }
    return n - 1;
}

// nonzero (with a message) when the time grid does not fit the model tables
static int build_instance(const ArgTrees &tr, const double *times, int ntimes,
                          const double *popsizes, double rho, double mu, char **seqs,
                          int new_chrom, HostInstance &hi)
{
    if (ntimes < 2 || ntimes > MAXT) {
        fprintf(stderr, "argweaver_amd: %d time points (supported: 2..%d)\n", ntimes, MAXT);
        return 1;
    }
    flatten(tr, hi.ft);
    const int nleaves = tr.num_leaves();
    hi.rows.resize(nleaves + 1);
    for (int i = 0; i < nleaves; i++) hi.rows[i] = seqs[tr.seqids[i]] + tr.start_coord;
    hi.rows[nleaves] = seqs[new_chrom] + tr.start_coord;
    std::vector<double> pop(ntimes);
    for (int i = 0; i < ntimes; i++) pop[i] = popsizes ? popsizes[i] : 0.0;
    model_init(hi.model, ntimes, times, pop.data(), rho, mu);
    return 0;
}

// The per-block inputs of the recombination sampler: the tree as arrays,
// its lineage counts and its states' (node, time) by state index.
struct RecombTree {
    std::vector<int> par, c0, c1, age, off;   // off[node]: first state of node
    Lineages L;
    TreeView view() const
    {
        int root = 0;
        while (par[root] != -1) root++;
        return TreeView{(int)par.size(), root, par.data(), c0.data(), c1.data(), age.data()};
    }
    void build(const LTree &lt, int T)
    {
        const int N = lt.nnodes;
        par.resize(N); c0.resize(N); c1.resize(N); age.resize(N); off.resize(N + 1);
        off[0] = 0;
        for (int i = 0; i < N; i++) {
            par[i] = lt.nodes[i].parent;
            c0[i] = lt.nodes[i].child[0];
            c1[i] = lt.nodes[i].child[1];
            age[i] = lt.nodes[i].age;
        }
        for (int i = 0; i < N; i++) {
            const int top = par[i] == -1 ? T - 2 : age[par[i]];
            off[i + 1] = off[i] + std::max(0, top - age[i] + 1);
        }
        lineages(T);
    }
    // count_lineages (local_tree.cpp:34-69) by a difference array over time:
    // the same counts in O(nodes + times)
    void lineages(int T)
    {
        int d[MAXT + 1] = {0};
        for (int j = 0; j < T; j++) L.nb[j] = L.nr[j] = L.nc[j] = 0;
        for (size_t i = 0; i < par.size(); i++) {
            const int p = par[i], page = p == -1 ? T - 2 : age[p];
            if (age[i] < page) {
                d[age[i]]++;
                d[page]--;
            }
            L.nr[page]++;
            L.nc[page]++;
            if (p == -1) L.nb[page]++;
        }
        for (int j = 0, run = 0; j < T; j++) {
            run += d[j];
            L.nb[j] += run;
            L.nr[j] += run;
            L.nc[j] += run;
        }
        L.nb[T - 1] = 1;
    }
    void state(int s, int &node, int &time) const
    {
        node = (int)(std::upper_bound(off.begin(), off.end(), s) - off.begin()) - 1;
        time = age[node] + (s - off[node]);
    }
};

// sample_recombinations (recomb.cpp:151-235) for external threading.  The
// draws are sequential (one glibc rand() stream), but what they need per
// column is fixed by the path: the rate of the waiting-time draw at every
// column whose state repeats the previous one.  Those rates (a TransMatrix
// per block) are computed first, over the blocks in parallel; the draws
// then run in the reference's order, building a block's tree arrays only
// where a recombination is placed.
static void sample_recombinations(const ArgTrees &tr, const Model &m, const int *path,
                                  std::vector<int> &rpos, std::vector<int> &rnode,
                                  std::vector<int> &rtime)
{
    const int T = m.T, nb = (int)tr.blocks.size(), sc = tr.start_coord;
    std::vector<int> bstart(nb + 1, sc);
    for (int b = 0; b < nb; b++) bstart[b + 1] = bstart[b] + tr.blocks[b].blocklen;
    std::vector<double> rate((size_t)tr.length(), 0.0);
    auto rates = [&](int b0, int b1) {
        RecombTree rt;
        TransVec tv;
        for (int b = b0; b < b1; b++) {
            const int start = bstart[b], end = bstart[b + 1];
            if (end - start < 2) continue;
            rt.build(tr.blocks[b].tree, T);
            calc_trans(rt.view(), m, rt.L, tv);
            int last = -1;
            double r = 0.0;
            for (int i = start + 1; i < end; i++) {
                const int pl = path[i - 1 - sc];
                if (path[i - sc] != pl) continue;
                if (pl != last) {
                    int node, a;
                    rt.state(pl, node, a);
                    const double self = tv.get_time(a, a, rt.age[node], true);
                    r = 1.0 - (tv.norecombs[a] / self);
                    last = pl;
                }
                rate[i - sc] = r;
            }
        }
    };
    const int nthr = nb < 256 ? 1 : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (nthr == 1) {
        rates(0, nb);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthr; t++)
            pool.emplace_back(rates, (int)((long long)nb * t / nthr), (int)((long long)nb * (t + 1) / nthr));
        for (auto &th : pool) th.join();
    }

    RecombTree rt;
    std::vector<int> cn, ct;
    std::vector<double> probs;
    for (int bi = 0; bi < nb; bi++) {
        const int start = bstart[bi], end = bstart[bi + 1];
        bool built = false;
        int next_recomb = -1;
        // the first column of every block is a switch (or the very first column)
        for (int i = start + 1; i < end; i++) {
            const int pi = path[i - sc], pl = path[i - 1 - sc];
            if (pi == pl) {
                if (i > next_recomb) {
                    double ev = -log(frand()) / rate[i - sc];
                    next_recomb = int(fmin(double(end), i + ev));
                }
                if (i < next_recomb) continue;
            }
            next_recomb = -1;
            if (!built) {
                rt.build(tr.blocks[bi].tree, T);
                built = true;
            }
            const std::vector<int> &par = rt.par, &age = rt.age;
            const Lineages &L = rt.L;
            const int root = rt.view().root;
            int sn, st, ln, ltm;
            rt.state(pi, sn, st);
            rt.state(pl, ln, ltm);
            // get_possible_recomb (recomb.cpp:124-148)
            cn.clear();
            ct.clear();
            int end_time = st < ltm ? st : ltm;
            if (sn == ln)
                for (int k = age[sn]; k <= end_time; k++) { cn.push_back(sn); ct.push_back(k); }
            for (int k = 0; k <= end_time; k++) { cn.push_back(-1); ct.push_back(k); }
            // recomb_prob_unnormalized (recomb.cpp:14-119)
            probs.clear();
            for (size_t x = 0; x < cn.size(); x++) {
                int k = ct[x], j = st;
                int root_time = age[root] > ltm ? age[root] : ltm;
                int rpa = (cn[x] == -1 || par[cn[x]] == -1 || cn[x] == ln) ? ltm : age[par[cn[x]]];
                int nbk = L.nb[k] + int(k < ltm);
                int nrk = L.nr[k] + int(k <= ltm) + int(k == ltm) - int(k == root_time);
                double precomb = nbk * m.steps[k] / nrk;
                double sum = 0.0;
                for (int y = k; y < j - 1; y++) {
                    int nbm = L.nb[y] + int(y < ltm) - int(y < rpa);
                    sum += (m.steps[y] * nbm / (2.0 * m.pop[y]));
                }
                double pcoal = 1.0;
                int nbj = L.nb[j] + int(j < ltm) - int(j < rpa);
                if (k == j) {
                    if (j < T - 2) pcoal = 1.0 - exp(-m.csteps[2 * j] * nbj / (2.0 * m.pop[j]));
                } else {
                    int y = j - 1;
                    int nbm = L.nb[y] + int(y < ltm) - int(y < rpa);
                    sum += (m.csteps[2 * y] * nbm / (2.0 * m.pop[y]));
                    if (j < T - 2)
                        pcoal = 1.0 - exp(-m.csteps[2 * j - 1] * nbm / (2.0 * m.pop[y]) -
                                          m.csteps[2 * j] * nbj / (2.0 * m.pop[j]));
                }
                int ncj = L.nc[j] - int(j <= rpa) - int(j == rpa) + int(j <= ltm) + int(j == ltm);
                pcoal /= ncj;
                probs.push_back(precomb * exp(-sum) * pcoal);
            }
            int pick = sample_weights(probs.data(), (int)probs.size());
            rpos.push_back(i);
            rnode.push_back(cn[pick]);
            rtime.push_back(ct[pick]);
        }
    }
}

// sample_arg_thread (sample_thread.cpp:578-633)
static int sample_arg_thread(ArgTrees &tr, const double *times, int ntimes, const double *popsizes,
                             double rho, double mu, char **seqs, int new_chrom)
{
    HostInstance hi;
    if (build_instance(tr, times, ntimes, popsizes, rho, mu, seqs, new_chrom, hi)) return 1;
    const int L = tr.length();
    std::vector<double> u(L);
    for (int i = 0; i < L; i++) u[i] = frand();   // one draw per column, last column first
    std::vector<int> path(L);
    RunStats st;
    if (run_instance(hi, u.data(), path.data(), nullptr, nullptr, &st)) return 1;
    if (g_log_level >= 1)
        fprintf(stderr, "forward+trace on %s: prep %.3f ms, emit %.3f ms, fw %.3f ms, tb %.3f ms (%d blocks)\n",
                device_name(), st.ms_prep, st.ms_emit, st.ms_forward, st.ms_traceback,
                hi.ft.nblocks);
    std::vector<int> rpos, rnode, rtime;
    sample_recombinations(tr, hi.model, path.data(), rpos, rnode, rtime);
    add_arg_thread(tr, ntimes, path.data(), new_chrom, rpos, rnode, rtime);
    return 0;
}

// false (with a message) when no HIP device is visible: the HMM entry
// points then return NULL instead of computing anything on the host
// The internal-branch instance of an ARG whose trees are in
// subtree/maintree form: per-block inputs prepared on the host
// (prepare_internal), forward and traceback on the device.
// start_state >= 0: the first column is 1 on that internal state of the
// first block (a conditioned start), else the internal prior
static int build_internal(const ArgTrees &tr, const double *times, int ntimes,
                          const double *popsizes, double rho, double mu, char **seqs, int minage,
                          HostInstance &hi, int start_state = -1)
{
    if (ntimes < 2 || ntimes > MAXT) {
        fprintf(stderr, "argweaver_amd: %d time points (supported: 2..%d)\n", ntimes, MAXT);
        return 1;
    }
    flatten(tr, hi.ft);
    hi.rows.clear();   // the device kernels read the prepared inputs only
    std::vector<double> pop(ntimes);
    for (int i = 0; i < ntimes; i++) pop[i] = popsizes ? popsizes[i] : 0.0;
    model_init(hi.model, ntimes, times, pop.data(), rho, mu);
    auto hp = std::make_shared<HostPrep>();
    if (prepare_internal(tr, hi.model, seqs, minage, *hp)) return 1;
    if (start_state >= 0) {
        // the first column given: 1 on the start state
        hp->prior.assign(hp->prior.size(), 0.0);
        hp->prior[lane_pos(start_state, hp->R)] = 1.0;
    }
    hi.prep = hp;
    return 0;
}

// forward and traceback of the internal branch (sample_thread.cpp:637-668,
// 981-1006): one uniform per column, last column first
static int internal_thread_path(const ArgTrees &tr, const double *times, int ntimes,
                                const double *popsizes, double rho, double mu, char **seqs,
                                int minage, std::vector<int> &path, HostInstance &hi)
{
    if (build_internal(tr, times, ntimes, popsizes, rho, mu, seqs, minage, hi)) return 1;
    const int L = tr.length();
    std::vector<double> u(L);
    for (int i = 0; i < L; i++) u[i] = frand();
    path.assign(L, 0);
    return run_instance(hi, u.data(), path.data(), nullptr, nullptr, nullptr);
}

// sample_arg_thread_internal (sample_thread.cpp:637-694): re-thread the
// removed subtree into the ARG
static int sample_arg_thread_internal(ArgTrees &tr, const double *times, int ntimes,
                                      const double *popsizes, double rho, double mu, char **seqs,
                                      int minage)
{
    HostInstance hi;
    std::vector<int> path;
    if (internal_thread_path(tr, times, ntimes, popsizes, rho, mu, seqs, minage, path, hi))
        return 1;
    std::vector<int> rpos, rnode, rtime;
    sample_recombinations_internal(tr, hi.model, minage, path.data(), rpos, rnode, rtime);
    return add_thread_path(tr, ntimes, minage, path.data(), rpos, rnode, rtime);
}

// index of state (node, time) among the internal states of tree t (-1 when
// absent)
static int internal_state_index(const LTree &t, int ntimes, int node, int time)
{
    std::vector<int> sn, st;
    internal_states(t, ntimes, 0, sn, st);
    for (size_t k = 0; k < sn.size(); k++)
        if (sn[k] == node && st[k] == time) return (int)k;
    return -1;
}

// cond_sample_arg_thread_internal (sample_thread.cpp:771-865): re-thread the
// removed subtree with the first column's state given (start_node != -1) and
// the last column's state given (end_node != -1); an end without states (a
// complete tree) is given as its placeholder.  The last column's uniform is
// drawn only when its state is sampled.
static int cond_sample_arg_thread_internal(ArgTrees &tr, const double *times, int ntimes,
                                           const double *popsizes, double rho, double mu,
                                           char **seqs, int start_node, int start_time,
                                           int end_node, int end_time)
{
    const LTree &first = tr.blocks.front().tree, &last = tr.blocks.back().tree;
    int j = -1;
    if (start_node != -1 && first.nodes[first.root].age >= ntimes) {
        j = internal_state_index(first, ntimes, start_node, start_time);
        if (j < 0) {
            fprintf(stderr, "argweaver_amd: conditioned start state not in the first tree\n");
            return 1;
        }
    }
    HostInstance hi;
    if (build_internal(tr, times, ntimes, popsizes, rho, mu, seqs, 0, hi, j)) return 1;
    int k = -1;
    if (last.nodes[last.root].age < ntimes) {
        k = 0;   // a complete tree: its placeholder state
    } else if (end_node != -1) {
        k = internal_state_index(last, ntimes, end_node, end_time);
        if (k < 0) {
            fprintf(stderr, "argweaver_amd: conditioned end state not in the last tree\n");
            return 1;
        }
    }
    hi.last_state = k;
    const int L = tr.length();
    std::vector<double> u(L, 0.0);
    for (int i = (k >= 0) ? 1 : 0; i < L; i++) u[i] = frand();
    std::vector<int> path(L, 0);
    if (run_instance(hi, u.data(), path.data(), nullptr, nullptr, nullptr)) return 1;
    std::vector<int> rpos, rnode, rtime;
    sample_recombinations_internal(tr, hi.model, 0, path.data(), rpos, rnode, rtime);
    return add_thread_path(tr, ntimes, 0, path.data(), rpos, rnode, rtime);
}

// resample_arg_region (sample_arg.cpp:429-527): niters re-threadings of a
// uniformly drawn branch over the region [rstart, rend), each conditioned on
// the states the branch held at the region's ends (open at an end of the
// ARG) and accepted with probability exp(npaths - npaths2); the region is
// then spliced back.  Returns the acceptance rate, -1 on failure.
static double resample_arg_region(ArgTrees &tr, const double *times, int ntimes,
                                  const double *popsizes, double rho, double mu, char **seqs,
                                  int rstart, int rend, int niters, bool open_ended)
{
    const int maxtime = ntimes + 1;
    if (rstart == rend) return 1.0;
    if (rstart < tr.start_coord || rend > tr.end_coord || rstart > rend) {
        fprintf(stderr, "argweaver_amd: region [%d, %d) outside the ARG [%d, %d)\n", rstart,
                rend, tr.start_coord, tr.end_coord);
        return -1;
    }
    const int arg_start = tr.start_coord;
    ArgTrees t2 = partition_trees(tr, rstart);
    ArgTrees t3 = partition_trees(t2, rend);
    // a zero-length last block is extended by one column while the region
    // is re-threaded
    const bool stub = t2.blocks.back().blocklen == 0;
    if (stub) {
        t2.blocks.back().blocklen++;
        t2.end_coord++;
    }
    int accepts = 0;
    for (int it = 0; it < niters; it++) {
        ArgTrees old = t2;
        const LTree start_tree = t2.blocks.front().tree, end_tree = t2.blocks.back().tree;
        std::vector<int> path(t2.blocks.size());
        const double npaths = removal_path_uniform(t2, path.data());
        remove_thread_path(t2, path.data(), maxtime, nullptr);
        int sn, st, en, et;
        find_state_sub_tree_internal(start_tree, t2.blocks.front().tree, maxtime, sn, st);
        find_state_sub_tree_internal(end_tree, t2.blocks.back().tree, maxtime, en, et);
        if (open_ended) {
            if (rstart == arg_start) sn = st = -1;
            if (rend == t3.end_coord) en = et = -1;
        }
        if (cond_sample_arg_thread_internal(t2, times, ntimes, popsizes, rho, mu, seqs, sn, st,
                                            en, et))
            return -1;
        const double npaths2 = count_removal_paths(t2);
        const bool accept = frand() < exp(npaths - npaths2);
        if (!accept) t2 = std::move(old);
        else accepts++;
    }
    if (stub) {
        t2.blocks.back().blocklen--;
        t2.end_coord--;
    }
    append_trees(tr, t2);
    append_trees(tr, t3);
    return accepts / double(niters);
}

// resample_arg_regions (sample_arg.cpp:532-552): windows of `window`
// columns every `step` columns
static double resample_arg_regions(ArgTrees &tr, const double *times, int ntimes,
                                   const double *popsizes, double rho, double mu, char **seqs,
                                   int window, int step, int niters)
{
    double rate = 0.0;
    int nwindows = 0;
    for (int start = tr.start_coord; start == tr.start_coord || start + window / 2 < tr.end_coord;
         start += step) {
        nwindows++;
        const int end = std::min(start + window, tr.end_coord);
        const double r = resample_arg_region(tr, times, ntimes, popsizes, rho, mu, seqs, start, end,
                                             niters, true);
        if (r < 0) return -1;
        rate += r;
    }
    return rate / nwindows;
}

static bool require_gpu(const char *fn)
{
    if (device_available()) return true;
    fprintf(stderr, "argweaver_amd: %s needs a HIP device and none is visible\n", fn);
    return false;
}

}  // namespace aw

using namespace aw;

extern "C" {

//---------------------------------------------------------------------------
// handles

LocalTrees *arghmm_new_trees(int **ptrees, int **ages, int **sprs, int *blocklens, int ntrees,
                             int nnodes, int start_coord)
{
    return reinterpret_cast<LocalTrees *>(
        trees_from_arrays(ptrees, ages, sprs, blocklens, ntrees, nnodes, start_coord));
}

LocalTrees *arghmm_copy_trees(LocalTrees *trees)
{
    return reinterpret_cast<LocalTrees *>(new ArgTrees(*reinterpret_cast<ArgTrees *>(trees)));
}

int get_local_trees_ntrees(LocalTrees *trees)
{
    return (int)reinterpret_cast<ArgTrees *>(trees)->blocks.size();
}

int get_local_trees_nnodes(LocalTrees *trees)
{
    return reinterpret_cast<ArgTrees *>(trees)->nnodes;
}

void get_local_trees_ptrees(LocalTrees *trees, int **ptrees, int **ages, int **sprs, int *blocklens)
{
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    const int nleaves = tr.num_leaves();
    std::vector<int> perm(tr.nnodes);
    for (int i = 0; i < nleaves; i++) perm[i] = tr.seqids[i];
    for (int i = nleaves; i < tr.nnodes; i++) perm[i] = i;
    if (!check_trees(tr)) {
        fprintf(stderr, "argweaver_amd: get_local_trees_ptrees on an invalid ARG\n");
        abort();
    }
    for (size_t i = 0; i < tr.blocks.size(); i++) {
        const LBlock &B = tr.blocks[i];
        for (int j = 0; j < B.tree.nnodes; j++) {
            int parent = B.tree.nodes[j].parent;
            if (parent != -1) parent = perm[parent];
            ptrees[i][perm[j]] = parent;
            ages[i][perm[j]] = B.tree.nodes[j].age;
        }
        blocklens[i] = B.blocklen;
        if (!spr_is_null(B.spr)) {
            sprs[i][0] = perm[B.spr.rn];
            sprs[i][1] = B.spr.rt;
            sprs[i][2] = perm[B.spr.cn];
            sprs[i][3] = B.spr.ct;
        } else {
            sprs[i][0] = B.spr.rn;
            sprs[i][1] = B.spr.rt;
            sprs[i][2] = B.spr.cn;
            sprs[i][3] = B.spr.ct;
        }
    }
}

void get_local_trees_blocks(LocalTrees *trees, int *starts, int *ends)
{
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    int end = tr.start_coord;
    for (size_t i = 0; i < tr.blocks.size(); i++) {
        starts[i] = end;
        end += tr.blocks[i].blocklen;
        ends[i] = end;
    }
}

void get_treelens(LocalTrees *trees, double *times, int ntimes, double *treelens)
{
    (void)ntimes;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    for (size_t i = 0; i < tr.blocks.size(); i++) {
        const LTree &t = tr.blocks[i].tree;
        double len = 0.0;
        for (int j = 0; j < t.nnodes; j++) {
            int p = t.nodes[j].parent;
            if (p != -1) len += times[t.nodes[p].age] - times[t.nodes[j].age];
        }
        treelens[i] = len;
    }
}

void delete_local_trees(LocalTrees *trees) { delete reinterpret_cast<ArgTrees *>(trees); }

//---------------------------------------------------------------------------
// state spaces

void arghmm_get_nstates(LocalTrees *trees, int ntimes, bool internal, int *nstates)
{
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    std::vector<int> sn, st;
    int i = 0;
    for (const LBlock &B : tr.blocks) {
        int S = 0;
        if (internal) {
            internal_states(B.tree, ntimes, 0, sn, st);
            S = (int)sn.size();
        } else {
            for (int x = 0; x < B.tree.nnodes; x++) {
                const int p = B.tree.nodes[x].parent;
                S += (p == -1 ? ntimes - 2 : B.tree.nodes[p].age) - B.tree.nodes[x].age + 1;
            }
        }
        for (int j = 0; j < B.blocklen; j++) nstates[i + j] = S;
        i += B.blocklen;
    }
}

intstate **get_state_spaces(LocalTrees *trees, int ntimes, bool internal)
{
    (void)internal;   // the reference ignores it here (states.cpp:230-251)
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    intstate **all = new intstate *[tr.blocks.size()];
    std::vector<int> sn, st;
    for (size_t i = 0; i < tr.blocks.size(); i++) {
        tree_states(tr.blocks[i].tree, ntimes, sn, st);
        all[i] = new intstate[sn.size()];
        for (size_t j = 0; j < sn.size(); j++) {
            all[i][j][0] = sn[j];
            all[i][j][1] = st[j];
        }
    }
    return all;
}

void delete_state_spaces(intstate **all_states, int ntrees)
{
    for (int i = 0; i < ntrees; i++) delete[] all_states[i];
    delete[] all_states;
}

//---------------------------------------------------------------------------
// threading HMM

double **arghmm_forward_alg(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                            double rho, double mu, char **seqs, int nseqs, int seqlen,
                            bool prior_given, double *prior, bool internal, bool slow)
{
    (void)slow;   // the dense variant computes the same table; one device kernel serves both
    if (!require_gpu("arghmm_forward_alg")) return nullptr;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    HostInstance hi;
    std::vector<int> Sv;
    long long W = 0;
    if (internal) {
        // the trees are in subtree/maintree form; the subtree is the thread
        if (build_internal(tr, times, ntimes, popsizes, rho, mu, seqs, 0, hi)) return nullptr;
        Sv = hi.prep->S;
        for (int b = 0; b < hi.ft.nblocks; b++) W += (long long)Sv[b] * hi.ft.len[b];
        prior_given = false;   // the prepared first column is the reference's internal prior
    } else {
        const int new_chrom = tr.num_leaves();   // ArgHmmMatrixIter default (matrices.h:224-226)
        if (new_chrom >= nseqs) {
            fprintf(stderr, "argweaver_amd: arghmm_forward_alg needs %d sequences\n",
                    new_chrom + 1);
            return nullptr;
        }
        if (build_instance(tr, times, ntimes, popsizes, rho, mu, seqs, new_chrom, hi))
            return nullptr;
        W = instance_states(hi.ft, ntimes, &Sv);
    }
    double *table = new double[W > 0 ? W : 1];
    if (run_instance(hi, nullptr, nullptr, table, nullptr, nullptr, Precision::FP64,
                     prior_given ? prior : nullptr)) {
        delete[] table;
        return nullptr;
    }
    // rows of the caller's table alias the contiguous block-major buffer
    double **fw = new double *[seqlen > 0 ? seqlen : 1];
    for (int i = 0; i < seqlen; i++) fw[i] = nullptr;
    long long off = 0;
    int pos = tr.start_coord;
    for (int b = 0; b < hi.ft.nblocks; b++) {
        for (int i = 0; i < hi.ft.len[b]; i++, pos++)
            if (pos >= 0 && pos < seqlen) fw[pos] = table + off + (long long)i * Sv[b];
        off += (long long)hi.ft.len[b] * Sv[b];
    }
    if (seqlen > 0 && fw[0] != table) {
        // keep the base pointer reachable for delete_forward_matrix
        fprintf(stderr, "argweaver_amd: forward table must start at coordinate 0\n");
        delete[] table;
        delete[] fw;
        return nullptr;
    }
    return fw;
}

void delete_forward_matrix(double **mat, int nrows)
{
    (void)nrows;
    if (!mat) return;
    delete[] mat[0];
    delete[] mat;
}

void delete_double_matrix(double **mat, int nrows)
{
    (void)nrows;
    if (!mat) return;
    delete[] mat[0];
    delete[] mat;
}

intstate *arghmm_sample_posterior(int **ptrees, int **ages, int **sprs, int *blocklens, int ntrees,
                                  int nnodes, double *times, int ntimes, double *popsizes,
                                  double rho, double mu, char **seqs, int nseqs, int seqlen,
                                  intstate *path)
{
    (void)nseqs;
    if (!require_gpu("arghmm_sample_posterior")) return nullptr;
    ArgTrees *tr = trees_from_arrays(ptrees, ages, sprs, blocklens, ntrees, nnodes, 0);
    HostInstance hi;
    if (build_instance(*tr, times, ntimes, popsizes, rho, mu, seqs, tr->num_leaves(), hi)) {
        delete tr;
        return nullptr;
    }
    std::vector<double> u(seqlen);
    for (int i = 0; i < seqlen; i++) u[i] = frand();
    std::vector<int> ipath(seqlen);
    if (run_instance(hi, u.data(), ipath.data(), nullptr, nullptr, nullptr)) {
        delete tr;
        return nullptr;
    }
    if (!path) path = new intstate[seqlen];
    // The reference converts with a shadowed block end (sample_thread.cpp:959-970):
    // every block writes the columns starting at start_coord.  Kept as is.
    std::vector<int> sn, st;
    for (const LBlock &B : tr->blocks) {
        tree_states(B.tree, ntimes, sn, st);
        for (int i = tr->start_coord; i < tr->start_coord + B.blocklen; i++) {
            int istate = ipath[i];
            path[i][0] = sn[istate];
            path[i][1] = st[istate];
        }
    }
    delete tr;
    return path;
}

LocalTrees *arghmm_sample_thread(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                                 double rho, double mu, char **seqs, int nseqs, int seqlen)
{
    (void)seqlen;
    if (!require_gpu("arghmm_sample_thread")) return nullptr;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    if (sample_arg_thread(tr, times, ntimes, popsizes, rho, mu, seqs, nseqs - 1)) return nullptr;
    return trees;
}

// sample_arg_seq (sample_arg.cpp:26-67), in input order; nonzero when a
// threading failed (the ARG then holds the leaves threaded so far)
static int sample_arg_seq(ArgTrees &tr, double *times, int ntimes, double *popsizes, double rho,
                           double mu, char **seqs, int nseqs, int seqlen)
{
    if (tr.num_leaves() == 0 || tr.blocks.empty()) {
        int start = tr.start_coord, end = tr.end_coord;
        if (end != seqlen) {
            start = 0;
            end = seqlen;
        }
        make_trunk(tr, start, end, 2 * nseqs - 1);
    }
    std::vector<char> has(nseqs, 0);
    for (int i = 0; i < tr.num_leaves(); i++) has[tr.seqids[i]] = 1;
    for (int i = 0; i < nseqs; i++)
        if (!has[i] && sample_arg_thread(tr, times, ntimes, popsizes, rho, mu, seqs, i)) return 1;
    return 0;
}

// resample_arg (sample_arg.cpp:71-107); nonzero when a threading failed
static int resample_arg(ArgTrees &tr, double *times, int ntimes, double *popsizes, double rho,
                         double mu, char **seqs, int nremove)
{
    const int nleaves = tr.num_leaves();
    if (nremove == 1) {
        for (int chrom = 0; chrom < nleaves; chrom++) {
            remove_arg_thread(tr, chrom);
            if (sample_arg_thread(tr, times, ntimes, popsizes, rho, mu, seqs, chrom)) return 1;
        }
        return 0;
    }
    if (nremove <= 0) return 0;
    if (nremove > nleaves - 1) nremove = nleaves - 1;
    std::vector<int> avail(nleaves);
    for (int i = 0; i < nleaves; i++) avail[i] = i;
    // shuffle with irand (common.h:109-140)
    for (int i = 1; i < nleaves; i++) {
        int j = int(rand() / float(RAND_MAX) * i);
        if (j == i) j = i - 1;
        std::swap(avail[i], avail[j]);
    }
    for (int i = 0; i < nremove; i++) remove_arg_thread(tr, avail[i]);
    for (int i = 0; i < nremove; i++)
        if (sample_arg_thread(tr, times, ntimes, popsizes, rho, mu, seqs, avail[i])) return 1;
    return 0;
}

LocalTrees *arghmm_sample_arg_seq(double *times, int ntimes, double *popsizes, double rho, double mu,
                                  char **seqs, int nseqs, int seqlen)
{
    if (!require_gpu("arghmm_sample_arg_seq")) return nullptr;
    ArgTrees *tr = new ArgTrees();
    if (sample_arg_seq(*tr, times, ntimes, popsizes, rho, mu, seqs, nseqs, seqlen)) {
        delete tr;
        return nullptr;
    }
    return reinterpret_cast<LocalTrees *>(tr);
}

LocalTrees *arghmm_sample_arg_refine(double *times, int ntimes, double *popsizes, double rho,
                                     double mu, char **seqs, int nseqs, int seqlen, int niters,
                                     int nremove)
{
    if (!require_gpu("arghmm_sample_arg_refine")) return nullptr;
    ArgTrees *tr = new ArgTrees();
    int rc = sample_arg_seq(*tr, times, ntimes, popsizes, rho, mu, seqs, nseqs, seqlen);
    for (int i = 0; i < niters && !rc; i++)
        rc = resample_arg(*tr, times, ntimes, popsizes, rho, mu, seqs, nremove);
    if (rc) {
        delete tr;
        return nullptr;
    }
    return reinterpret_cast<LocalTrees *>(tr);
}

LocalTrees *arghmm_resample_arg(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                                double rho, double mu, char **seqs, int nseqs, int seqlen,
                                int niters, int nremove)
{
    (void)seqlen;
    if (!require_gpu("arghmm_resample_arg")) return nullptr;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    // arghmm_complete_arg (sample_arg.cpp:655-662)
    for (int c = tr.num_leaves(); c < nseqs; c++)
        if (sample_arg_thread(tr, times, ntimes, popsizes, rho, mu, seqs, c)) return nullptr;
    for (int i = 0; i < niters; i++)
        if (resample_arg(tr, times, ntimes, popsizes, rho, mu, seqs, nremove)) return nullptr;
    return trees;
}

void arghmm_sample_arg_thread_internal(LocalTrees *trees, double *times, int ntimes,
                                       double *popsizes, double rho, double mu, char **seqs,
                                       int nseqs, int seqlen, int *thread_path)
{
    (void)nseqs; (void)seqlen;
    if (!require_gpu("arghmm_sample_arg_thread_internal")) return;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    HostInstance hi;
    std::vector<int> path;
    if (internal_thread_path(tr, times, ntimes, popsizes, rho, mu, seqs, 0, path, hi)) {
        fprintf(stderr, "argweaver_amd: arghmm_sample_arg_thread_internal failed\n");
        return;
    }
    std::copy(path.begin(), path.end(), thread_path);
}

// resample_arg_all (sample_arg.cpp:111-127): a random branch at a random
// position, its removal path, then the internal re-threading
static int resample_arg_all(ArgTrees &tr, const double *times, int ntimes,
                            const double *popsizes, double rho, double mu, char **seqs,
                            double prob_path_switch)
{
    const int maxtime = ntimes + 1;
    std::vector<int> path(tr.blocks.size());
    const int node = irand(tr.nnodes);
    const int pos = irand(tr.start_coord, tr.end_coord);
    removal_path_at(tr, node, pos, path.data(), prob_path_switch);
    remove_thread_path(tr, path.data(), maxtime, nullptr);
    return sample_arg_thread_internal(tr, times, ntimes, popsizes, rho, mu, seqs, 0);
}

// resample_arg_recomb (sample_arg.cpp:211-223)
static int resample_arg_recomb(ArgTrees &tr, const double *times, int ntimes,
                               const double *popsizes, double rho, double mu, char **seqs,
                               double recomb_preference)
{
    std::vector<int> path(tr.blocks.size());
    removal_path_recomb(tr, recomb_preference, path.data());
    remove_thread_path(tr, path.data(), ntimes + 1, nullptr);
    return sample_arg_thread_internal(tr, times, ntimes, popsizes, rho, mu, seqs, 0);
}

// resample_arg_leaf (sample_arg.cpp:131-149), phased data
static int resample_arg_leaf(ArgTrees &tr, const double *times, int ntimes,
                             const double *popsizes, double rho, double mu, char **seqs)
{
    std::vector<int> path(tr.blocks.size());
    const int node = irand(tr.num_leaves());
    removal_leaf_path(tr, node, path.data());
    remove_thread_path(tr, path.data(), ntimes + 1, nullptr);
    return sample_arg_thread_internal(tr, times, ntimes, popsizes, rho, mu, seqs, 0);
}

LocalTrees *arghmm_resample_arg_region(LocalTrees *trees, double *times, int ntimes,
                                       double *popsizes, double rho, double mu, char **seqs,
                                       int nseqs, int seqlen, int region_start, int region_end,
                                       int niters)
{
    (void)nseqs;
    (void)seqlen;
    if (!require_gpu("arghmm_resample_arg_region")) return nullptr;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    if (resample_arg_region(tr, times, ntimes, popsizes, rho, mu, seqs, region_start, region_end,
                            niters, true) < 0)
        return nullptr;
    return trees;
}

// resample_arg_mcmc_all (sample_arg.cpp:190-203) per iteration: a leaf's
// re-threading with probability frac_leaf = 0.5, else windows of `window`
// columns every window / 2 columns, niters2 re-threadings each
LocalTrees *arghmm_resample_mcmc_arg(LocalTrees *trees, double *times, int ntimes,
                                     double *popsizes, double rho, double mu, char **seqs,
                                     int nseqs, int seqlen, int niters, int niters2, int window)
{
    (void)seqlen;
    if (!require_gpu("arghmm_resample_mcmc_arg")) return nullptr;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    const double frac_leaf = 0.5;
    const int step = window / 2;
    for (int c = tr.num_leaves(); c < nseqs; c++)
        if (sample_arg_thread(tr, times, ntimes, popsizes, rho, mu, seqs, c)) return nullptr;
    for (int i = 0; i < niters; i++) {
        if (frand() < frac_leaf) {
            if (resample_arg_leaf(tr, times, ntimes, popsizes, rho, mu, seqs)) return nullptr;
        } else if (resample_arg_regions(tr, times, ntimes, popsizes, rho, mu, seqs, window, step,
                                        niters2) < 0) {
            return nullptr;
        }
    }
    return trees;
}

LocalTrees *arghmm_resample_all_arg(LocalTrees *trees, double *times, int ntimes,
                                    double *popsizes, double rho, double mu, char **seqs,
                                    int nseqs, int seqlen, int niters, double prob_path_switch)
{
    (void)seqlen;
    if (!require_gpu("arghmm_resample_all_arg")) return nullptr;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    for (int c = tr.num_leaves(); c < nseqs; c++)
        if (sample_arg_thread(tr, times, ntimes, popsizes, rho, mu, seqs, c)) return nullptr;
    for (int i = 0; i < niters; i++)
        if (resample_arg_all(tr, times, ntimes, popsizes, rho, mu, seqs, prob_path_switch))
            return nullptr;
    return trees;
}

LocalTrees *arghmm_resample_climb_arg(LocalTrees *trees, double *times, int ntimes,
                                      double *popsizes, double rho, double mu, char **seqs,
                                      int nseqs, int seqlen, int niters, double recomb_preference)
{
    (void)seqlen;
    if (!require_gpu("arghmm_resample_climb_arg")) return nullptr;
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    for (int c = tr.num_leaves(); c < nseqs; c++)
        if (sample_arg_thread(tr, times, ntimes, popsizes, rho, mu, seqs, c)) return nullptr;
    for (int i = 0; i < niters; i++)
        if (resample_arg_recomb(tr, times, ntimes, popsizes, rho, mu, seqs, recomb_preference))
            return nullptr;
    return trees;
}

void delete_path(int *path) { delete[] path; }

void setLogLevel(int level) { g_log_level = level; }

//---------------------------------------------------------------------------
// extensions of this library (no counterpart in the reference)

LocalTrees *awamd_remove_thread(LocalTrees *trees, int seqid)
{
    remove_arg_thread(*reinterpret_cast<ArgTrees *>(trees), seqid);
    return trees;
}

int awamd_thread_forward(LocalTrees *trees, const double *times, int ntimes,
                         const double *popsizes, double rho, double mu, char **seqs, int new_chrom,
                         const double *uniforms, int precision, double *alpha_out, int *path_out,
                         double *lnl_out, double *ms_out)
{
    if (!device_available()) {
        fprintf(stderr, "argweaver_amd: awamd_thread_forward needs a HIP device\n");
        return -1;
    }
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    HostInstance hi;
    if (build_instance(tr, times, ntimes, popsizes, rho, mu, seqs, new_chrom, hi)) return -1;
    RunStats st;
    int rc = run_instance(hi, uniforms, path_out, alpha_out, lnl_out, &st,
                          precision == 1 ? Precision::FP32 : Precision::FP64);
    if (ms_out) {
        ms_out[0] = st.ms_prep;
        ms_out[1] = st.ms_emit;
        ms_out[2] = st.ms_forward;
        ms_out[3] = st.ms_traceback;
    }
    return rc;
}

int awamd_thread_forward_noalloc(LocalTrees *trees, const double *times, int ntimes,
                                 const double *popsizes, double rho, double mu, char **seqs,
                                 int new_chrom, const double *uniforms, int precision,
                                 double *alpha_out, int *path_out, double *lnl_out,
                                 double *ms_out)
{
    if (!device_available()) {
        fprintf(stderr, "argweaver_amd: awamd_thread_forward_noalloc needs a HIP device\n");
        return -1;
    }
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    HostInstance hi;
    if (build_instance(tr, times, ntimes, popsizes, rho, mu, seqs, new_chrom, hi)) return -1;
    RunStats st;
    int rc = run_instance_noalloc(hi, uniforms, path_out, alpha_out, lnl_out, &st,
                                  precision == 1 ? Precision::FP32 : Precision::FP64);
    if (ms_out) {
        ms_out[0] = st.ms_prep;
        ms_out[1] = st.ms_emit;
        ms_out[2] = st.ms_forward;
        ms_out[3] = st.ms_traceback;
    }
    return rc;
}

long long awamd_num_column_states(LocalTrees *trees, int ntimes, int *nstates_per_block)
{
    const ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    long long W = 0;
    for (size_t b = 0; b < tr.blocks.size(); b++) {
        const LTree &t = tr.blocks[b].tree;
        int S = 0;
        for (int i = 0; i < t.nnodes; i++) {
            const int p = t.nodes[i].parent;
            S += (p == -1 ? ntimes - 2 : t.nodes[p].age) - t.nodes[i].age + 1;
        }
        if (nstates_per_block) nstates_per_block[b] = S;
        W += (long long)S * tr.blocks[b].blocklen;
    }
    return W;
}

int awamd_get_mappings(LocalTrees *trees, int *mappings)
{
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    const int N = tr.nnodes;
    for (size_t b = 0; b < tr.blocks.size(); b++)
        for (int i = 0; i < N; i++)
            mappings[b * N + i] = tr.blocks[b].mapping.empty() ? -1 : tr.blocks[b].mapping[i];
    return 0;
}

int awamd_get_seqids(LocalTrees *trees, int *seqids)
{
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    for (size_t i = 0; i < tr.seqids.size(); i++) seqids[i] = tr.seqids[i];
    return (int)tr.seqids.size();
}

int awamd_device_available(void) { return device_available() ? 1 : 0; }

int awamd_device_info(char *buf, int n) { return device_info(buf, n); }

//---------------------------------------------------------------------------
// batches of independent threading instances

struct AwBatch {
    Batch batch;
    int tb_warm = BatchOptions().tile_warm;
    int tb_opts = BatchOptions().tile_opts;
    int tb_tail = BatchOptions().tail_tile_cols;
};

void *awamd_batch_new(void) { return new AwBatch(); }

void awamd_batch_free(void *b) { delete reinterpret_cast<AwBatch *>(b); }

int awamd_batch_add(void *b, LocalTrees *trees, const double *times, int ntimes,
                    const double *popsizes, double rho, double mu, char **seqs, int new_chrom,
                    const double *uniforms)
{
    AwBatch &ab = *reinterpret_cast<AwBatch *>(b);
    HostInstance hi;
    if (build_instance(*reinterpret_cast<ArgTrees *>(trees), times, ntimes, popsizes, rho, mu,
                       seqs, new_chrom, hi))
        return -1;
    return ab.batch.add(std::move(hi), uniforms);
}

int awamd_batch_clear(void *b)
{
    reinterpret_cast<AwBatch *>(b)->batch.clear();
    return 0;
}

int awamd_batch_upload(void *b)
{
    if (!device_available()) {
        fprintf(stderr, "argweaver_amd: awamd_batch_upload needs a HIP device\n");
        return -1;
    }
    return reinterpret_cast<AwBatch *>(b)->batch.upload();
}

int awamd_batch_run_opts(void *b, int precision, int tile_cols, int traceback, int fixup,
                         double *stats)
{
    AwBatch &ab = *reinterpret_cast<AwBatch *>(b);
    BatchOptions opt;
    opt.prec = precision == 1 ? Precision::FP32 : Precision::FP64;
    opt.tile_cols = tile_cols;
    opt.tile_warm = ab.tb_warm;
    opt.tile_opts = ab.tb_opts;
    opt.tail_tile_cols = ab.tb_tail;
    opt.traceback = traceback != 0;
    opt.fixup = fixup != 0;
    RunStats st;
    int rc = ab.batch.run(opt, &st);
    if (stats) {
        stats[0] = st.ms_prep;
        stats[1] = st.ms_emit;
        stats[2] = st.ms_forward;
        stats[3] = st.ms_traceback;
        stats[4] = st.ms_total;
        stats[5] = (double)st.W;
        stats[6] = st.nvar;
        stats[7] = st.tb_passes;
        stats[8] = (double)st.tb_rerun_cols;
    }
    return rc;
}

int awamd_batch_set_traceback(void *b, int warm, int options)
{
    AwBatch &ab = *reinterpret_cast<AwBatch *>(b);
    if (warm < 0 || options < 1 || options > 4) {
        fprintf(stderr, "argweaver_amd: traceback warm-up %d / options %d out of range\n", warm,
                options);
        return -1;
    }
    ab.tb_warm = warm;
    ab.tb_opts = options;
    return 0;
}

int awamd_batch_set_tail_tiles(void *b, int cols)
{
    if (cols < 0) {
        fprintf(stderr, "argweaver_amd: tail traceback tile length %d out of range\n", cols);
        return -1;
    }
    reinterpret_cast<AwBatch *>(b)->tb_tail = cols;
    return 0;
}

int awamd_batch_run(void *b, int precision, int tile_cols, int traceback, double *stats)
{
    return awamd_batch_run_opts(b, precision, tile_cols, traceback, 1, stats);
}

int awamd_batch_run_ex(void *b, int precision, int tile_cols, int traceback, int fixup,
                       int fw_tile, int fw_warm, double fw_tol, double *stats)
{
    AwBatch &ab = *reinterpret_cast<AwBatch *>(b);
    BatchOptions opt;
    opt.prec = precision == 1 ? Precision::FP32 : Precision::FP64;
    opt.tile_cols = tile_cols;
    opt.tile_warm = ab.tb_warm;
    opt.tile_opts = ab.tb_opts;
    opt.tail_tile_cols = ab.tb_tail;
    opt.traceback = traceback != 0;
    opt.fixup = fixup != 0;
    opt.fw_tile = fw_tile;
    if (fw_warm > 0) opt.fw_warm = fw_warm;
    if (fw_tol > 0) opt.fw_tol = fw_tol;
    RunStats st;
    int rc = ab.batch.run(opt, &st);
    if (stats) {
        const double v[14] = {st.ms_prep, st.ms_emit, st.ms_forward, st.ms_traceback, st.ms_total,
                              (double)st.W, (double)st.nvar, (double)st.tb_passes,
                              (double)st.tb_rerun_cols, (double)st.fw_tiles, (double)st.fw_reruns,
                              st.ms_fw_main, st.ms_tb_tail, st.ms_tb_side};
        std::copy(v, v + 14, stats);
    }
    return rc;
}

int awamd_batch_tune(void *b, int precision, const int *cand, int ncand, double *ms_out)
{
    AwBatch &ab = *reinterpret_cast<AwBatch *>(b);
    return ab.batch.tune_tiles(precision == 1 ? Precision::FP32 : Precision::FP64, cand, ncand,
                               ms_out, ab.tb_warm, ab.tb_opts);
}

int awamd_batch_tune_forward(void *b, const int *tiles, const int *warms, int ncand, double tol,
                             double *ms_out)
{
    return reinterpret_cast<AwBatch *>(b)->batch.tune_forward(tiles, warms, ncand, tol, ms_out);
}

int awamd_batch_path(void *b, int i, int *path)
{
    return reinterpret_cast<AwBatch *>(b)->batch.copy_path(i, path);
}

int awamd_batch_alpha(void *b, int i, double *alpha)
{
    return reinterpret_cast<AwBatch *>(b)->batch.copy_alpha(i, alpha);
}

int awamd_batch_alpha_range(void *b, int i, int c0, int c1, double *alpha)
{
    return reinterpret_cast<AwBatch *>(b)->batch.copy_alpha_range(i, c0, c1, alpha);
}

int awamd_batch_norms(void *b, int i, double *norms)
{
    return reinterpret_cast<AwBatch *>(b)->batch.copy_norms(i, norms);
}

double awamd_batch_lnl(void *b, int i) { return reinterpret_cast<AwBatch *>(b)->batch.lnl(i); }

long long awamd_batch_W(void *b, int i) { return reinterpret_cast<AwBatch *>(b)->batch.W(i); }

LocalTrees *awamd_add_thread(LocalTrees *trees, int ntimes, const int *path, int seqid,
                             int nrecomb, const int *recomb_pos, const int *recomb_node,
                             const int *recomb_time)
{
    std::vector<int> rp(recomb_pos, recomb_pos + nrecomb), rn(recomb_node, recomb_node + nrecomb),
        rt(recomb_time, recomb_time + nrecomb);
    add_arg_thread(*reinterpret_cast<ArgTrees *>(trees), ntimes, path, seqid, rp, rn, rt);
    return trees;
}

int awamd_sample_recombinations(LocalTrees *trees, const double *times, int ntimes,
                                const double *popsizes, double rho, double mu, const int *path,
                                int maxrecomb, int *recomb_pos, int *recomb_node, int *recomb_time)
{
    ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    Model m;
    std::vector<double> pop(ntimes);
    for (int i = 0; i < ntimes; i++) pop[i] = popsizes[i];
    model_init(m, ntimes, times, pop.data(), rho, mu);
    std::vector<int> rp, rn, rt;
    sample_recombinations(tr, m, path, rp, rn, rt);
    for (size_t i = 0; i < rp.size() && (int)i < maxrecomb; i++) {
        recomb_pos[i] = rp[i];
        recomb_node[i] = rn[i];
        recomb_time[i] = rt[i];
    }
    return (int)rp.size();
}

}  // extern "C"

namespace aw {
ArgTrees *simulate_arg(int k, int seqlen, const double *times, int T, const double *popsizes,
                       double rho, double mu, uint64_t seed, char *seqs);
}

extern "C" LocalTrees *awamd_simulate(int k, int seqlen, const double *times, int ntimes,
                                      const double *popsizes, double rho, double mu,
                                      unsigned long long seed, char *seqs)
{
    return reinterpret_cast<LocalTrees *>(
        aw::simulate_arg(k, seqlen, times, ntimes, popsizes, rho, mu, seed, seqs));
}

//---------------------------------------------------------------------------
// continuous-batching server

extern "C" {

void *awamd_server_new_ex(double hbm_budget_gb, int precision, int tile_cols, int max_batch,
                          int drivers)
{
    if (!device_available()) {
        fprintf(stderr, "argweaver_amd: awamd_server_new needs a HIP device\n");
        return nullptr;
    }
    if (drivers < 1 || drivers > 4) {
        fprintf(stderr, "argweaver_amd: %d server drivers (1 .. 4)\n", drivers);
        return nullptr;
    }
    ServeOptions o;
    o.hbm_budget = (size_t)(hbm_budget_gb * 1073741824.0);
    o.max_batch = max_batch;
    o.drivers = drivers;
    o.run.prec = precision == 1 ? Precision::FP32 : Precision::FP64;
    o.run.tile_cols = tile_cols;
    // FP32 requests longer than 50000 columns run as forward tiles
    if (precision == 1) {
        o.run.fw_tile = 50000;
        o.run.fw_warm = 20000;
        o.run.fw_tol = 1e-4;
    }
    return new Server(o);
}

void *awamd_server_new(double hbm_budget_gb, int precision, int tile_cols, int max_batch)
{
    return awamd_server_new_ex(hbm_budget_gb, precision, tile_cols, max_batch, ServeOptions().drivers);
}

void awamd_server_free(void *server) { delete reinterpret_cast<Server *>(server); }

int awamd_server_submit(void *server, LocalTrees *trees, const double *times, int ntimes,
                        const double *popsizes, double rho, double mu, char **seqs,
                        int new_chrom, const double *uniforms)
{
    HostInstance hi;
    if (build_instance(*reinterpret_cast<ArgTrees *>(trees), times, ntimes, popsizes, rho, mu,
                       seqs, new_chrom, hi))
        return -1;
    return reinterpret_cast<Server *>(server)->submit(hi, uniforms);
}

int awamd_server_wait(void *server, int id, double timeout_s, int *path, double *lnl,
                      double *times3)
{
    ServeResult r;
    const int w = reinterpret_cast<Server *>(server)->wait(id, &r, timeout_s);
    if (w == 2) return -2;
    if (w) return -1;
    if (path) std::copy(r.path.begin(), r.path.end(), path);
    if (lnl) *lnl = r.lnl;
    if (times3) {
        times3[0] = r.t_submit;
        times3[1] = r.t_start;
        times3[2] = r.t_done;
    }
    return r.status;
}

int awamd_server_wait_ex(void *server, int id, double timeout_s, int *path, double *lnl,
                         double *info5)
{
    ServeResult r;
    const int w = reinterpret_cast<Server *>(server)->wait(id, &r, timeout_s);
    if (w == 2) return -2;
    if (w) return -1;
    if (path) std::copy(r.path.begin(), r.path.end(), path);
    if (lnl) *lnl = r.lnl;
    if (info5) {
        info5[0] = r.t_submit;
        info5[1] = r.t_start;
        info5[2] = r.t_done;
        info5[3] = r.batch_size;
        info5[4] = r.driver;
    }
    return r.status;
}

double awamd_server_now(void *server) { return reinterpret_cast<Server *>(server)->now(); }

}  // extern "C"
