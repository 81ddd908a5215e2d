// stats_gpu.hip -- the chain statistics an MCMC run reports every iteration,
// on the device: log P(sequences | ARG) (calc_arg_likelihood,
// src/argweaver/total_prob.cpp:19-45, with likelihood_tree,
// src/argweaver/emit.cpp:399-449).
//
// The reference walks the columns of every local tree, prunes each variant
// column through the tree (JC69, uniform root, branch lengths floored at a
// tenth of the first time step) and reuses the first invariant column's
// likelihood for every later invariant column of the same tree.  Here:
//   k_lnl_class  one thread per column: is every row equal at the column
//   k_lnl_first  one thread per local tree: its first invariant column
//   k_lnl_eval   one wave per column that needs a pruning (variant, or its
//                tree's first invariant column): level-synchronous pruning
//                with the four partial likelihoods of every node in LDS,
//                one node per lane and level
//   k_lnl_sum    one workgroup per 256 columns: the column's log-likelihood
//                (its own, or its tree's first invariant column's) summed in
//                a fixed order; the host adds the workgroup sums in order.
// The terms are the reference's; only the summation order differs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/argweaver_amd.h"
#include "aw_math.h"
#include "local_trees.h"

namespace aw {
namespace {

#define LNL_OK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "argweaver_amd: %s: %s\n", #x, hipGetErrorString(e_));      \
            return 1;                                                                    \
        }                                                                                \
    } while (0)

struct LnlArgs {
    int B, N, L, nrows, nleaves;
    double mintime, mu;
    const double *times;            // [T]
    const int *parent, *c0, *c1, *age, *height;   // [B*N]
    const int *hmax;                // [B] largest node height of the tree
    const int *bstart;              // [B+1] column of each tree's start, then L
    const char *rows;               // [nrows*L]: row j is leaf j's sequence
    int *invariant;                 // [L]
    int *first;                     // [B] first invariant column, -1 for none
    double *lk;                     // [L] likelihood of the pruned columns
    double *part;                   // [ceil(L/256)] workgroup sums
};

__device__ int owner_block(const int *bstart, int B, int col)
{
    int lo = 0, hi = B - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (bstart[mid] <= col) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_lnl_class(LnlArgs a)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.L) return;
    const char r0 = a.rows[c];
    int inv = 1;
    for (int j = 1; j < a.nrows && inv; j++) inv = a.rows[(size_t)j * a.L + c] == r0;
    a.invariant[c] = inv;
}

__global__ __launch_bounds__(256) void k_lnl_first(LnlArgs a)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= a.B) return;
    int f = -1;
    for (int c = a.bstart[b]; c < a.bstart[b + 1]; c++)
        if (a.invariant[c]) {
            f = c;
            break;
        }
    a.first[b] = f;
}

constexpr int LNL_MAXN = MAXN;

// one wave per column
__global__ __launch_bounds__(64) void k_lnl_eval(LnlArgs a)
{
    __shared__ double inner[LNL_MAXN][4];
    __shared__ double mut[LNL_MAXN], nomut[LNL_MAXN];
    const int c = blockIdx.x, lane = threadIdx.x;
    const int b = owner_block(a.bstart, a.B, c);
    // an invariant column other than its tree's first takes that one's value
    if (a.invariant[c] && a.first[b] != c) return;   // uniform per wave
    const int N = a.N;
    const int *par = a.parent + (size_t)b * N, *ch0 = a.c0 + (size_t)b * N;
    const int *ch1 = a.c1 + (size_t)b * N, *ag = a.age + (size_t)b * N;
    const int *hg = a.height + (size_t)b * N;
    int root = -1;
    for (int i = lane; i < N; i += 64) {
        const int p = par[i];
        if (p == -1) {
            root = i;
            mut[i] = nomut[i] = 0.0;
        } else {
            const double t = fmax(a.times[ag[p]] - a.times[ag[i]], a.mintime);
            mut[i] = prob_branch(t, a.mu, true);
            nomut[i] = prob_branch(t, a.mu, false);
        }
    }
    // the root's index from whichever lane found it
    for (int off = 32; off > 0; off >>= 1) root = max(root, __shfl_xor(root, off));
    __syncthreads();
    const int H = a.hmax[b];
    for (int h = 0; h <= H; h++) {
        for (int i = lane; i < N; i += 64) {
            if (hg[i] != h) continue;
            if (ch0[i] == -1) {
                const int code = base_code(a.rows[(size_t)i * a.L + c]);   // -1: 'N'
                for (int x = 0; x < 4; x++) inner[i][x] = (code < 0 || x == code) ? 1.0 : 0.0;
            } else {
                const int u = ch0[i], v = ch1[i];
                for (int x = 0; x < 4; x++) {
                    double p1 = 0.0, p2 = 0.0;
                    for (int y = 0; y < 4; y++) {
                        p1 += inner[u][y] * (x == y ? nomut[u] : mut[u]);
                        p2 += inner[v][y] * (x == y ? nomut[v] : mut[v]);
                    }
                    inner[i][x] = p1 * p2;
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        double lk = 0.0;
        for (int x = 0; x < 4; x++) lk += inner[root][x] * .25;
        a.lk[c] = lk;
    }
}

__global__ __launch_bounds__(256) void k_lnl_sum(LnlArgs a)
{
    __shared__ double s[256];
    const int c = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (c < a.L) {
        const int b = owner_block(a.bstart, a.B, c);
        v = log(a.invariant[c] ? a.lk[a.first[b]] : a.lk[c]);
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[blockIdx.x] = s[0];
}

// node heights (leaves 0, a node one more than its higher child) of every
// tree, and each tree's largest
void tree_heights(const FlatTrees &ft, std::vector<int> &height, std::vector<int> &hmax)
{
    const int B = ft.nblocks, N = ft.nnodes;
    height.assign((size_t)B * N, 0);
    hmax.assign(B, 0);
    parallel_ranges(B, [&](int lo, int hi) {
        std::vector<int> pending(N);
        std::vector<int> ready;
        for (int b = lo; b < hi; b++) {
            const int *par = &ft.parent[(size_t)b * N], *ch0 = &ft.child0[(size_t)b * N];
            int *h = &height[(size_t)b * N];
            ready.clear();
            for (int i = 0; i < N; i++) {
                pending[i] = ch0[i] == -1 ? 0 : 2;
                if (!pending[i]) ready.push_back(i);
            }
            int top = 0;
            for (size_t k = 0; k < ready.size(); k++) {
                const int i = ready[k];
                top = std::max(top, h[i]);
                const int p = par[i];
                if (p == -1) continue;
                h[p] = std::max(h[p], h[i] + 1);
                if (--pending[p] == 0) ready.push_back(p);
            }
            hmax[b] = top;
        }
    });
}

struct DevBufs {
    std::vector<void *> ptrs;
    ~DevBufs()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T *&p, size_t n)
    {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return 1;
        ptrs.push_back(q);
        p = static_cast<T *>(q);
        return 0;
    }
};

int arg_loglik_device(const ArgTrees &tr, const double *times, int ntimes, double mu, char **seqs,
                      int nseqs, int seqlen, double *out)
{
    FlatTrees ft;
    flatten(tr, ft);
    const int B = ft.nblocks, N = ft.nnodes, L = seqlen;
    if (tr.length() != seqlen) {
        fprintf(stderr, "argweaver_amd: awamd_arg_likelihood: ARG length %d, %d columns\n",
                tr.length(), seqlen);
        return 1;
    }
    if (N > LNL_MAXN) {
        fprintf(stderr, "argweaver_amd: awamd_arg_likelihood: %d nodes (limit %d)\n", N, LNL_MAXN);
        return 1;
    }
    std::vector<int> height, hmax;
    tree_heights(ft, height, hmax);
    std::vector<int> bstart(B + 1);
    for (int b = 0; b < B; b++) bstart[b] = ft.start[b];
    bstart[B] = L;
    // row j is leaf j's sequence (seqids) for the tree's leaves, then the rest
    std::vector<const char *> rowp(nseqs);
    for (int j = 0; j < nseqs; j++) rowp[j] = seqs[j < (int)tr.seqids.size() ? tr.seqids[j] : j];

    DevBufs m;
    LnlArgs a{};
    a.B = B; a.N = N; a.L = L; a.nrows = nseqs; a.nleaves = tr.num_leaves();
    a.mintime = times[1] * .1;
    a.mu = mu;
    double *dt;
    int *dp, *dc0, *dc1, *dag, *dh, *dhm, *dbs, *dinv, *dfirst;
    char *drows;
    double *dlk, *dpart;
    const int nwg = (L + 255) / 256;
    if (m.alloc(dt, ntimes) || m.alloc(dp, (size_t)B * N) || m.alloc(dc0, (size_t)B * N) ||
        m.alloc(dc1, (size_t)B * N) || m.alloc(dag, (size_t)B * N) || m.alloc(dh, (size_t)B * N) ||
        m.alloc(dhm, B) || m.alloc(dbs, B + 1) || m.alloc(drows, (size_t)nseqs * L) ||
        m.alloc(dinv, L) || m.alloc(dfirst, B) || m.alloc(dlk, L) || m.alloc(dpart, nwg)) {
        fprintf(stderr, "argweaver_amd: awamd_arg_likelihood: device allocation failed\n");
        return 1;
    }
    LNL_OK(hipMemcpy(dt, times, sizeof(double) * ntimes, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dp, ft.parent.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dc0, ft.child0.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dc1, ft.child1.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dag, ft.age.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dh, height.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dhm, hmax.data(), sizeof(int) * B, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dbs, bstart.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice));
    for (int j = 0; j < nseqs; j++)
        LNL_OK(hipMemcpy(drows + (size_t)j * L, rowp[j], L, hipMemcpyHostToDevice));
    a.times = dt; a.parent = dp; a.c0 = dc0; a.c1 = dc1; a.age = dag; a.height = dh;
    a.hmax = dhm; a.bstart = dbs; a.rows = drows; a.invariant = dinv; a.first = dfirst;
    a.lk = dlk; a.part = dpart;
    hipLaunchKernelGGL(k_lnl_class, dim3(nwg), dim3(256), 0, 0, a);
    hipLaunchKernelGGL(k_lnl_first, dim3((B + 255) / 256), dim3(256), 0, 0, a);
    hipLaunchKernelGGL(k_lnl_eval, dim3(L), dim3(64), 0, 0, a);
    hipLaunchKernelGGL(k_lnl_sum, dim3(nwg), dim3(256), 0, 0, a);
    LNL_OK(hipGetLastError());
    std::vector<double> part(nwg);
    LNL_OK(hipMemcpy(part.data(), dpart, sizeof(double) * nwg, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (double x : part) s += x;
    *out = s;
    return 0;
}

// log P(ARG) (calc_arg_prior, src/argweaver/total_prob.cpp:262-299): a
// term per local tree, each on a thread of its own: the probability of the
// tree's recombination-free stretch and, when a tree follows, of the SPR
// into it (calc_spr_prob, total_prob.cpp:210-257).
struct PriorArgs {
    int B, N;
    const Model *m;
    const int *parent, *c0, *c1, *age;   // [B*N]
    const int *spr;                      // [4B] the SPR into block b
    const int *len;                      // [B]
    const int *has_next;                 // [B] a block follows before the end
    double *term;                        // [B]
    double *part;                        // [ceil(B/256)]
};

__device__ double spr_logprob_dev(const TreeView &t, const Model &m, int rn, int rt, int ct,
                                  double treelen)
{
    const int T = m.T;
    const int root_age = t.age[t.root];
    const double treelen_b = treelen + m.steps[root_age];
    Lineages L;
    count_lineages(t, T, L);
    L.nr[root_age]--;
    double lnl = 0.0;
    const int k = rt;
    lnl += log(L.nb[k] * m.steps[k] / (L.nr[k] * treelen_b));
    const int broken_age = t.age[t.parent[rn]];
    const int j = ct;
    auto rate = [&](int i) {
        const int nb = L.nb[i / 2] - int(i / 2 < broken_age);
        return m.csteps[i] * nb / (2.0 * m.pop[i / 2]);
    };
    const int ncoals_j = L.nc[j] - int(j <= broken_age) - int(j == broken_age);
    lnl -= log((double)ncoals_j);
    if (j < T - 2) lnl += log(1.0 - exp(-rate(2 * j) - (j > k ? rate(2 * j - 1) : 0.0)));
    for (int x = 2 * k; x < 2 * j - 1; x++) lnl -= rate(x);
    return lnl;
}

__global__ __launch_bounds__(256) void k_prior_block(PriorArgs a)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= a.B) return;
    const int N = a.N;
    const int *par = a.parent + (size_t)b * N;
    int root = 0;
    while (par[root] != -1) root++;
    const TreeView t{N, root, par, a.c0 + (size_t)b * N, a.c1 + (size_t)b * N,
                     a.age + (size_t)b * N};
    const Model &m = *a.m;
    const double treelen = tree_length(t, m.times);
    const double rr = fmax(m.rho * treelen, m.rho);
    double v;
    if (a.has_next[b]) {
        const int *sp = a.spr + 4 * (size_t)(b + 1);
        v = log(rr) - rr * a.len[b] + spr_logprob_dev(t, m, sp[0], sp[1], sp[3], treelen);
    } else {
        v = -rr * a.len[b];
    }
    a.term[b] = v;
}

__global__ __launch_bounds__(256) void k_sum_terms(const double *term, int n, double *part)
{
    __shared__ double s[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    s[threadIdx.x] = i < n ? term[i] : 0.0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

int arg_logprior_device(const ArgTrees &tr, const Model &model, double *out)
{
    FlatTrees ft;
    flatten(tr, ft);
    const int B = ft.nblocks, N = ft.nnodes;
    std::vector<int> has_next(B);
    long long end = tr.start_coord;
    for (int b = 0; b < B; b++) {
        end += ft.len[b];
        has_next[b] = end < tr.end_coord ? 1 : 0;
    }
    DevBufs m;
    Model *dm;
    int *dp, *dc0, *dc1, *dag, *dspr, *dlen, *dnext;
    double *dterm, *dpart;
    const int nwg = (B + 255) / 256;
    if (m.alloc(dm, 1) || m.alloc(dp, (size_t)B * N) || m.alloc(dc0, (size_t)B * N) ||
        m.alloc(dc1, (size_t)B * N) || m.alloc(dag, (size_t)B * N) || m.alloc(dspr, 4 * (size_t)B) ||
        m.alloc(dlen, B) || m.alloc(dnext, B) || m.alloc(dterm, B) || m.alloc(dpart, nwg)) {
        fprintf(stderr, "argweaver_amd: awamd_arg_prior: device allocation failed\n");
        return 1;
    }
    LNL_OK(hipMemcpy(dm, &model, sizeof(Model), hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dp, ft.parent.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dc0, ft.child0.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dc1, ft.child1.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dag, ft.age.data(), sizeof(int) * B * N, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dspr, ft.spr.data(), sizeof(int) * 4 * B, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dlen, ft.len.data(), sizeof(int) * B, hipMemcpyHostToDevice));
    LNL_OK(hipMemcpy(dnext, has_next.data(), sizeof(int) * B, hipMemcpyHostToDevice));
    PriorArgs a{B, N, dm, dp, dc0, dc1, dag, dspr, dlen, dnext, dterm, dpart};
    hipLaunchKernelGGL(k_prior_block, dim3(nwg), dim3(256), 0, 0, a);
    hipLaunchKernelGGL(k_sum_terms, dim3(nwg), dim3(256), 0, 0, (const double *)dterm, B, dpart);
    LNL_OK(hipGetLastError());
    std::vector<double> part(nwg);
    LNL_OK(hipMemcpy(part.data(), dpart, sizeof(double) * nwg, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (double x : part) s += x;
    *out = s;
    return 0;
}

}  // namespace
}  // namespace aw

using namespace aw;

extern "C" double awamd_arg_likelihood(LocalTrees *trees, const double *times, int ntimes,
                                       double mu, char **seqs, int nseqs, int seqlen)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        fprintf(stderr, "argweaver_amd: awamd_arg_likelihood needs a HIP device\n");
        return NAN;
    }
    if (ntimes < 3 || ntimes > MAXT) {
        fprintf(stderr, "argweaver_amd: awamd_arg_likelihood: ntimes %d outside 3..%d\n", ntimes,
                MAXT);
        return NAN;
    }
    const ArgTrees &tr = *reinterpret_cast<ArgTrees *>(trees);
    if (tr.nnodes < 3) return log(.25) * seqlen;
    double v = NAN;
    if (arg_loglik_device(tr, times, ntimes, mu, seqs, nseqs, seqlen, &v)) return NAN;
    return v;
}

extern "C" double awamd_arg_prior(LocalTrees *trees, const double *times, int ntimes,
                                  const double *popsizes, double rho)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        fprintf(stderr, "argweaver_amd: awamd_arg_prior needs a HIP device\n");
        return NAN;
    }
    if (ntimes < 3 || ntimes > MAXT) {
        fprintf(stderr, "argweaver_amd: awamd_arg_prior: ntimes %d outside 3..%d\n", ntimes, MAXT);
        return NAN;
    }
    std::vector<double> pops(ntimes, 1.0);
    if (popsizes)
        for (int i = 0; i < ntimes; i++) pops[i] = popsizes[i];
    Model m;
    model_init(m, ntimes, times, pops.data(), rho, 0.0);
    double v = NAN;
    if (arg_logprior_device(*reinterpret_cast<ArgTrees *>(trees), m, &v)) return NAN;
    return v;
}
