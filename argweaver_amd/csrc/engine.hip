// engine.hip -- gfx950 kernels of the threading HMM and their host driver.
//
// Everything runs on one HIP stream over a Batch of independent instances
// whose inputs (local trees, sequences, uniforms) are resident in HBM.
//
//   k_classify_b    one lane per column: masked / invariant / variant
//   k_scan_b        one workgroup per instance: variant-column numbering
//   k_prep_b        one workgroup per local-tree block: state space, lineage
//                   counts, TransMatrix vectors, the (a,b) and (c,a,b)
//                   transition tables, the switch matrix from the previous
//                   block, emission terms, first-column prior
//   k_var_emit_b    one wave per variant column: Felsenstein partials and the
//                   emission row over the block's states
//   k_prep32_b      one wave per block: FP32 tables and the time-grouped lane
//                   assignment used by the FP32 forward
//
// Plain FP64 path (the drop-in, reference arithmetic):
//   k_forward64_b   one workgroup per instance, previous column in LDS,
//                   per-time sums by wave reductions, two barriers per column
//   k_traceback64_b one wave per instance, next row prefetched into
//                   registers, CDF by a wave scan
//
// FP32 column-tiled path (normalised FP32 table, FP64 normalisers and CDF):
//   k_forward32     one wave per (instance, column tile): each lane owns the
//                   states of one time bin, so per-time sums are a segmented
//                   wave scan; one LDS sync per column.  Tiles after the first
//                   start from a uniform column and are then re-run from the
//                   true boundary column until they agree with the stored
//                   speculative columns (fix-up passes).
//   k_traceback32   one wave per tile, speculative start state; fix-up passes
//                   re-run from the true boundary state until the re-sampled
//                   path coalesces with the stored one (exact from there on).
//   k_lnl_b         per-instance sum of log normalisers
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "engine.h"

namespace aw {

#define HIP_OK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "argweaver_amd: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

// LDS ordering inside a single-wave workgroup: complete this wave's LDS
// operations and keep the compiler from moving memory accesses across.
__device__ inline void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Pointers read out of DevInst are generic; accesses through them compile to
// flat instructions, which count against the LDS wait counter too, so every
// LDS wait would also drain outstanding HBM stores.  gp() restores the
// global address space so loads and stores become global_* instructions.
#define AW_GLOBAL __attribute__((address_space(1)))
template <class X>
__device__ inline AW_GLOBAL X *gp(X *p)
{
    return (AW_GLOBAL X *)p;
}

// Read-only data indexed by wave-uniform values (block bounds, column
// classes, uniforms) goes through the constant address space so it is
// fetched with scalar loads, which do not wait on outstanding vector stores.
#define AW_CONST __attribute__((address_space(4)))
template <class X>
__device__ inline AW_CONST const X *cp(const X *p)
{
    return (AW_CONST const X *)p;
}

// index i such that off[i] <= x < off[i+1]
template <class PT, class X>
__device__ inline int find_owner(PT off, int n, X x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

//===========================================================================
// column classification

__global__ void k_classify_b(const DevInst *__restrict__ D, const long long *coloff, int ninst)
{
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= coloff[ninst]) return;
    int i = find_owner(coloff, ninst, g);
    const DevInst &d = D[i];
    int col = (int)(g - coloff[i]);
    char c0 = gp(d.seq)[col];
    bool inv = true;
    for (int r = 1; r < d.nrows; r++)
        if (gp(d.seq)[(size_t)r * d.L + col] != c0) { inv = false; break; }
    gp(d.vflag)[col] = inv ? 0 : 1;
    gp(d.cls)[col] = inv ? (c0 == 'N' ? -2 : -1) : 0;
}

// variant-column numbering of one instance: each of 1024 threads owns a
// contiguous run of columns
__global__ __launch_bounds__(1024) void k_scan_b(const DevInst *__restrict__ D, int *nvar_out)
{
    __shared__ int part[1024];
    const DevInst &d = D[blockIdx.x];
    const int tid = threadIdx.x;
    const int per = (d.L + 1023) / 1024;
    const int s0 = min(d.L, tid * per), e0 = min(d.L, s0 + per);
    int loc = 0;
    for (int i = s0; i < e0; i++) loc += gp(d.vflag)[i];
    part[tid] = loc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (tid >= off) ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - loc;
    for (int i = s0; i < e0; i++) {
        if (gp(d.vflag)[i]) {
            gp(d.cls)[i] = run;
            gp(d.vcol)[run] = i;
            run++;
        }
    }
    if (tid == 1023) nvar_out[blockIdx.x] = part[1023];
}

//===========================================================================
// per-block preparation

struct PrepShared {
    Model M;
    int par[MAXN], ch0[MAXN], ch1[MAXN], ag[MAXN];
    int ppar[MAXN], pch0[MAXN], pch1[MAXN], pag[MAXN];
    int map[MAXN];
    int root, proot;
    Lineages L, PL;
    TransVec tv;
    NodeIndex ni, pni;
    SwitchCtx cx;
    double mainlen;
    int recombsrc, recoalsrc;
    int node3, rparent;
};

__device__ inline int node_of_state(const NodeIndex &ni, int n, int k)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ni.first[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_prep_b(const DevInst *__restrict__ D,
                                                const int *__restrict__ blkoff, int ninst)
{
    __shared__ PrepShared sh;
    const int inst = find_owner(blkoff, ninst, (int)blockIdx.x);
    const DevInst &d = D[inst];
    const int b = blockIdx.x - blkoff[inst];
    const int tid = threadIdx.x, NT = blockDim.x;
    const int T = d.T, N = d.N, T1 = T - 1;

    if (tid == 0) sh.M = *d.model;
    for (int i = tid; i < N; i += NT) {
        size_t x = (size_t)b * N + i;
        sh.par[i] = gp(d.parent)[x];
        sh.ch0[i] = gp(d.c0)[x];
        sh.ch1[i] = gp(d.c1)[x];
        sh.ag[i] = gp(d.age)[x];
        if (b > 0) {
            size_t y = (size_t)(b - 1) * N + i;
            sh.ppar[i] = gp(d.parent)[y];
            sh.pch0[i] = gp(d.c0)[y];
            sh.pch1[i] = gp(d.c1)[y];
            sh.pag[i] = gp(d.age)[y];
            sh.map[i] = gp(d.map)[x];
        }
    }
    __syncthreads();
    TreeView t{N, -1, sh.par, sh.ch0, sh.ch1, sh.ag};
    TreeView pt{N, -1, sh.ppar, sh.pch0, sh.pch1, sh.pag};
    Spr spr{0, 0, 0, 0};
    if (b > 0) spr = Spr{gp(d.spr)[4 * b], gp(d.spr)[4 * b + 1], gp(d.spr)[4 * b + 2], gp(d.spr)[4 * b + 3]};
    if (tid == 0) {
        int r = 0;
        for (int i = 0; i < N; i++) if (sh.par[i] == -1) r = i;
        sh.root = r;
        t.root = r;
        count_lineages(t, T, sh.L);
        calc_trans(t, sh.M, sh.L, sh.tv);
        build_node_index(t, T, sh.ni);
        sh.mainlen = emit_mainlen(t, sh.M);
        gp(d.roots)[b] = r;
        int order[MAXN];
        postorder(t, order);
        for (int i = 0; i < N; i++) gp(d.order)[(size_t)b * N + i] = order[i];
        if (b > 0) {
            int pr = 0;
            for (int i = 0; i < N; i++) if (sh.ppar[i] == -1) pr = i;
            sh.proot = pr;
            pt.root = pr;
            count_lineages(pt, T, sh.PL);
            build_node_index(pt, T, sh.pni);
            switch_ctx(pt, sh.M, sh.PL, spr, sh.cx);
        }
    }
    __syncthreads();
    t.root = sh.root;
    pt.root = (b > 0) ? sh.proot : -1;
    const Model &M = sh.M;
    const int so = cp(d.soff)[b], S = cp(d.soff)[b + 1] - so;

    for (int x = tid; x < 9 * T; x += NT) {
        int v = x / T, a = x % T;
        double val;
        switch (v) {
        case 0: val = sh.tv.D[a]; break;
        case 1: val = sh.tv.E[a]; break;
        case 2: val = sh.tv.lnB[a]; break;
        case 3: val = sh.tv.lnE2[a]; break;
        case 4: val = sh.tv.lnNegG1[a]; break;
        case 5: val = sh.tv.G2[a]; break;
        case 6: val = sh.tv.G3[a]; break;
        case 7: val = sh.tv.lnG4[a]; break;
        default: val = sh.tv.norecombs[a]; break;
        }
        gp(d.tv)[(size_t)b * 9 * T + x] = val;
    }

    for (int n = tid; n < N; n += NT) {
        gp(d.nfirst)[(size_t)b * N + n] = sh.ni.first[n];
        gp(d.nhi)[(size_t)b * N + n] = sh.ni.hi[n];
        for (int a = sh.ag[n]; a <= sh.ni.hi[n]; a++) {
            int k = sh.ni.first[n] + a - sh.ag[n];
            gp(d.snode)[so + k] = n;
            gp(d.stime)[so + k] = a;
        }
        // branch mutation probabilities for pruning (prob_tree_mutation)
        double mut = 0.0, nomut = 0.0;
        if (n != sh.root) {
            const double mintime = M.times[1] * .1;
            double tt = fmax(M.times[sh.ag[sh.par[n]]] - M.times[sh.ag[n]], mintime);
            mut = prob_branch(tt, M.mu, true);
            nomut = prob_branch(tt, M.mu, false);
        }
        gp(d.nmut)[2 * ((size_t)b * N + n)] = mut;
        gp(d.nmut)[2 * ((size_t)b * N + n) + 1] = nomut;
    }

    for (int k = tid; k < S; k += NT) {
        int n = node_of_state(sh.ni, N, k);
        int a = sh.ag[n] + k - sh.ni.first[n];
        EmitTerms e;
        emit_terms(t, M, sh.mainlen, n, a, e);
        gp(d.einv)[so + k] = e.inv;
        for (int x = 0; x < 3; x++) {
            gp(d.eterm)[6 * (size_t)(so + k) + x] = e.mut[x];
            gp(d.eterm)[6 * (size_t)(so + k) + 3 + x] = e.nomut[x];
        }
        if (b == 0) gp(d.prior)[k] = prior_time(a, sh.L, M);
    }

    const int TT = T1 * T1;
    for (int x = tid; x < TT; x += NT) {
        int a = x / T1, bb = x % T1;
        gp(d.tm)[(size_t)b * TT + x] = sh.tv.get_time(a, bb, 0, false);
    }
    for (int x = tid; x < d.TC; x += NT) {
        int c = 0, base = 0;
        while (base + (T1 - c) * (T1 - c) <= x) {
            base += (T1 - c) * (T1 - c);
            c++;
        }
        int r = x - base, w = T1 - c;
        int a = c + r / w, bb = c + r % w;
        double same = sh.tv.get_time(a, bb, c, true);
        gp(d.tm2s)[(size_t)b * d.TC + x] = same;
        gp(d.tm2)[(size_t)b * d.TC + x] = same - sh.tv.get_time(a, bb, 0, false);
    }

    // states grouped by time (counting sort, stable in state order)
    if (tid == 0) {
        int cnt[MAXT + 1];
        for (int a = 0; a <= T; a++) cnt[a] = 0;
        for (int n = 0; n < N; n++)
            for (int a = sh.ag[n]; a <= sh.ni.hi[n]; a++) cnt[a + 1]++;
        for (int a = 0; a < T; a++) cnt[a + 1] += cnt[a];
        for (int a = 0; a <= T; a++) gp(d.tstart)[(size_t)b * (T + 1) + a] = cnt[a];
        for (int n = 0; n < N; n++)
            for (int a = sh.ag[n]; a <= sh.ni.hi[n]; a++)
                gp(d.tperm)[so + cnt[a]++] = sh.ni.first[n] + a - sh.ag[n];
    }

    if (b == 0) return;

    // switch matrix from block b-1 (calc_transition_probs_switch)
    const int so1 = cp(d.soff)[b - 1], S1 = so - so1;
    for (int j = tid; j < S1; j += NT) {
        int n1 = node_of_state(sh.pni, N, j);
        int t1 = sh.pag[n1] + j - sh.pni.first[n1];
        int tgt = sw_determ(pt, t, sh.ni, spr, sh.map, n1, t1);
        gp(d.determ)[so1 + j] = tgt;
        gp(d.determp)[so1 + j] = tgt >= 0 ? sw_determprob(pt, M, sh.PL, spr, sh.cx, n1, t1) : 0.0;
    }
    for (int k = tid; k < S; k += NT) {
        gp(d.recombrow)[so + k] = 0.0;
        gp(d.recoalrow)[so + k] = 0.0;
    }
    if (tid == 0) {
        int rs = lookup(pt, sh.pni, spr.rn, spr.rt);
        int cs = lookup(pt, sh.pni, spr.cn, spr.ct);
        sh.recombsrc = rs;
        sh.recoalsrc = cs;
        gp(d.srcs)[2 * b] = rs;
        gp(d.srcs)[2 * b + 1] = cs;
        if (cs != -1) {
            int node1 = spr.cn;
            int last_parent = sh.ppar[spr.rn];
            if (last_parent == node1) {
                int other = (sh.pch1[last_parent] == spr.rn) ? sh.pch0[last_parent] : sh.pch1[last_parent];
                sh.node3 = sh.map[other];
            } else {
                sh.node3 = sh.map[node1];
            }
            sh.rparent = sh.par[sh.map[spr.rn]];
        }
    }
    __syncthreads();
    if (tid == 0 && sh.recombsrc != -1) {
        // get_recomb_transition_switch: stay and escape targets
        int parent = sh.ppar[spr.rn];
        int time2 = sh.pag[parent];
        int other = pt.sibling(spr.rn);
        int node2 = (other == spr.cn) ? sh.par[sh.map[other]] : sh.map[other];
        int stay = lookup(t, sh.ni, sh.map[spr.rn], spr.rt);
        int esc = lookup(t, sh.ni, node2, time2);
        if (stay != -1)
            gp(d.recombrow)[so + stay] = sw_recomb_recoal(pt, M, sh.PL, spr, spr.ct, spr.rn, spr.rt,
                                                      sh.cx.rpa, sh.cx.last_len);
        if (esc != -1)
            gp(d.recombrow)[so + esc] = sw_recomb_recoal(pt, M, sh.PL, spr, spr.ct, spr.rn, spr.rt,
                                                     spr.rt, sh.cx.last_len);
    }
    if (sh.recoalsrc != -1) {
        const int time1 = spr.ct;
        const int mrn = sh.map[spr.rn];
        for (int k = tid; k < S; k += NT) {
            int n2 = node_of_state(sh.ni, N, k);
            int t2 = sh.ag[n2] + k - sh.ni.first[n2];
            if (!((n2 == mrn && t2 >= spr.rt) || (n2 == sh.node3 && t2 == time1) ||
                  (n2 == sh.rparent && t2 == time1)))
                continue;
            gp(d.recoalrow)[so + k] = sw_recomb_recoal(pt, M, sh.PL, spr, t2, spr.cn, spr.ct,
                                                   sh.cx.rpa, sh.cx.last_len);
        }
    }
    __syncthreads();
    // inverse of the deterministic map, sources ascending per target
    if (tid == 0) {
        auto ptr = gp(d.invptr) + so + b;
        for (int k = 0; k <= S; k++) ptr[k] = 0;
        for (int j = 0; j < S1; j++) {
            int k = gp(d.determ)[so1 + j];
            if (j != sh.recombsrc && j != sh.recoalsrc && k != -1) ptr[k + 1]++;
        }
        for (int k = 0; k < S; k++) ptr[k + 1] += ptr[k];
        for (int j = 0; j < S1; j++) {
            int k = gp(d.determ)[so1 + j];
            if (j != sh.recombsrc && j != sh.recoalsrc && k != -1) gp(d.invidx)[so1 + ptr[k]++] = j;
        }
        for (int k = S; k > 0; k--) ptr[k] = ptr[k - 1];
        ptr[0] = 0;
    }
}

//===========================================================================
// emissions of variant columns

__global__ __launch_bounds__(64) void k_var_emit_b(const DevInst *__restrict__ D,
                                                   const int *__restrict__ voff, int ninst)
{
    __shared__ int par[MAXN], ch0[MAXN], ch1[MAXN], ag[MAXN], order[MAXN];
    __shared__ char bases[MAXN];
    __shared__ double muts[MAXN], nomuts[MAXN];
    __shared__ double inner[4 * MAXN], outer[4 * MAXN];
    __shared__ double in2[4];
    const int inst = find_owner(voff, ninst, (int)blockIdx.x);
    const DevInst &d = D[inst];
    const int v = blockIdx.x - voff[inst];
    const int lane = threadIdx.x;
    const int N = d.N;
    const int col = gp(d.vcol)[v];
    const int b = find_owner(cp(d.bstart), d.B, col);
    const int nleaves = (N + 1) / 2;
    for (int i = lane; i < N; i += 64) {
        size_t x = (size_t)b * N + i;
        par[i] = gp(d.parent)[x];
        ch0[i] = gp(d.c0)[x];
        ch1[i] = gp(d.c1)[x];
        ag[i] = gp(d.age)[x];
        order[i] = gp(d.order)[x];
        muts[i] = gp(d.nmut)[2 * x];
        nomuts[i] = gp(d.nmut)[2 * x + 1];
        if (i < nleaves) bases[i] = gp(d.seq)[(size_t)i * d.L + col];
    }
    __syncthreads();
    const int root = gp(d.roots)[b];
    if (lane == 0) {
        TreeView t{N, root, par, ch0, ch1, ag};
        pruning(t, order, bases, muts, nomuts, inner, outer);
        leaf_partial(gp(d.seq)[(size_t)nleaves * d.L + col], in2);
    }
    __syncthreads();
    const int so = cp(d.soff)[b], S = cp(d.soff)[b + 1] - so;
    auto row = gp(d.vemit) + (size_t)v * d.Smax;
    for (int k = lane; k < S; k += 64) {
        int n = gp(d.snode)[so + k];
        const double *et = d.eterm + 6 * (size_t)(so + k);
        row[k] = emit_variant(in2, inner + 4 * n, outer + 4 * n, n == root, et, et + 3);
    }
}

//===========================================================================
// FP32 tables and lane assignment: lanes are grouped by time bin so that a
// lane's slots all share one time; R slots per lane is the smallest value
// for which the bins fit in 64 lanes.

__global__ __launch_bounds__(64) void k_prep32_b(const DevInst *__restrict__ D,
                                                 const int *__restrict__ blkoff, int ninst,
                                                 int *overflow)
{
    const int inst = find_owner(blkoff, ninst, (int)blockIdx.x);
    const DevInst &d = D[inst];
    const int b = blockIdx.x - blkoff[inst];
    const int lane = threadIdx.x;
    const int T = d.T, T1 = T - 1, TT = T1 * T1, RM = d.RM;
    const int so = cp(d.soff)[b], S = cp(d.soff)[b + 1] - so;
    for (int x = lane; x < TT; x += 64) gp(d.tm32)[(size_t)b * TT + x] = (float)gp(d.tm)[(size_t)b * TT + x];
    for (int x = lane; x < d.TC; x += 64) gp(d.tm2_32)[(size_t)b * d.TC + x] = (float)gp(d.tm2)[(size_t)b * d.TC + x];
    for (int k = lane; k < S; k += 64) gp(d.einv32)[so + k] = (float)gp(d.einv)[so + k];
    for (int x = lane; x < 64 * RM; x += 64) gp(d.lslot)[(size_t)b * 64 * RM + x] = -1;
    gp(d.lkey)[(size_t)b * 64 + lane] = T1;
    __syncthreads();
    if (lane == 0) {
        const int *ts = d.tstart + (size_t)b * (T + 1);
        int R = 1;
        for (; R <= RM; R++) {
            int lanes = 0;
            for (int a = 0; a < T1; a++) lanes += (ts[a + 1] - ts[a] + R - 1) / R;
            if (lanes <= 64) break;
        }
        if (R > RM) {
            atomicMax(overflow, R);
            return;
        }
        int l = 0;
        for (int a = 0; a < T1; a++) {
            for (int x = ts[a]; x < ts[a + 1]; x += R) {
                int e = min(ts[a + 1], x + R);
                for (int y = x; y < e; y++)
                    gp(d.lslot)[((size_t)b * 64 + l) * RM + (y - x)] = gp(d.tperm)[so + y];
                gp(d.lkey)[(size_t)b * 64 + l] = a;
                l++;
            }
        }
    }
}

//===========================================================================
// plain FP64 forward, one workgroup per instance

__device__ inline double wave_sum(double x)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

template <int NT>
__device__ inline void time_sums(const double *col, const int *tperm, const int *tstart, int T1,
                                 double *G)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int a = w; a < T1; a += NW) {
        double s = 0.0;
        for (int x = tstart[a] + lane; x < tstart[a + 1]; x += 64) s += col[tperm[x]];
        s = wave_sum(s);
        if (lane == 0) G[a] = s;
    }
}

template <int NT, int R>
__global__ __launch_bounds__(NT) void k_forward64_b(const DevInst *__restrict__ D,
                                                    const int *__restrict__ list)
{
    extern __shared__ double dyn[];
    const DevInst &d = D[list[blockIdx.x]];
    const int tid = threadIdx.x, lane = tid & 63;
    const int T1 = d.T - 1, TT = T1 * T1, N = d.N;
    double *bufA = dyn;
    double *bufB = bufA + d.Smax;
    double *tm2 = bufB + d.Smax;
    double *G = tm2 + d.TC;                      // 32 doubles
    int *tperm = (int *)(G + 32);                // Smax ints
    int *tstart = tperm + d.Smax;                // 33 ints

    double *cur = bufA, *nxt = bufB;
    double sc = 1.0;        // alpha of the column held in cur = cur * sc
    int prevS = 0;
    long long prev_row = 0;
    double tmc[MAXT - 1];   // lane b < T1 keeps column b of the (a,b) table

    for (int b = 0; b < d.B; b++) {
        const int so = cp(d.soff)[b], S = cp(d.soff)[b + 1] - so;
        const int start = cp(d.bstart)[b], len = cp(d.blen)[b];
        const long long aoff = cp(d.aoff)[b];

        for (int x = tid; x < d.TC; x += NT) tm2[x] = gp(d.tm2)[(size_t)b * d.TC + x];
        for (int x = tid; x < S; x += NT) tperm[x] = gp(d.tperm)[so + x];
        for (int x = tid; x <= d.T; x += NT) tstart[x] = gp(d.tstart)[(size_t)b * (d.T + 1) + x];
#pragma unroll
        for (int a = 0; a < MAXT - 1; a++)
            tmc[a] = (a < T1 && lane < T1) ? gp(d.tm)[(size_t)b * TT + a * T1 + lane] : 0.0;

        int sb[R], sj1[R], slen[R], sbase[R], sstride[R];
        double se[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            int k = tid + NT * r;
            sb[r] = 0; sj1[r] = 0; slen[r] = 0; sbase[r] = 0; sstride[r] = 1; se[r] = 0.0;
            if (k < S) {
                int n = gp(d.snode)[so + k];
                int bt = gp(d.stime)[so + k];
                int lo = gp(d.age)[(size_t)b * N + n];
                int hi = gp(d.nhi)[(size_t)b * N + n];
                sb[r] = bt;
                sj1[r] = gp(d.nfirst)[(size_t)b * N + n];
                slen[r] = hi - lo + 1;
                sstride[r] = T1 - lo;
                sbase[r] = tc_base(lo, T1) + (bt - lo);
                se[r] = gp(d.einv)[so + k];
            }
        }
        int cls_next = (len > 1) ? cp(d.cls)[start + 1] : 0;

        // column `start`: prior (block 0) or switch from the previous block
        {
            const int cls = cp(d.cls)[start];
            if (b == 0) {
                for (int k = tid; k < S; k += NT) nxt[k] = gp(d.prior)[k];
            } else {
                const int rs = cp(d.srcs)[2 * b], cs = cp(d.srcs)[2 * b + 1];
                const int so1 = cp(d.soff)[b - 1];
                auto ptr = gp(d.invptr) + so + b;
                const double prs = (rs != -1) ? cur[rs] * sc : 0.0;
                const double pcs = (cs != -1) ? cur[cs] * sc : 0.0;
                for (int k = tid; k < S; k += NT) {
                    double acc = 0.0;
                    for (int x = ptr[k]; x < ptr[k + 1]; x++) {
                        int j = gp(d.invidx)[so1 + x];
                        acc += (cur[j] * sc) * gp(d.determp)[so1 + j];
                    }
                    double rr = gp(d.recombrow)[so + k], cr = gp(d.recoalrow)[so + k];
                    if (rs != -1 && rr > 0.0) acc += prs * rr;
                    if (cs != -1 && cr > 0.0) acc += pcs * cr;
                    double e = (cls == -1) ? gp(d.einv)[so + k]
                             : (cls == -2) ? 1.0 : gp(d.vemit)[(size_t)cls * d.Smax + k];
                    nxt[k] = acc * e;
                }
            }
            __syncthreads();
            if (b > 0)
                for (int k = tid; k < prevS; k += NT) gp(d.alpha)[prev_row + k] = cur[k] * sc;
            time_sums<NT>(nxt, tperm, tstart, T1, G);
            __syncthreads();
            double norm = 0.0;
            for (int a = 0; a < T1; a++) norm += G[a];
            if (b == 0) {
                sc = 1.0;   // the prior column is kept unnormalised
                if (tid == 0) gp(d.norms)[0] = (d.L == 1) ? norm : 1.0;
            } else {
                sc = 1.0 / norm;
                if (tid == 0) gp(d.norms)[start] = norm;
            }
            double *tmp = cur; cur = nxt; nxt = tmp;
        }

        for (int i = 1; i < len; i++) {
            const int col = start + i;
            const int cls = cls_next;
            if (i + 1 < len) cls_next = cp(d.cls)[col + 1];
            const long long row = aoff + (long long)(i - 1) * S;
            double tfb = 0.0;
#pragma unroll
            for (int a = 0; a < MAXT - 1; a++)
                if (a < T1) tfb += tmc[a] * G[a];
            tfb *= sc;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int k = tid + NT * r;
                double tfv = __shfl(tfb, sb[r], 64);
                if (k < S) {
                    gp(d.alpha)[row + k] = cur[k] * sc;
                    double s = 0.0;
                    const double *coef = tm2 + sbase[r];
                    const double *src = cur + sj1[r];
                    const int st = sstride[r];
                    for (int x = 0; x < slen[r]; x++) s += coef[x * st] * src[x];
                    double e = (cls == -1) ? se[r]
                             : (cls == -2) ? 1.0 : gp(d.vemit)[(size_t)cls * d.Smax + k];
                    nxt[k] = (tfv + sc * s) * e;
                }
            }
            __syncthreads();
            time_sums<NT>(nxt, tperm, tstart, T1, G);
            __syncthreads();
            double norm = 0.0;
            for (int a = 0; a < T1; a++) norm += G[a];
            sc = 1.0 / norm;
            if (tid == 0) gp(d.norms)[col] = norm;
            double *tmp = cur; cur = nxt; nxt = tmp;
        }
        prevS = S;
        prev_row = aoff + (long long)(len - 1) * S;
        __syncthreads();
    }
    for (int k = tid; k < prevS; k += NT) gp(d.alpha)[prev_row + k] = cur[k] * sc;
}

//===========================================================================
// traceback sampling helper

// first index j of w[0..n) with cumulative sum >= u * total (sample(),
// common.h:272-290); each lane owns a contiguous chunk of the weights
__device__ inline int wave_sample(const double *w, int n, double u)
{
    const int lane = threadIdx.x & 63;
    const int chunk = (n + 63) / 64;
    const int s0 = min(n, lane * chunk), e0 = min(n, s0 + chunk);
    double loc = 0.0;
    for (int j = s0; j < e0; j++) loc += w[j];
    double incl = loc;
    for (int off = 1; off < 64; off <<= 1) {
        double y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    double excl = __shfl_up(incl, 1, 64);
    if (lane == 0) excl = 0.0;
    double mine = excl + loc;
    double total = __shfl(mine, 63, 64);
    double pick = u * total;
    unsigned long long m = __ballot(mine >= pick && e0 > s0);
    if (m == 0) return n - 1;
    int owner = __ffsll((long long)m) - 1;
    int res = n - 1;
    if (lane == owner) {
        double p = 0.0;
        res = e0 - 1;
        for (int j = s0; j < e0; j++) {
            p += w[j];
            if (excl + p >= pick) { res = j; break; }
        }
    }
    return __shfl(res, owner, 64);
}

// column table layout shared by both tracebacks
struct TbTables {
    double *w, *tm, *tm2s;
    int *snode, *stime;
};

__device__ inline void tb_load_tables(const DevInst &d, int b, const TbTables &tt)
{
    const int lane = threadIdx.x;
    const int T1 = d.T - 1, TT = T1 * T1;
    const int so = cp(d.soff)[b], S = cp(d.soff)[b + 1] - so;
    for (int x = lane; x < TT; x += 64) tt.tm[x] = gp(d.tm)[(size_t)b * TT + x];
    for (int x = lane; x < d.TC; x += 64) tt.tm2s[x] = gp(d.tm2s)[(size_t)b * d.TC + x];
    for (int x = lane; x < S; x += 64) {
        tt.snode[x] = gp(d.snode)[so + x];
        tt.stime[x] = gp(d.stime)[so + x];
    }
}

// weight of the within-block transition j -> knext
__device__ inline double tb_trans(const DevInst &d, const TbTables &tt, int b, int j, int nk,
                                  int bk, int ck)
{
    const int T1 = d.T - 1;
    int a = tt.stime[j];
    if (tt.snode[j] == nk) return tt.tm2s[tc_base(ck, T1) + (a - ck) * (T1 - ck) + (bk - ck)];
    return tt.tm[a * T1 + bk];
}

// weight of the switch transition j (block b-1) -> knext (block b)
__device__ inline double tb_switch(const DevInst &d, int b, int j, int knext)
{
    const int so = cp(d.soff)[b], so1 = cp(d.soff)[b - 1];
    const int rs = cp(d.srcs)[2 * b], cs = cp(d.srcs)[2 * b + 1];
    if (j == cs) return gp(d.recoalrow)[so + knext];
    if (j == rs) return gp(d.recombrow)[so + knext];
    return (gp(d.determ)[so1 + j] == knext) ? gp(d.determp)[so1 + j] : 0.0;
}

//===========================================================================
// traceback over a column range, one wave; the next row is prefetched into
// registers while the current one is sampled

template <class PT, class X, int RT>
__device__ inline void load_row(PT row, int S, X (&reg)[RT])
{
    const int lane = threadIdx.x;
#pragma unroll
    for (int r = 0; r < RT; r++) {
        int j = lane + 64 * r;
        reg[r] = (j < S) ? row[j] : X(0);
    }
}

// Sample path[col] for col = hi down to lo given the state knext at column
// hi+1 (in block kblk).  mode 1 stops at the first column whose sample
// equals the stored path (the chains have coalesced) and returns it; else
// returns -1.  Uniform of column col: u[L-1-col].
template <class AT, int RT>
__device__ int tb_range(const DevInst &d, AW_GLOBAL const AT *alpha, const TbTables &tt,
                        int hi, int lo, int knext, int kblk, int mode)
{
    const int lane = threadIdx.x;
    const int L = d.L, N = d.N;
    auto u = cp(d.u);
    auto path = gp(d.path);
    int bc = find_owner(cp(d.bstart), d.B, hi);
    int bc_start = cp(d.bstart)[bc], bc_S = cp(d.soff)[bc + 1] - cp(d.soff)[bc];
    long long bc_aoff = cp(d.aoff)[bc];
    // the block before, fetched ahead of need
    int pb_start = 0, pb_S = 0;
    long long pb_aoff = 0;
    if (bc > 0) {
        pb_start = cp(d.bstart)[bc - 1];
        pb_S = cp(d.soff)[bc] - cp(d.soff)[bc - 1];
        pb_aoff = cp(d.aoff)[bc - 1];
    }
    int tblk = -1;
    AT pf[RT], pn[RT];
    if (hi >= lo) load_row(alpha + bc_aoff + (long long)(hi - bc_start) * bc_S, bc_S, pn);
    int result = -1;
    for (int col = hi; col >= lo; col--) {
#pragma unroll
        for (int r = 0; r < RT; r++) pf[r] = pn[r];
        const int oldp = (mode == 1) ? path[col] : -1;
        if (col - 1 >= lo) {
            if (col - 1 >= bc_start)
                load_row(alpha + bc_aoff + (long long)(col - 1 - bc_start) * bc_S, bc_S, pn);
            else
                load_row(alpha + pb_aoff + (long long)(col - 1 - pb_start) * pb_S, pb_S, pn);
        }
        if (bc != kblk) {
            // col is the last column of block bc = kblk-1: switch into kblk
#pragma unroll
            for (int r = 0; r < RT; r++) {
                int j = lane + 64 * r;
                if (j < bc_S) tt.w[j] = (double)pf[r] * tb_switch(d, kblk, j, knext);
            }
        } else {
            if (tblk != bc) {
                tb_load_tables(d, bc, tt);
                tblk = bc;
                __syncthreads();
            }
            const int nk = tt.snode[knext], bk = tt.stime[knext];
            const int ck = gp(d.age)[(size_t)bc * N + nk];
#pragma unroll
            for (int r = 0; r < RT; r++) {
                int j = lane + 64 * r;
                if (j < bc_S) tt.w[j] = (double)pf[r] * tb_trans(d, tt, bc, j, nk, bk, ck);
            }
        }
        __syncthreads();
        knext = wave_sample(tt.w, bc_S, u[L - 1 - col]);
        kblk = bc;
        if (mode == 1 && oldp == knext) {
            result = col;
            break;
        }
        if (lane == 0) path[col] = knext;
        __syncthreads();
        if (col - 1 < bc_start && col - 1 >= lo) {
            bc--;
            bc_start = pb_start;
            bc_S = pb_S;
            bc_aoff = pb_aoff;
            if (bc > 0) {
                pb_start = cp(d.bstart)[bc - 1];
                pb_S = cp(d.soff)[bc] - cp(d.soff)[bc - 1];
                pb_aoff = cp(d.aoff)[bc - 1];
            }
        }
    }
    return result;
}

// sample the last column of the instance from its forward column alone
template <class AT, int RT>
__device__ int tb_sample_column(const DevInst &d, AW_GLOBAL const AT *alpha,
                                const TbTables &tt, int col, double uval, int *blk)
{
    const int lane = threadIdx.x;
    const int b = find_owner(cp(d.bstart), d.B, col);
    const int S = cp(d.soff)[b + 1] - cp(d.soff)[b];
    AT pf[RT];
    load_row(alpha + cp(d.aoff)[b] + (long long)(col - cp(d.bstart)[b]) * S, S, pf);
#pragma unroll
    for (int r = 0; r < RT; r++)
        if (lane + 64 * r < S) tt.w[lane + 64 * r] = (double)pf[r];
    __syncthreads();
    int k = wave_sample(tt.w, S, uval);
    __syncthreads();
    *blk = b;
    return k;
}

__device__ inline TbTables tb_layout(double *dyn, const DevInst &d)
{
    const int T1 = d.T - 1;
    TbTables tt;
    tt.w = dyn;
    tt.tm = tt.w + d.Smax;
    tt.tm2s = tt.tm + T1 * T1;
    tt.snode = (int *)(tt.tm2s + d.TC);
    tt.stime = tt.snode + d.Smax;
    return tt;
}

// plain FP64 traceback of whole instances
template <int RT>
__global__ __launch_bounds__(64) void k_traceback64_b(const DevInst *__restrict__ D,
                                                      const int *__restrict__ list)
{
    extern __shared__ double dyn[];
    const DevInst &d = D[list[blockIdx.x]];
    TbTables tt = tb_layout(dyn, d);
    const int L = d.L;
    int kblk;
    int k = tb_sample_column<double, RT>(d, gp(d.alpha), tt, L - 1, cp(d.u)[0], &kblk);
    if (threadIdx.x == 0) gp(d.path)[L - 1] = k;
    __syncthreads();
    tb_range<double, RT>(d, gp(d.alpha), tt, L - 2, 0, k, kblk, 0);
}

//===========================================================================
// FP32 column-tiled forward

struct Chunk {
    int inst, c0, c1, b0;   // columns [c0, c1) of instance inst; block of c0
    int first, last;        // first / last tile of its instance
};

// per-lane view of one block for the FP32 forward
template <int RM>
struct Lane32 {
    int key;
    int slot[RM], j1[RM], ln[RM], cb[RM], cs[RM];
    float ei[RM], u[RM];
    float tmc[MAXT - 1];
    bool gend;
};

template <int RM>
__device__ inline void lane32_setup(const DevInst &d, int b, Lane32<RM> &ls, float *t2, float *G)
{
    const int lane = threadIdx.x;
    const int T1 = d.T - 1, TT = T1 * T1, N = d.N;
    const int so = cp(d.soff)[b];
    // the instance's slot table has d.RM <= RM slots per lane (the kernel is
    // instantiated for the largest RM of the batch); the rest stay empty
    const int IR = d.RM;
    ls.key = gp(d.lkey)[(size_t)b * 64 + lane];
    int knext = __shfl_down(ls.key, 1, 64);
    ls.gend = (ls.key < T1) && (lane == 63 || knext != ls.key);
#pragma unroll
    for (int r = 0; r < RM; r++) {
        int k = (r < IR) ? gp(d.lslot)[((size_t)b * 64 + lane) * IR + r] : -1;
        ls.slot[r] = k;
        ls.j1[r] = 0; ls.ln[r] = 0; ls.cb[r] = 0; ls.cs[r] = 1; ls.ei[r] = 0.f; ls.u[r] = 0.f;
        if (k >= 0) {
            int n = gp(d.snode)[so + k];
            int lo = gp(d.age)[(size_t)b * N + n];
            int hi = gp(d.nhi)[(size_t)b * N + n];
            ls.j1[r] = gp(d.nfirst)[(size_t)b * N + n];
            ls.ln[r] = hi - lo + 1;
            ls.cs[r] = T1 - lo;
            ls.cb[r] = tc_base(lo, T1) + (ls.key - lo);
            ls.ei[r] = gp(d.einv32)[so + k];
        }
    }
#pragma unroll
    for (int a = 0; a < MAXT - 1; a++)
        ls.tmc[a] = (a < T1 && ls.key < T1) ? gp(d.tm32)[(size_t)b * TT + a * T1 + ls.key] : 0.f;
    for (int x = lane; x < d.TC; x += 64) t2[x] = gp(d.tm2_32)[(size_t)b * d.TC + x];
    for (int a = lane; a < 32; a += 64) G[a] = 0.f;
}

// per-time sums of the lanes' partials: segmented scan over lanes grouped by key
__device__ inline void seg_reduce_store(float part, int key, bool gend, float *G)
{
    const int lane = threadIdx.x;
    float v = part;
    for (int off = 1; off < 64; off <<= 1) {
        float vu = __shfl_up(v, off, 64);
        int ku = __shfl_up(key, off, 64);
        if (lane >= off && ku == key) v += vu;
    }
    if (gend) G[key] = v;
}

template <int RM>
__global__ __launch_bounds__(64) void k_forward32(const DevInst *__restrict__ D,
                                                  const Chunk *__restrict__ chunks, int mode,
                                                  const float *__restrict__ snap, int sstride,
                                                  int *__restrict__ flags, float tol_rel,
                                                  float tol_abs)
{
    extern __shared__ float sm32[];
    const Chunk ch = chunks[blockIdx.x];
    const DevInst &d = D[ch.inst];
    const int lane = threadIdx.x;
    const int T1 = d.T - 1, Smax = d.Smax, L = d.L;
    auto cls_a = cp(d.cls);
    auto vemit = gp(d.vemit);
    auto alpha = gp(d.alpha32);
    auto norms = gp(d.norms);
    float *A = sm32, *Bn = sm32 + Smax;
    float *t2 = Bn + Smax;
    float *G = t2 + d.TC;
    Lane32<RM> ls;

    int b = ch.b0;
    int so = cp(d.soff)[b], S = cp(d.soff)[b + 1] - so;
    int bstart = cp(d.bstart)[b], bend = bstart + cp(d.blen)[b];
    long long aoff = cp(d.aoff)[b];
    float sc = 1.f, tfb = 0.f;
    int col = ch.c0;

    // column c0 - 1 into A (uniform guess, or the true column in a fix-up pass)
    if (col > 0) {
        const int bp = (col - 1 >= bstart) ? b : b - 1;
        const int Sp = cp(d.soff)[bp + 1] - cp(d.soff)[bp];
        if (mode == 0) {
            for (int k = lane; k < Sp; k += 64) A[k] = 1.f;
            sc = 1.f / Sp;
        } else {
            for (int k = lane; k < Sp; k += 64) A[k] = snap[(size_t)blockIdx.x * sstride + k];
            sc = 1.f;
        }
    }
    lane32_setup<RM>(d, b, ls, t2, G);
    wave_lds_sync();
    if (col > 0 && col != bstart) {
        float part = 0.f;
#pragma unroll
        for (int r = 0; r < RM; r++)
            if (ls.slot[r] >= 0) part += A[ls.slot[r]];
        seg_reduce_store(part, ls.key, ls.gend, G);
        wave_lds_sync();
        float t = 0.f;
#pragma unroll
        for (int a = 0; a < MAXT - 1; a++)
            if (a < T1) t += ls.tmc[a] * G[a];
        tfb = t * sc;
    }
    // next block's bounds, fetched ahead of need
    int nb_S = 0, nb_so = 0, nb_len = 0;
    long long nb_aoff = 0;
    if (b + 1 < d.B) {
        nb_so = cp(d.soff)[b + 1];
        nb_S = cp(d.soff)[b + 2] - nb_so;
        nb_len = cp(d.blen)[b + 1];
        nb_aoff = cp(d.aoff)[b + 1];
    }

    int cls_next = cls_a[col];
    int converged = -1;
    for (; col < ch.c1; col++) {
        if (col == bend) {
            b++;
            so = nb_so;
            S = nb_S;
            bstart = bend;
            bend = bstart + nb_len;
            aoff = nb_aoff;
            lane32_setup<RM>(d, b, ls, t2, G);
            if (b + 1 < d.B) {
                nb_so = cp(d.soff)[b + 1];
                nb_S = cp(d.soff)[b + 2] - nb_so;
                nb_len = cp(d.blen)[b + 1];
                nb_aoff = cp(d.aoff)[b + 1];
            }
            wave_lds_sync();
        }
        const int cls = cls_next;
        if (col + 1 < ch.c1) cls_next = cls_a[col + 1];
        const long long row = aoff + (long long)(col - bstart) * S;
        float ov[RM];
        if (mode == 1) {
#pragma unroll
            for (int r = 0; r < RM; r++) ov[r] = (ls.slot[r] >= 0) ? alpha[row + ls.slot[r]] : 0.f;
        }
        float part = 0.f;
        const bool is_prior = (col == 0);
        const bool is_switch = (col > 0 && col == bstart);
        int rs = -1, csrc = -1, so1 = 0;
        float prs = 0.f, pcs = 0.f;
        AW_GLOBAL const int *ptr = nullptr;
        if (is_switch) {
            rs = cp(d.srcs)[2 * b];
            csrc = cp(d.srcs)[2 * b + 1];
            so1 = cp(d.soff)[b - 1];
            ptr = gp(d.invptr) + so + b;
            prs = (rs != -1) ? A[rs] * sc : 0.f;
            pcs = (csrc != -1) ? A[csrc] * sc : 0.f;
        }
#pragma unroll
        for (int r = 0; r < RM; r++) {
            const int k = ls.slot[r];
            if (k < 0) continue;
            float v;
            if (is_prior) {
                v = (float)gp(d.prior)[k];
            } else {
                float e = (cls == -1) ? ls.ei[r]
                        : (cls == -2) ? 1.f : (float)vemit[(size_t)cls * Smax + k];
                if (is_switch) {
                    float acc = 0.f;
                    for (int x = ptr[k]; x < ptr[k + 1]; x++) {
                        int j = gp(d.invidx)[so1 + x];
                        acc += (A[j] * sc) * (float)gp(d.determp)[so1 + j];
                    }
                    float rr = (float)gp(d.recombrow)[so + k], cr = (float)gp(d.recoalrow)[so + k];
                    if (rs != -1 && rr > 0.f) acc += prs * rr;
                    if (csrc != -1 && cr > 0.f) acc += pcs * cr;
                    v = acc * e;
                } else {
                    // four independent partial sums per pass keep the LDS
                    // loads of a slot in flight together
                    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
                    const float *cf = t2 + ls.cb[r];
                    const float *src = A + ls.j1[r];
                    const int st = ls.cs[r], ln = ls.ln[r];
                    for (int x = 0; x < ln; x += 4) {
                        float c0 = cf[x * st], c1 = cf[(x + 1) * st];
                        float c2 = cf[(x + 2) * st], c3 = cf[(x + 3) * st];
                        float v0 = src[x], v1 = src[x + 1], v2 = src[x + 2], v3 = src[x + 3];
                        s0 += c0 * v0;
                        s1 += (x + 1 < ln) ? c1 * v1 : 0.f;
                        s2 += (x + 2 < ln) ? c2 * v2 : 0.f;
                        s3 += (x + 3 < ln) ? c3 * v3 : 0.f;
                    }
                    v = (tfb + sc * ((s0 + s1) + (s2 + s3))) * e;
                }
            }
            ls.u[r] = v;
            Bn[k] = v;
            part += v;
        }
        seg_reduce_store(part, ls.key, ls.gend, G);
        wave_lds_sync();
        float norm = 0.f;
#pragma unroll
        for (int a = 0; a < MAXT - 1; a++)
            if (a < T1) norm += G[a];
        float nsc;
        double nrm;
        if (is_prior) {
            nsc = 1.f;
            nrm = (L == 1) ? (double)norm : 1.0;
        } else {
            nsc = 1.f / norm;
            nrm = (double)norm;
        }
        float t = 0.f;
#pragma unroll
        for (int a = 0; a < MAXT - 1; a++)
            if (a < T1) t += ls.tmc[a] * G[a];
        tfb = t * nsc;
        sc = nsc;
        if (lane == 0) norms[col] = nrm;
        if (mode == 1) {
            bool ok = true;
#pragma unroll
            for (int r = 0; r < RM; r++) {
                if (ls.slot[r] < 0) continue;
                float nv = ls.u[r] * sc;
                if (!(fabsf(nv - ov[r]) <= tol_rel * fabsf(ov[r]) + tol_abs)) ok = false;
            }
            if (__all(ok)) {
                converged = col;
                break;
            }
        }
#pragma unroll
        for (int r = 0; r < RM; r++)
            if (ls.slot[r] >= 0) alpha[row + ls.slot[r]] = ls.u[r] * sc;
        float *tmp = A; A = Bn; Bn = tmp;
    }
    if (mode == 1 && lane == 0) flags[blockIdx.x] = converged;
}

// boundary column of every tile after the first: its previous column
__global__ __launch_bounds__(64) void k_fw_snapshot(const DevInst *__restrict__ D,
                                                    const Chunk *__restrict__ chunks,
                                                    float *__restrict__ snap, int sstride)
{
    const Chunk ch = chunks[blockIdx.x];
    if (ch.first) return;
    const DevInst &d = D[ch.inst];
    const int col = ch.c0 - 1;
    const int b = find_owner(cp(d.bstart), d.B, col);
    const int S = cp(d.soff)[b + 1] - cp(d.soff)[b];
    const long long row = cp(d.aoff)[b] + (long long)(col - cp(d.bstart)[b]) * S;
    for (int k = threadIdx.x; k < S; k += 64) snap[(size_t)blockIdx.x * sstride + k] = gp(d.alpha32)[row + k];
}

//===========================================================================
// FP32 column-tiled traceback (FP64 weights and CDF)

template <int RT>
__global__ __launch_bounds__(64) void k_traceback32(const DevInst *__restrict__ D,
                                                    const Chunk *__restrict__ chunks, int mode,
                                                    const int *__restrict__ snap,
                                                    int *__restrict__ flags)
{
    extern __shared__ double dyn[];
    const Chunk ch = chunks[blockIdx.x];
    const DevInst &d = D[ch.inst];
    TbTables tt = tb_layout(dyn, d);
    const int L = d.L;
    int knext, kblk, hi = ch.c1 - 1;
    if (ch.last) {
        knext = tb_sample_column<float, RT>(d, gp(d.alpha32), tt, L - 1, cp(d.u)[0], &kblk);
        if (threadIdx.x == 0) gp(d.path)[L - 1] = knext;
        __syncthreads();
        hi = L - 2;
    } else if (mode == 0) {
        // speculative state at column c1: sample that column alone
        knext = tb_sample_column<float, RT>(d, gp(d.alpha32), tt, ch.c1, cp(d.u)[L - 1 - ch.c1], &kblk);
    } else {
        knext = snap[blockIdx.x];
        kblk = find_owner(cp(d.bstart), d.B, ch.c1);
    }
    int coal = tb_range<float, RT>(d, gp(d.alpha32), tt, hi, ch.c0, knext, kblk, mode);
    if (mode == 1 && threadIdx.x == 0) flags[blockIdx.x] = coal;
}

__global__ __launch_bounds__(64) void k_tb_snapshot(const DevInst *__restrict__ D,
                                                    const Chunk *__restrict__ chunks,
                                                    int *__restrict__ snap)
{
    const Chunk ch = chunks[blockIdx.x];
    if (ch.last || threadIdx.x != 0) return;
    snap[blockIdx.x] = D[ch.inst].path[ch.c1];
}

// log-likelihood: sum of log column normalisers
__global__ __launch_bounds__(256) void k_lnl_b(const DevInst *__restrict__ D)
{
    __shared__ double part[256];
    const DevInst &d = D[blockIdx.x];
    double s = 0.0;
    for (int c = threadIdx.x; c < d.L; c += 256) s += log(gp(d.norms)[c]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *d.lnl = part[0];
}

// FP64 copy of an FP32 table
__global__ void k_widen(const float *__restrict__ src, double *__restrict__ dst, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}

//===========================================================================
// host driver

namespace {

struct Ctx {
    bool init = false;
    int dev = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    hipEvent_t ev[8];
};

Ctx &ctx()
{
    static Ctx c;
    return c;
}

int ctx_init()
{
    Ctx &c = ctx();
    if (c.init) return 0;
    HIP_OK(hipGetDevice(&c.dev));
    HIP_OK(hipGetDeviceProperties(&c.prop, c.dev));
    HIP_OK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    for (auto &e : c.ev) HIP_OK(hipEventCreate(&e));
    c.init = true;
    return 0;
}

// grow-only device buffer
struct DevBuf {
    char *p = nullptr;
    size_t cap = 0;
    int ensure(size_t n)
    {
        if (n <= cap) return 0;
        if (p) HIP_OK(hipFree(p));
        p = nullptr;
        cap = 0;
        HIP_OK(hipMalloc(&p, n));
        cap = n;
        return 0;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// bump allocator over a size plan: first pass counts, second pass assigns
struct Plan {
    size_t used = 0;
    char *base = nullptr;
    template <class X>
    X *take(size_t n)
    {
        size_t off = (used + 255) & ~size_t(255);
        used = off + std::max<size_t>(n, 1) * sizeof(X);
        return base ? reinterpret_cast<X *>(base + off) : nullptr;
    }
};

}  // namespace

struct Batch::Impl {
    std::vector<HostInstance> hosts;
    std::vector<std::vector<double>> us, priors;
    std::vector<std::vector<int>> Sv;
    std::vector<std::vector<int>> soff;
    std::vector<std::vector<long long>> aoff;
    std::vector<DevInst> h;
    std::vector<long long> W;
    std::vector<double> lnl;
    std::vector<int> nvar;
    std::vector<int> RMv;
    std::vector<Chunk> chunks;
    DevBuf in, work, tab, aux, vbuf;
    DevInst *dD = nullptr;
    long long *d_coloff = nullptr;
    int *d_blkoff = nullptr, *d_voff = nullptr, *d_nvar = nullptr, *d_list = nullptr;
    int *d_flags = nullptr, *d_ovf = nullptr, *d_tsnap = nullptr;
    float *d_fsnap = nullptr;
    Chunk *d_chunks = nullptr;
    bool uploaded = false;
    Precision prec_done = Precision::FP64;

    ~Impl()
    {
        in.release();
        work.release();
        tab.release();
        aux.release();
        vbuf.release();
    }
};

Batch::Batch() : impl_(new Impl) {}
Batch::~Batch() { delete impl_; }
int Batch::size() const { return (int)impl_->hosts.size(); }
double Batch::lnl(int i) const { return impl_->lnl[i]; }
long long Batch::W(int i) const { return impl_->W[i]; }

int Batch::add(const HostInstance &hi, const double *u, const double *prior)
{
    Impl &m = *impl_;
    m.hosts.push_back(hi);
    m.us.emplace_back();
    if (u) m.us.back().assign(u, u + hi.ft.seqlen);
    m.priors.emplace_back();
    std::vector<int> Sv;
    long long W = instance_states(hi.ft, hi.model.T, &Sv);
    if (prior) m.priors.back().assign(prior, prior + Sv[0]);
    m.W.push_back(W);
    m.Sv.push_back(Sv);
    m.uploaded = false;
    return (int)m.hosts.size() - 1;
}

// Inputs: trees, sequences, uniforms and the model of every instance.
int Batch::upload()
{
    if (ctx_init()) return 1;
    Ctx &c = ctx();
    Impl &m = *impl_;
    const int n = (int)m.hosts.size();
    m.h.assign(n, DevInst{});
    m.soff.assign(n, {});
    m.aoff.assign(n, {});
    m.lnl.assign(n, 0.0);
    m.nvar.assign(n, 0);
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.in.p : nullptr;
        for (int i = 0; i < n; i++) {
            const HostInstance &hi = m.hosts[i];
            const FlatTrees &ft = hi.ft;
            const int B = ft.nblocks, N = ft.nnodes, L = ft.seqlen, T = hi.model.T;
            DevInst &d = m.h[i];
            if (pass == 0) {
                if (N > MAXN || T > MAXT) {
                    fprintf(stderr, "argweaver_amd: instance exceeds nnodes<=%d / ntimes<=%d\n",
                            MAXN, MAXT);
                    return 1;
                }
                std::vector<int> &so = m.soff[i];
                std::vector<long long> &ao = m.aoff[i];
                so.assign(B + 1, 0);
                ao.assign(B + 1, 0);
                int Smax = 1;
                for (int b = 0; b < B; b++) {
                    so[b + 1] = so[b] + m.Sv[i][b];
                    ao[b + 1] = ao[b] + (long long)m.Sv[i][b] * ft.len[b];
                    Smax = std::max(Smax, m.Sv[i][b]);
                }
                d.T = T; d.N = N; d.B = B; d.L = L; d.nrows = (int)hi.rows.size();
                d.Smax = Smax; d.TC = tc_base(T - 1, T - 1); d.W = m.W[i];
            }
            d.model = p.take<Model>(1);
            d.parent = p.take<int>((size_t)B * N);
            d.c0 = p.take<int>((size_t)B * N);
            d.c1 = p.take<int>((size_t)B * N);
            d.age = p.take<int>((size_t)B * N);
            d.map = p.take<int>((size_t)B * N);
            d.spr = p.take<int>(4 * (size_t)B);
            d.bstart = p.take<int>(B);
            d.blen = p.take<int>(B);
            d.soff = p.take<int>(B + 1);
            d.aoff = p.take<long long>(B + 1);
            d.seq = p.take<char>((size_t)d.nrows * L);
            d.u = p.take<double>(L);
            if (pass == 1) {
                hipStream_t s = c.stream;
                HIP_OK(hipMemcpyAsync((void *)d.model, &hi.model, sizeof(Model), hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.parent, ft.parent.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.c0, ft.child0.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.c1, ft.child1.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.age, ft.age.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.map, ft.mapping.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.spr, ft.spr.data(), sizeof(int) * 4 * B, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.bstart, ft.start.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.blen, ft.len.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.soff, m.soff[i].data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync((void *)d.aoff, m.aoff[i].data(), sizeof(long long) * (B + 1), hipMemcpyHostToDevice, s));
                for (int r = 0; r < d.nrows; r++)
                    HIP_OK(hipMemcpyAsync((char *)d.seq + (size_t)r * L, hi.rows[r], L, hipMemcpyHostToDevice, s));
                if (!m.us[i].empty())
                    HIP_OK(hipMemcpyAsync((void *)d.u, m.us[i].data(), sizeof(double) * L, hipMemcpyHostToDevice, s));
            }
        }
        if (pass == 0 && m.in.ensure(p.used + 4096)) return 1;
    }
    HIP_OK(hipStreamSynchronize(c.stream));
    // host rows may be released now
    for (auto &hi : m.hosts) hi.rows.clear();
    m.uploaded = true;
    return 0;
}

static size_t lds_fw64(const DevInst &d)
{
    return sizeof(double) * (2 * (size_t)d.Smax + d.TC + 32) + sizeof(int) * ((size_t)d.Smax + 33);
}
static size_t lds_tb(const DevInst &d)
{
    const int T1 = d.T - 1;
    return sizeof(double) * ((size_t)d.Smax + T1 * T1 + d.TC) + sizeof(int) * (2 * (size_t)d.Smax);
}
static size_t lds_fw32(const DevInst &d)
{
    // + 4*(T-1) floats: the 4-way gather reads up to three strides past a run
    return sizeof(float) * (2 * (size_t)d.Smax + d.TC + 32 + 4 * (d.T - 1) + 8);
}

int Batch::run(const BatchOptions &opt, RunStats *st)
{
    if (ctx_init()) return 1;
    Ctx &c = ctx();
    Impl &m = *impl_;
    if (!m.uploaded && upload()) return 1;
    const int n = (int)m.hosts.size();
    hipStream_t s = c.stream;
    const bool fp32 = opt.prec == Precision::FP32;

    // classification and variant numbering: one host sync for the counts
    std::vector<long long> coloff(n + 1, 0);
    std::vector<int> blkoff(n + 1, 0);
    for (int i = 0; i < n; i++) {
        coloff[i + 1] = coloff[i] + m.h[i].L;
        blkoff[i + 1] = blkoff[i] + m.h[i].B;
    }
    // work buffers, first pass sized without the variant rows
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.work.p : nullptr;
        for (int i = 0; i < n; i++) {
            DevInst &d = m.h[i];
            const int B = d.B, N = d.N, L = d.L, T = d.T, T1 = T - 1;
            const int Stot = m.soff[i][B];
            d.vflag = p.take<int>(L);
            d.cls = p.take<int>(L);
            d.vcol = p.take<int>(L);
            d.snode = p.take<int>(Stot);
            d.stime = p.take<int>(Stot);
            d.nfirst = p.take<int>((size_t)B * N);
            d.nhi = p.take<int>((size_t)B * N);
            d.tperm = p.take<int>(Stot);
            d.tstart = p.take<int>((size_t)B * (T + 1));
            d.tv = p.take<double>((size_t)B * 9 * T);
            d.einv = p.take<double>(Stot);
            d.eterm = p.take<double>((size_t)Stot * 6);
            d.determ = p.take<int>(Stot);
            d.determp = p.take<double>(Stot);
            d.recombrow = p.take<double>(Stot);
            d.recoalrow = p.take<double>(Stot);
            d.srcs = p.take<int>(2 * (size_t)B);
            d.invptr = p.take<int>((size_t)Stot + B);
            d.invidx = p.take<int>(Stot);
            d.prior = p.take<double>(d.Smax);
            d.order = p.take<int>((size_t)B * N);
            d.roots = p.take<int>(B);
            d.nmut = p.take<double>((size_t)B * N * 2);
            d.norms = p.take<double>(L);
            d.path = p.take<int>(L);
            d.lnl = p.take<double>(1);
            (void)T1;
        }
        if (pass == 0 && m.work.ensure(p.used + 4096)) return 1;
    }
    // transition tables and the forward table
    std::vector<int> RMv(n, 8);
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.tab.p : nullptr;
        for (int i = 0; i < n; i++) {
            DevInst &d = m.h[i];
            const int B = d.B, T1 = d.T - 1;
            const int Stot = m.soff[i][B];
            d.tm = p.take<double>((size_t)B * T1 * T1);
            d.tm2s = p.take<double>((size_t)B * d.TC);
            if (fp32) {
                // worst case slots per lane: all states in the largest bin
                // sum_a ceil(cnt_a/R) <= S/R + T1 fits 64 lanes when R >= S/(64-T1)
                int RM = 8;
                while (RM < 64 && (long long)RM * (64 - T1) < d.Smax) RM *= 2;
                RMv[i] = RM;
                d.RM = RM;
                d.tm2 = p.take<double>((size_t)B * d.TC);
                d.tm32 = p.take<float>((size_t)B * T1 * T1);
                d.tm2_32 = p.take<float>((size_t)B * d.TC);
                d.einv32 = p.take<float>(Stot);
                d.lslot = p.take<int>((size_t)B * 64 * RM);
                d.lkey = p.take<int>((size_t)B * 64);
                d.alpha32 = p.take<float>(d.W);
                d.alpha = nullptr;
            } else {
                d.tm2 = p.take<double>((size_t)B * d.TC);
                d.alpha = p.take<double>(d.W);
                d.alpha32 = nullptr;
            }
        }
        if (pass == 0 && m.tab.ensure(p.used + 4096)) return 1;
    }
    // auxiliary arrays
    {
        size_t nchunk_max = 0;
        std::vector<Chunk> &chs = m.chunks;
        chs.clear();
        for (int i = 0; i < n; i++) {
            const DevInst &d = m.h[i];
            const int tile = (fp32 && opt.chunk_cols > 0) ? opt.chunk_cols : d.L;
            const int nt = (d.L + tile - 1) / tile;
            const std::vector<int> &bs = m.hosts[i].ft.start;
            int bcur = 0;
            for (int t = 0; t < nt; t++) {
                Chunk k;
                k.inst = i;
                k.c0 = t * tile;
                k.c1 = std::min(d.L, (t + 1) * tile);
                while (bcur + 1 < d.B && bs[bcur + 1] <= k.c0) bcur++;
                k.b0 = bcur;
                k.first = (t == 0);
                k.last = (t == nt - 1);
                chs.push_back(k);
            }
        }
        nchunk_max = chs.size();
        int Smax_all = 1;
        for (auto &d : m.h) Smax_all = std::max(Smax_all, d.Smax);
        Plan p;
        for (int pass = 0; pass < 2; pass++) {
            p = Plan();
            p.base = pass ? m.aux.p : nullptr;
            m.dD = p.take<DevInst>(n);
            m.d_coloff = p.take<long long>(n + 1);
            m.d_blkoff = p.take<int>(n + 1);
            m.d_voff = p.take<int>(n + 1);
            m.d_nvar = p.take<int>(n);
            m.d_list = p.take<int>(n);
            m.d_flags = p.take<int>(nchunk_max);
            m.d_ovf = p.take<int>(1);
            m.d_tsnap = p.take<int>(nchunk_max);
            m.d_chunks = p.take<Chunk>(nchunk_max);
            m.d_fsnap = p.take<float>(fp32 ? nchunk_max * (size_t)Smax_all : 1);
            if (pass == 0 && m.aux.ensure(p.used + 4096)) return 1;
        }
    }

    (void)hipEventRecord(c.ev[0], s);
    HIP_OK(hipMemcpyAsync(m.d_coloff, coloff.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(m.d_blkoff, blkoff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(m.dD, m.h.data(), sizeof(DevInst) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_classify_b, dim3((unsigned)((coloff[n] + 255) / 256)), dim3(256), 0, s,
                       m.dD, m.d_coloff, n);
    hipLaunchKernelGGL(k_scan_b, dim3(n), dim3(1024), 0, s, m.dD, m.d_nvar);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(m.nvar.data(), m.d_nvar, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));

    // variant rows, allocated now that the counts are known
    std::vector<int> voff(n + 1, 0);
    {
        Plan p;
        for (int pass = 0; pass < 2; pass++) {
            p = Plan();
            p.base = pass ? m.vbuf.p : nullptr;
            for (int i = 0; i < n; i++) {
                m.h[i].nvar = m.nvar[i];
                m.h[i].vemit = p.take<double>((size_t)std::max(m.nvar[i], 1) * m.h[i].Smax);
                voff[i + 1] = voff[i] + m.nvar[i];
            }
            if (pass == 0 && m.vbuf.ensure(p.used + 4096)) return 1;
        }
    }
    HIP_OK(hipMemcpyAsync(m.d_voff, voff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(m.dD, m.h.data(), sizeof(DevInst) * n, hipMemcpyHostToDevice, s));

    (void)hipEventRecord(c.ev[1], s);
    hipLaunchKernelGGL(k_prep_b, dim3(blkoff[n]), dim3(256), 0, s, m.dD, m.d_blkoff, n);
    HIP_OK(hipGetLastError());
    for (int i = 0; i < n; i++)
        if (!m.priors[i].empty())
            HIP_OK(hipMemcpyAsync(m.h[i].prior, m.priors[i].data(), sizeof(double) * m.priors[i].size(),
                                  hipMemcpyHostToDevice, s));
    if (fp32) {
        HIP_OK(hipMemsetAsync(m.d_ovf, 0, sizeof(int), s));
        hipLaunchKernelGGL(k_prep32_b, dim3(blkoff[n]), dim3(64), 0, s, m.dD, m.d_blkoff, n, m.d_ovf);
        HIP_OK(hipGetLastError());
    }
    (void)hipEventRecord(c.ev[2], s);
    if (voff[n] > 0) hipLaunchKernelGGL(k_var_emit_b, dim3(voff[n]), dim3(64), 0, s, m.dD, m.d_voff, n);
    HIP_OK(hipGetLastError());
    (void)hipEventRecord(c.ev[3], s);

    RunStats rs;
    if (!fp32) {
        // group instances by the slot count of the plain forward
        std::vector<int> lists[6];
        for (int i = 0; i < n; i++) {
            const int S = m.h[i].Smax;
            int g = S <= 256 ? 0 : S <= 512 ? 1 : S <= 1024 ? 2 : S <= 2048 ? 3 : S <= 4096 ? 4 : 5;
            if (g == 5 || lds_fw64(m.h[i]) > (size_t)c.prop.sharedMemPerBlock) {
                fprintf(stderr, "argweaver_amd: %d states per block exceed the FP64 kernel\n", S);
                return 1;
            }
            lists[g].push_back(i);
        }
        std::vector<int> flat;
        for (int g = 0; g < 5; g++) flat.insert(flat.end(), lists[g].begin(), lists[g].end());
        HIP_OK(hipMemcpyAsync(m.d_list, flat.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        int at = 0;
        for (int g = 0; g < 5; g++) {
            const int cnt = (int)lists[g].size();
            if (!cnt) continue;
            size_t sh = 0;
            for (int i : lists[g]) sh = std::max(sh, lds_fw64(m.h[i]));
            int *lp = m.d_list + at;
#define FW64(NT, R)                                                                          \
    do {                                                                                     \
        auto kern = k_forward64_b<NT, R>;                                                    \
        HIP_OK(hipFuncSetAttribute((const void *)kern,                                       \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));    \
        hipLaunchKernelGGL(kern, dim3(cnt), dim3(NT), sh, s, m.dD, lp);                      \
    } while (0)
            if (g == 0) FW64(256, 1);
            else if (g == 1) FW64(256, 2);
            else if (g == 2) FW64(256, 4);
            else if (g == 3) FW64(256, 8);
            else FW64(256, 16);
#undef FW64
            HIP_OK(hipGetLastError());
            at += cnt;
        }
        (void)hipEventRecord(c.ev[4], s);
        hipLaunchKernelGGL(k_lnl_b, dim3(n), dim3(256), 0, s, m.dD);
        if (opt.traceback) {
            at = 0;
            for (int g = 0; g < 5; g++) {
                const int cnt = (int)lists[g].size();
                if (!cnt) continue;
                size_t sh = 0;
                for (int i : lists[g]) sh = std::max(sh, lds_tb(m.h[i]));
                int *lp = m.d_list + at;
#define TB64(RT)                                                                             \
    do {                                                                                     \
        auto kern = k_traceback64_b<RT>;                                                     \
        HIP_OK(hipFuncSetAttribute((const void *)kern,                                       \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));    \
        hipLaunchKernelGGL(kern, dim3(cnt), dim3(64), sh, s, m.dD, lp);                      \
    } while (0)
                if (g == 0) TB64(4);
                else if (g == 1) TB64(8);
                else if (g == 2) TB64(16);
                else if (g == 3) TB64(32);
                else TB64(64);
#undef TB64
                HIP_OK(hipGetLastError());
                at += cnt;
            }
        }
        (void)hipEventRecord(c.ev[5], s);
    } else {
        // FP32 column tiles
        int ovf = 0;
        HIP_OK(hipMemcpyAsync(&ovf, m.d_ovf, sizeof(int), hipMemcpyDeviceToHost, s));
        const int nch = (int)m.chunks.size();
        HIP_OK(hipMemcpyAsync(m.d_chunks, m.chunks.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice, s));
        HIP_OK(hipStreamSynchronize(s));
        if (ovf) {
            fprintf(stderr, "argweaver_amd: a block needs %d slots per lane\n", ovf);
            return 1;
        }
        int RM = 8, RT = 4, sstride = 1;
        size_t shf = 0, sht = 0;
        for (int i = 0; i < n; i++) {
            RM = std::max(RM, m.h[i].RM);
            sstride = std::max(sstride, m.h[i].Smax);
            RT = std::max(RT, (m.h[i].Smax + 63) / 64);
            shf = std::max(shf, lds_fw32(m.h[i]));
            sht = std::max(sht, lds_tb(m.h[i]));
        }
        auto fw = [&](int mode, const Chunk *chp, int cnt, const float *snap, int *flags) -> int {
#define FW32(R)                                                                              \
    do {                                                                                     \
        auto kern = k_forward32<R>;                                                          \
        HIP_OK(hipFuncSetAttribute((const void *)kern,                                       \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shf));   \
        hipLaunchKernelGGL(kern, dim3(cnt), dim3(64), shf, s, m.dD, chp, mode, snap, sstride, flags, \
                           opt.tol_rel, opt.tol_abs);                                  \
    } while (0)
            if (RM <= 8) FW32(8);
            else if (RM <= 16) FW32(16);
            else if (RM <= 32) FW32(32);
            else FW32(64);
#undef FW32
            HIP_OK(hipGetLastError());
            return 0;
        };
        auto tb = [&](int mode, const Chunk *chp, int cnt, const int *snap, int *flags) -> int {
#define TB32(R)                                                                              \
    do {                                                                                     \
        auto kern = k_traceback32<R>;                                                        \
        HIP_OK(hipFuncSetAttribute((const void *)kern,                                       \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)sht));   \
        hipLaunchKernelGGL(kern, dim3(cnt), dim3(64), sht, s, m.dD, chp, mode, snap, flags); \
    } while (0)
            if (RT <= 4) TB32(4);
            else if (RT <= 8) TB32(8);
            else if (RT <= 16) TB32(16);
            else if (RT <= 32) TB32(32);
            else TB32(64);
#undef TB32
            HIP_OK(hipGetLastError());
            return 0;
        };
        // forward: speculative pass over every tile, then fix-up passes
        if (fw(0, m.d_chunks, nch, nullptr, nullptr)) return 1;
        std::vector<int> redo;
        for (int j = 0; j < nch; j++)
            if (!m.chunks[j].first) redo.push_back(j);
        std::vector<Chunk> sub;
        std::vector<int> flags;
        Chunk *d_sub = m.d_chunks;   // re-used region after the full list
        (void)d_sub;
        if (!opt.fixup) redo.clear();
        while (!redo.empty()) {
            rs.fw_passes++;
            sub.clear();
            for (int j : redo) sub.push_back(m.chunks[j]);
            const int ns = (int)sub.size();
            // sub-list after the full list inside the chunk buffer is not
            // available; reuse the front of the buffer and restore later
            HIP_OK(hipMemcpyAsync(m.d_chunks, sub.data(), sizeof(Chunk) * ns, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_fw_snapshot, dim3(ns), dim3(64), 0, s, m.dD, m.d_chunks, m.d_fsnap, sstride);
            if (fw(1, m.d_chunks, ns, m.d_fsnap, m.d_flags)) return 1;
            flags.resize(ns);
            HIP_OK(hipMemcpyAsync(flags.data(), m.d_flags, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            std::vector<int> next;
            for (int x = 0; x < ns; x++) {
                const int j = redo[x];
                const Chunk &k = m.chunks[j];
                rs.fw_rerun_cols += (flags[x] < 0 ? k.c1 : flags[x] + 1) - k.c0;
                if (flags[x] < 0 && !k.last) next.push_back(j + 1);
            }
            redo.swap(next);
        }
        HIP_OK(hipMemcpyAsync(m.d_chunks, m.chunks.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice, s));
        (void)hipEventRecord(c.ev[4], s);
        hipLaunchKernelGGL(k_lnl_b, dim3(n), dim3(256), 0, s, m.dD);
        if (opt.traceback) {
            if (tb(0, m.d_chunks, nch, nullptr, nullptr)) return 1;
            redo.clear();
            for (int j = 0; j < nch; j++)
                if (!m.chunks[j].last) redo.push_back(j);
            if (!opt.fixup) redo.clear();
            while (!redo.empty()) {
                rs.tb_passes++;
                sub.clear();
                for (int j : redo) sub.push_back(m.chunks[j]);
                const int ns = (int)sub.size();
                HIP_OK(hipMemcpyAsync(m.d_chunks, sub.data(), sizeof(Chunk) * ns, hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_tb_snapshot, dim3(ns), dim3(64), 0, s, m.dD, m.d_chunks, m.d_tsnap);
                if (tb(1, m.d_chunks, ns, m.d_tsnap, m.d_flags)) return 1;
                flags.resize(ns);
                HIP_OK(hipMemcpyAsync(flags.data(), m.d_flags, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
                HIP_OK(hipStreamSynchronize(s));
                std::vector<int> next;
                for (int x = 0; x < ns; x++) {
                    const int j = redo[x];
                    const Chunk &k = m.chunks[j];
                    rs.tb_rerun_cols += k.c1 - (flags[x] < 0 ? k.c0 : flags[x]);
                    if (flags[x] < 0 && !k.first) next.push_back(j - 1);
                }
                redo.swap(next);
            }
            HIP_OK(hipMemcpyAsync(m.d_chunks, m.chunks.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice, s));
        }
        (void)hipEventRecord(c.ev[5], s);
    }
    std::vector<double> lnl(n);
    for (int i = 0; i < n; i++)
        HIP_OK(hipMemcpyAsync(&m.lnl[i], m.h[i].lnl, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    m.prec_done = opt.prec;
    float ms;
    (void)hipEventElapsedTime(&ms, c.ev[1], c.ev[2]); rs.ms_prep = ms;
    (void)hipEventElapsedTime(&ms, c.ev[2], c.ev[3]); rs.ms_emit = ms;
    (void)hipEventElapsedTime(&ms, c.ev[3], c.ev[4]); rs.ms_forward = ms;
    (void)hipEventElapsedTime(&ms, c.ev[4], c.ev[5]); rs.ms_traceback = ms;
    (void)hipEventElapsedTime(&ms, c.ev[0], c.ev[5]); rs.ms_total = ms;
    for (int i = 0; i < n; i++) {
        rs.W += m.W[i];
        rs.nvar += m.nvar[i];
    }
    if (st) *st = rs;
    return 0;
}

int Batch::copy_path(int i, int *path) const
{
    const Impl &m = *impl_;
    HIP_OK(hipMemcpy(path, m.h[i].path, sizeof(int) * m.h[i].L, hipMemcpyDeviceToHost));
    return 0;
}

int Batch::copy_norms(int i, double *norms) const
{
    const Impl &m = *impl_;
    HIP_OK(hipMemcpy(norms, m.h[i].norms, sizeof(double) * m.h[i].L, hipMemcpyDeviceToHost));
    return 0;
}

int Batch::copy_alpha(int i, double *alpha) const
{
    const Impl &m = *impl_;
    const DevInst &d = m.h[i];
    if (m.prec_done == Precision::FP64) {
        HIP_OK(hipMemcpy(alpha, d.alpha, sizeof(double) * d.W, hipMemcpyDeviceToHost));
        return 0;
    }
    std::vector<float> tmp(d.W);
    HIP_OK(hipMemcpy(tmp.data(), d.alpha32, sizeof(float) * d.W, hipMemcpyDeviceToHost));
    for (long long x = 0; x < d.W; x++) alpha[x] = tmp[x];
    return 0;
}

long long instance_states(const FlatTrees &ft, int T, std::vector<int> *Sv)
{
    long long W = 0;
    if (Sv) Sv->resize(ft.nblocks);
    for (int b = 0; b < ft.nblocks; b++) {
        const int N = ft.nnodes;
        const int *par = &ft.parent[(size_t)b * N];
        const int *age = &ft.age[(size_t)b * N];
        int S = 0;
        for (int i = 0; i < N; i++) {
            int top = par[i] == -1 ? T - 2 : age[par[i]];
            S += top - age[i] + 1;
        }
        if (Sv) (*Sv)[b] = S;
        W += (long long)S * ft.len[b];
    }
    return W;
}

bool device_available()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return false;
    return n > 0;
}

const char *device_name()
{
    if (ctx_init()) return "unavailable";
    return ctx().prop.gcnArchName;
}

int device_info(char *buf, int n)
{
    if (ctx_init()) return 1;
    const hipDeviceProp_t &p = ctx().prop;
    snprintf(buf, n, "%s %s CUs=%d lds_per_block=%zu lds_per_cu=%zu clock_khz=%d mem_gb=%.1f",
             p.name, p.gcnArchName, p.multiProcessorCount, (size_t)p.sharedMemPerBlock,
             (size_t)p.maxSharedMemoryPerMultiProcessor, p.clockRate,
             p.totalGlobalMem / 1073741824.0);
    return 0;
}

// one instance = a batch of one; the batch object is kept between calls so
// device buffers are reused
int run_instance(const HostInstance &hi, const double *u, int *path, double *alpha, double *lnl,
                 RunStats *stats, Precision prec, const double *prior, int chunk_cols)
{
    Batch bt;
    bt.add(hi, u, prior);
    if (bt.upload()) return 1;
    BatchOptions opt;
    opt.prec = prec;
    opt.traceback = (u != nullptr);
    opt.chunk_cols = chunk_cols;
    if (bt.run(opt, stats)) return 1;
    if (path && u && bt.copy_path(0, path)) return 1;
    if (alpha && bt.copy_alpha(0, alpha)) return 1;
    if (lnl) *lnl = bt.lnl(0);
    return 0;
}

}  // namespace aw
