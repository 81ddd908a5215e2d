// engine.hip -- gfx950 kernels of the threading HMM and their host driver.
//
// Pipeline for one instance (one HIP stream, inputs resident in HBM):
//   k_classify    one lane per column: masked / invariant / variant
//   k_scan        single-workgroup exclusive scan of the variant flags
//   k_prep        one workgroup per local-tree block: state space, lineage
//                 counts, TransMatrix vectors, the (a,b) and (c,a,b) tables,
//                 the switch matrix from the previous block, emission terms
//   k_var_emit    one wave per variant column: Felsenstein partials and the
//                 emission row over the block's states
//   k_forward64   one workgroup per instance: the column recursion with the
//                 previous column staged in LDS, per-time sums by wave
//                 reductions, two barriers per column; writes the FP64 table
//   k_traceback64 one wave per instance: reverse sampling with one uniform
//                 per column (CDF by a wave scan)
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

//===========================================================================
// column classification

__global__ void k_classify(DevInst d)
{
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= d.L) return;
    char c0 = d.seq[col];
    bool inv = true;
    for (int r = 1; r < d.nrows; r++)
        if (d.seq[(size_t)r * d.L + col] != c0) { inv = false; break; }
    d.vflag[col] = inv ? 0 : 1;
    d.cls[col] = inv ? (c0 == 'N' ? -2 : -1) : 0;
}

// exclusive scan of vflag into cls (variant columns only) and vcol; one
// workgroup of 1024 threads, each owning a contiguous run of columns
__global__ __launch_bounds__(1024) void k_scan(DevInst d, int *nvar_out)
{
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const int per = (d.L + 1023) / 1024;
    const int s0 = tid * per, e0 = min(d.L, s0 + per);
    int loc = 0;
    for (int i = s0; i < e0; i++) loc += d.vflag[i];
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
        if (d.vflag[i]) {
            d.cls[i] = run;
            d.vcol[run] = i;
            run++;
        }
    }
    if (tid == 1023) *nvar_out = part[1023];
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
    // last node whose first state index is <= k (first[] is nondecreasing)
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ni.first[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_prep(DevInst d)
{
    __shared__ PrepShared sh;
    const int b = blockIdx.x;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int T = d.T, N = d.N, T1 = T - 1;

    if (tid == 0) sh.M = *d.model;
    for (int i = tid; i < N; i += NT) {
        size_t x = (size_t)b * N + i;
        sh.par[i] = d.parent[x];
        sh.ch0[i] = d.c0[x];
        sh.ch1[i] = d.c1[x];
        sh.ag[i] = d.age[x];
        if (b > 0) {
            size_t y = (size_t)(b - 1) * N + i;
            sh.ppar[i] = d.parent[y];
            sh.pch0[i] = d.c0[y];
            sh.pch1[i] = d.c1[y];
            sh.pag[i] = d.age[y];
            sh.map[i] = d.map[x];
        }
    }
    __syncthreads();
    TreeView t{N, -1, sh.par, sh.ch0, sh.ch1, sh.ag};
    TreeView pt{N, -1, sh.ppar, sh.pch0, sh.pch1, sh.pag};
    Spr spr{0, 0, 0, 0};
    if (b > 0) spr = Spr{d.spr[4 * b], d.spr[4 * b + 1], d.spr[4 * b + 2], d.spr[4 * b + 3]};
    if (tid == 0) {
        int r = 0;
        for (int i = 0; i < N; i++) if (sh.par[i] == -1) r = i;
        sh.root = r;
        t.root = r;
        count_lineages(t, T, sh.L);
        calc_trans(t, sh.M, sh.L, sh.tv);
        build_node_index(t, T, sh.ni);
        sh.mainlen = emit_mainlen(t, sh.M);
        d.roots[b] = r;
        int order[MAXN];
        postorder(t, order);
        for (int i = 0; i < N; i++) d.order[(size_t)b * N + i] = order[i];
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
    const int so = d.soff[b], S = d.soff[b + 1] - so;

    // transition vectors, kept for host-side consumers and the FP32 path
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
        d.tv[(size_t)b * 9 * T + x] = val;
    }

    // state list and node index
    for (int n = tid; n < N; n += NT) {
        d.nfirst[(size_t)b * N + n] = sh.ni.first[n];
        d.nhi[(size_t)b * N + n] = sh.ni.hi[n];
        for (int a = sh.ag[n]; a <= sh.ni.hi[n]; a++) {
            int k = sh.ni.first[n] + a - sh.ag[n];
            d.snode[so + k] = n;
            d.stime[so + k] = a;
        }
        // branch mutation probabilities for pruning (prob_tree_mutation)
        double mut = 0.0, nomut = 0.0;
        if (n != sh.root) {
            const double mintime = M.times[1] * .1;
            double tt = fmax(M.times[sh.ag[sh.par[n]]] - M.times[sh.ag[n]], mintime);
            mut = prob_branch(tt, M.mu, true);
            nomut = prob_branch(tt, M.mu, false);
        }
        d.nmut[2 * ((size_t)b * N + n)] = mut;
        d.nmut[2 * ((size_t)b * N + n) + 1] = nomut;
    }

    // per-state emission terms and first-column prior
    for (int k = tid; k < S; k += NT) {
        int n = node_of_state(sh.ni, N, k);
        int a = sh.ag[n] + k - sh.ni.first[n];
        EmitTerms e;
        emit_terms(t, M, sh.mainlen, n, a, e);
        d.einv[so + k] = e.inv;
        for (int x = 0; x < 3; x++) {
            d.eterm[6 * (size_t)(so + k) + x] = e.mut[x];
            d.eterm[6 * (size_t)(so + k) + 3 + x] = e.nomut[x];
        }
        if (b == 0) d.prior[k] = prior_time(a, sh.L, M);
    }

    // (a,b) table and the (c,a,b) same-branch tables
    const int TT = T1 * T1;
    for (int x = tid; x < TT; x += NT) {
        int a = x / T1, bb = x % T1;
        d.tm[(size_t)b * TT + x] = sh.tv.get_time(a, bb, 0, false);
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
        d.tm2s[(size_t)b * d.TC + x] = same;
        d.tm2[(size_t)b * d.TC + x] = same - sh.tv.get_time(a, bb, 0, false);
    }

    // states grouped by time (counting sort, stable in state order)
    if (tid == 0) {
        int cnt[MAXT + 1];
        for (int a = 0; a <= T; a++) cnt[a] = 0;
        for (int n = 0; n < N; n++)
            for (int a = sh.ag[n]; a <= sh.ni.hi[n]; a++) cnt[a + 1]++;
        for (int a = 0; a < T; a++) cnt[a + 1] += cnt[a];
        for (int a = 0; a <= T; a++) d.tstart[(size_t)b * (T + 1) + a] = cnt[a];
        for (int n = 0; n < N; n++)
            for (int a = sh.ag[n]; a <= sh.ni.hi[n]; a++)
                d.tperm[so + cnt[a]++] = sh.ni.first[n] + a - sh.ag[n];
    }

    if (b == 0) return;

    // switch matrix from block b-1 (calc_transition_probs_switch)
    const int so1 = d.soff[b - 1], S1 = so - so1;
    for (int j = tid; j < S1; j += NT) {
        int n1 = node_of_state(sh.pni, N, j);
        int t1 = sh.pag[n1] + j - sh.pni.first[n1];
        int tgt = sw_determ(pt, t, sh.ni, spr, sh.map, n1, t1);
        d.determ[so1 + j] = tgt;
        d.determp[so1 + j] = tgt >= 0 ? sw_determprob(pt, M, sh.PL, spr, sh.cx, n1, t1) : 0.0;
    }
    for (int k = tid; k < S; k += NT) {
        d.recombrow[so + k] = 0.0;
        d.recoalrow[so + k] = 0.0;
    }
    if (tid == 0) {
        int rs = lookup(pt, sh.pni, spr.rn, spr.rt);
        int cs = lookup(pt, sh.pni, spr.cn, spr.ct);
        sh.recombsrc = rs;
        sh.recoalsrc = cs;
        d.srcs[2 * b] = rs;
        d.srcs[2 * b + 1] = cs;
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
            d.recombrow[so + stay] = sw_recomb_recoal(pt, M, sh.PL, spr, spr.ct, spr.rn, spr.rt,
                                                      sh.cx.rpa, sh.cx.last_len);
        if (esc != -1)
            d.recombrow[so + esc] = sw_recomb_recoal(pt, M, sh.PL, spr, spr.ct, spr.rn, spr.rt,
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
            d.recoalrow[so + k] = sw_recomb_recoal(pt, M, sh.PL, spr, t2, spr.cn, spr.ct,
                                                   sh.cx.rpa, sh.cx.last_len);
        }
    }
    __syncthreads();
    // inverse of the deterministic map, sources ascending per target
    if (tid == 0) {
        int *ptr = d.invptr + so + b;
        for (int k = 0; k <= S; k++) ptr[k] = 0;
        for (int j = 0; j < S1; j++) {
            int k = d.determ[so1 + j];
            if (j != sh.recombsrc && j != sh.recoalsrc && k != -1) ptr[k + 1]++;
        }
        for (int k = 0; k < S; k++) ptr[k + 1] += ptr[k];
        for (int j = 0; j < S1; j++) {
            int k = d.determ[so1 + j];
            if (j != sh.recombsrc && j != sh.recoalsrc && k != -1) d.invidx[so1 + ptr[k]++] = j;
        }
        for (int k = S; k > 0; k--) ptr[k] = ptr[k - 1];
        ptr[0] = 0;
    }
}

//===========================================================================
// emissions of variant columns

__global__ __launch_bounds__(64) void k_var_emit(DevInst d)
{
    __shared__ int par[MAXN], ch0[MAXN], ch1[MAXN], ag[MAXN], order[MAXN];
    __shared__ char bases[MAXN];
    __shared__ double muts[MAXN], nomuts[MAXN];
    __shared__ double inner[4 * MAXN], outer[4 * MAXN];
    __shared__ double in2[4];
    const int v = blockIdx.x;
    const int lane = threadIdx.x;
    const int N = d.N;
    const int col = d.vcol[v];
    // block containing the column
    int lo = 0, hi = d.B - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (d.bstart[mid] <= col) lo = mid;
        else hi = mid - 1;
    }
    const int b = lo;
    const int nleaves = (N + 1) / 2;
    for (int i = lane; i < N; i += 64) {
        size_t x = (size_t)b * N + i;
        par[i] = d.parent[x];
        ch0[i] = d.c0[x];
        ch1[i] = d.c1[x];
        ag[i] = d.age[x];
        order[i] = d.order[x];
        muts[i] = d.nmut[2 * x];
        nomuts[i] = d.nmut[2 * x + 1];
        if (i < nleaves) bases[i] = d.seq[(size_t)i * d.L + col];
    }
    __syncthreads();
    const int root = d.roots[b];
    if (lane == 0) {
        TreeView t{N, root, par, ch0, ch1, ag};
        pruning(t, order, bases, muts, nomuts, inner, outer);
        leaf_partial(d.seq[(size_t)nleaves * d.L + col], in2);
    }
    __syncthreads();
    const int so = d.soff[b], S = d.soff[b + 1] - so;
    double *row = d.vemit + (size_t)v * d.Smax;
    for (int k = lane; k < S; k += 64) {
        int n = d.snode[so + k];
        const double *et = d.eterm + 6 * (size_t)(so + k);
        row[k] = emit_variant(in2, inner + 4 * n, outer + 4 * n, n == root, et, et + 3);
    }
}

//===========================================================================
// forward recursion, FP64, one workgroup per instance

__device__ inline double wave_sum(double x)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ inline double col_emit(const DevInst &d, int cls, double einv, int k)
{
    if (cls == -1) return einv;
    if (cls == -2) return 1.0;
    return d.vemit[(size_t)cls * d.Smax + k];
}

// per-time sums of an LDS column into G[0..T1): wave w handles bins w, w+NW, ...
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
__global__ __launch_bounds__(NT) void k_forward64(DevInst d)
{
    extern __shared__ double dyn[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int T1 = d.T - 1, TT = T1 * T1, N = d.N;
    double *bufA = dyn;
    double *bufB = bufA + d.Smax;
    double *tm = bufB + d.Smax;
    double *tm2 = tm + TT;
    double *G = tm2 + d.TC;                       // 32 doubles
    int *tperm = (int *)(G + 32);                 // Smax ints
    int *tstart = tperm + d.Smax;                 // T+1 ints
    int *tcoff = tstart + 33;                     // T1 ints

    for (int c = tid; c < T1; c += NT) tcoff[c] = tc_base(c, T1);

    double *cur = bufA, *nxt = bufB;
    double sc = 1.0;       // alpha of the column held in cur = cur * sc
    double lnl = 0.0;
    int prevS = 0;
    long long prev_row = 0;

    for (int b = 0; b < d.B; b++) {
        const int so = d.soff[b], S = d.soff[b + 1] - so;
        const int start = d.bstart[b], len = d.blen[b];
        const long long aoff = d.aoff[b];

        // tables of this block into LDS (all reads of the previous block's
        // tables finished before the last barrier of its final column)
        for (int x = tid; x < TT; x += NT) tm[x] = d.tm[(size_t)b * TT + x];
        for (int x = tid; x < d.TC; x += NT) tm2[x] = d.tm2[(size_t)b * d.TC + x];
        for (int x = tid; x < S; x += NT) tperm[x] = d.tperm[so + x];
        for (int x = tid; x <= d.T; x += NT) tstart[x] = d.tstart[(size_t)b * (d.T + 1) + x];

        // per-slot state data
        int sb[R], sj1[R], slen[R], sbase[R], sstride[R];
        double se[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            int k = tid + NT * r;
            sb[r] = 0; sj1[r] = 0; slen[r] = 0; sbase[r] = 0; sstride[r] = 1; se[r] = 0.0;
            if (k < S) {
                int n = d.snode[so + k];
                int bt = d.stime[so + k];
                int lo = d.age[(size_t)b * N + n];
                int hi = d.nhi[(size_t)b * N + n];
                sb[r] = bt;
                sj1[r] = d.nfirst[(size_t)b * N + n];
                slen[r] = hi - lo + 1;
                sstride[r] = T1 - lo;
                sbase[r] = tc_base(lo, T1) + (bt - lo);
                se[r] = d.einv[so + k];
            }
        }

        // column `start`: prior (block 0) or switch from the previous block
        {
            const int cls = d.cls[start];
            if (b == 0) {
                for (int k = tid; k < S; k += NT) nxt[k] = d.prior[k];
            } else {
                const int rs = d.srcs[2 * b], cs = d.srcs[2 * b + 1];
                const int so1 = d.soff[b - 1];
                const int *ptr = d.invptr + so + b;
                const double prs = (rs != -1) ? cur[rs] * sc : 0.0;
                const double pcs = (cs != -1) ? cur[cs] * sc : 0.0;
                for (int k = tid; k < S; k += NT) {
                    double acc = 0.0;
                    for (int x = ptr[k]; x < ptr[k + 1]; x++) {
                        int j = d.invidx[so1 + x];
                        acc += (cur[j] * sc) * d.determp[so1 + j];
                    }
                    double rr = d.recombrow[so + k], cr = d.recoalrow[so + k];
                    if (rs != -1 && rr > 0.0) acc += prs * rr;
                    if (cs != -1 && cr > 0.0) acc += pcs * cr;
                    nxt[k] = acc * col_emit(d, cls, d.einv[so + k], k);
                }
            }
            __syncthreads();
            // last row of the previous block, now that its LDS column is consumed
            if (b > 0)
                for (int k = tid; k < prevS; k += NT) d.alpha[prev_row + k] = cur[k] * sc;
            time_sums<NT>(nxt, tperm, tstart, T1, G);
            __syncthreads();
            double norm = 0.0;
            for (int a = 0; a < T1; a++) norm += G[a];
            if (b == 0) {
                sc = 1.0;   // the prior column is kept unnormalised
                if (d.L == 1) lnl = log(norm);
            } else {
                sc = 1.0 / norm;
                lnl += log(norm);
            }
            double *tmp = cur; cur = nxt; nxt = tmp;
        }

        for (int i = 1; i < len; i++) {
            const int col = start + i;
            const int cls = d.cls[col];
            // row i-1 of this block
            const long long row = aoff + (long long)(i - 1) * S;
            // T1 x T1 product of the per-time sums, one bin per lane
            double tfb = 0.0;
            if (lane < T1) {
                for (int a = 0; a < T1; a++) tfb += tm[a * T1 + lane] * G[a];
                tfb *= sc;
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int k = tid + NT * r;
                double tfv = __shfl(tfb, sb[r], 64);
                if (k < S) {
                    d.alpha[row + k] = cur[k] * sc;
                    double s = 0.0;
                    const double *coef = tm2 + sbase[r];
                    const double *src = cur + sj1[r];
                    const int st = sstride[r];
                    for (int x = 0; x < slen[r]; x++) s += coef[x * st] * src[x];
                    double e = (cls == -1) ? se[r] : col_emit(d, cls, 0.0, k);
                    nxt[k] = (tfv + sc * s) * e;
                }
            }
            __syncthreads();
            time_sums<NT>(nxt, tperm, tstart, T1, G);
            __syncthreads();
            double norm = 0.0;
            for (int a = 0; a < T1; a++) norm += G[a];
            sc = 1.0 / norm;
            lnl += log(norm);
            double *tmp = cur; cur = nxt; nxt = tmp;
        }
        prevS = S;
        prev_row = aoff + (long long)(len - 1) * S;
        __syncthreads();
    }
    for (int k = tid; k < prevS; k += NT) d.alpha[prev_row + k] = cur[k] * sc;
    if (tid == 0) *d.lnl = lnl;
}

//===========================================================================
// stochastic traceback, one wave per instance

// first index j of w[0..n) with cumulative sum >= u * total (sample(),
// common.h:272-290); each lane owns a contiguous chunk
__device__ inline int wave_sample(const double *w, int n, double u)
{
    const int lane = threadIdx.x & 63;
    const int chunk = (n + 63) / 64;
    const int s0 = min(n, lane * chunk), e0 = min(n, s0 + chunk);
    double loc = 0.0;
    for (int j = s0; j < e0; j++) loc += w[j];
    // exclusive scan of the lane sums
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

__global__ __launch_bounds__(64) void k_traceback64(DevInst d)
{
    extern __shared__ double dyn[];
    const int lane = threadIdx.x;
    const int T1 = d.T - 1, TT = T1 * T1, N = d.N;
    double *w = dyn;                 // Smax
    double *tm = w + d.Smax;         // TT
    double *tm2s = tm + TT;          // TC
    int *snode = (int *)(tm2s + d.TC);
    int *stime = snode + d.Smax;
    int *tcoff = stime + d.Smax;     // T1

    for (int c = lane; c < T1; c += 64) tcoff[c] = tc_base(c, T1);
    const int L = d.L;
    int knext;
    {
        const int b = d.B - 1;
        const int S = d.soff[b + 1] - d.soff[b];
        const long long row = d.aoff[b] + (long long)(d.blen[b] - 1) * S;
        for (int j = lane; j < S; j += 64) w[j] = d.alpha[row + j];
        __syncthreads();
        knext = wave_sample(w, S, d.u[0]);
        if (lane == 0) d.path[L - 1] = knext;
        __syncthreads();
    }
    for (int b = d.B - 1; b >= 0; b--) {
        const int so = d.soff[b], S = d.soff[b + 1] - so;
        const int start = d.bstart[b], len = d.blen[b];
        for (int x = lane; x < TT; x += 64) tm[x] = d.tm[(size_t)b * TT + x];
        for (int x = lane; x < d.TC; x += 64) tm2s[x] = d.tm2s[(size_t)b * d.TC + x];
        for (int x = lane; x < S; x += 64) {
            snode[x] = d.snode[so + x];
            stime[x] = d.stime[so + x];
        }
        __syncthreads();
        for (int i = len - 2; i >= 0; i--) {
            const int col = start + i;
            const int nk = snode[knext], bk = stime[knext];
            const int ck = d.age[(size_t)b * N + nk];
            const int cbase = tcoff[ck] + (bk - ck), cstride = T1 - ck;
            const long long row = d.aoff[b] + (long long)i * S;
            for (int j = lane; j < S; j += 64) {
                int a = stime[j];
                double tr = (snode[j] == nk) ? tm2s[cbase + (a - ck) * cstride] : tm[a * T1 + bk];
                w[j] = d.alpha[row + j] * tr;
            }
            __syncthreads();
            knext = wave_sample(w, S, d.u[L - 1 - col]);
            if (lane == 0) d.path[col] = knext;
            __syncthreads();
        }
        if (b > 0) {
            // switch column: last column of block b-1
            const int so1 = d.soff[b - 1], S1 = so - so1;
            const int rs = d.srcs[2 * b], cs = d.srcs[2 * b + 1];
            const long long row = d.aoff[b - 1] + (long long)(d.blen[b - 1] - 1) * S1;
            const double rr = d.recombrow[so + knext], cr = d.recoalrow[so + knext];
            for (int j = lane; j < S1; j += 64) {
                double g;
                if (j == cs) g = cr;
                else if (j == rs) g = rr;
                else g = (d.determ[so1 + j] == knext) ? d.determp[so1 + j] : 0.0;
                w[j] = d.alpha[row + j] * g;
            }
            __syncthreads();
            const int col = start - 1;
            knext = wave_sample(w, S1, d.u[L - 1 - col]);
            if (lane == 0) d.path[col] = knext;
            __syncthreads();
        }
    }
}

//===========================================================================
// host driver

namespace {

struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    int reserve(size_t bytes)
    {
        if (bytes <= cap) return 0;
        if (base) HIP_OK(hipFree(base));
        base = nullptr;
        size_t c = std::max(bytes, cap * 3 / 2);
        HIP_OK(hipMalloc(&base, c));
        cap = c;
        return 0;
    }
    template <class X>
    X *take(size_t n)
    {
        size_t off = (used + 255) & ~size_t(255);
        used = off + n * sizeof(X);
        return reinterpret_cast<X *>(base + off);
    }
};

struct Ctx {
    bool init = false;
    int dev = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    Arena in, work;
    int *nvar_dev = nullptr;
    int *nvar_host = nullptr;
    hipEvent_t ev[6];
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
    HIP_OK(hipMalloc(&c.nvar_dev, sizeof(int)));
    HIP_OK(hipHostMalloc(&c.nvar_host, sizeof(int), hipHostMallocDefault));
    for (auto &e : c.ev) HIP_OK(hipEventCreate(&e));
    c.init = true;
    return 0;
}

template <class X>
size_t bytes_of(size_t n)
{
    return ((n * sizeof(X)) + 255) & ~size_t(255);
}

}  // namespace

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

template <int NT, int R>
static int launch_forward(const DevInst &d, size_t shmem, hipStream_t s)
{
    auto kern = k_forward64<NT, R>;
    HIP_OK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shmem));
    hipLaunchKernelGGL(kern, dim3(1), dim3(NT), shmem, s, d);
    HIP_OK(hipGetLastError());
    return 0;
}

int run_instance(const HostInstance &hi, const double *u, int *path, double *alpha, double *lnl,
                 RunStats *stats, Precision prec, const double *prior)
{
    (void)prec;
    if (ctx_init()) return 1;
    Ctx &c = ctx();
    const FlatTrees &ft = hi.ft;
    const int T = hi.model.T, T1 = T - 1;
    const int B = ft.nblocks, N = ft.nnodes, L = ft.seqlen;
    const int nrows = (int)hi.rows.size();
    if (N > MAXN || T > MAXT) {
        fprintf(stderr, "argweaver_amd: instance exceeds nnodes<=%d / ntimes<=%d\n", MAXN, MAXT);
        return 1;
    }
    std::vector<int> Sv;
    long long W = instance_states(ft, T, &Sv);
    std::vector<int> soff(B + 1, 0);
    std::vector<long long> aoff(B + 1, 0);
    int Smax = 1;
    for (int b = 0; b < B; b++) {
        soff[b + 1] = soff[b] + Sv[b];
        aoff[b + 1] = aoff[b] + (long long)Sv[b] * ft.len[b];
        Smax = std::max(Smax, Sv[b]);
    }
    const int Stot = soff[B];
    const int TC = tc_base(T1, T1);

    // inputs
    size_t in_bytes = bytes_of<Model>(1) + 8 * bytes_of<int>((size_t)B * N) +
                      bytes_of<int>(4 * B) + 2 * bytes_of<int>(B) + bytes_of<int>(B + 1) +
                      bytes_of<long long>(B + 1) + bytes_of<char>((size_t)nrows * L) +
                      bytes_of<double>(L) + 4096;
    if (c.in.reserve(in_bytes)) return 1;
    c.in.used = 0;
    DevInst d{};
    d.T = T; d.N = N; d.B = B; d.L = L; d.nrows = nrows; d.Smax = Smax; d.TC = TC;
    Model *dm = c.in.take<Model>(1);
    int *dpar = c.in.take<int>((size_t)B * N), *dc0 = c.in.take<int>((size_t)B * N);
    int *dc1 = c.in.take<int>((size_t)B * N), *dage = c.in.take<int>((size_t)B * N);
    int *dmap = c.in.take<int>((size_t)B * N), *dspr = c.in.take<int>(4 * B);
    int *dbs = c.in.take<int>(B), *dbl = c.in.take<int>(B), *dsoff = c.in.take<int>(B + 1);
    long long *daoff = c.in.take<long long>(B + 1);
    char *dseq = c.in.take<char>((size_t)nrows * L);
    double *du = c.in.take<double>(L);
    HIP_OK(hipMemcpyAsync(dm, &hi.model, sizeof(Model), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dpar, ft.parent.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dc0, ft.child0.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dc1, ft.child1.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dage, ft.age.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dmap, ft.mapping.data(), sizeof(int) * B * N, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dspr, ft.spr.data(), sizeof(int) * 4 * B, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dbs, ft.start.data(), sizeof(int) * B, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dbl, ft.len.data(), sizeof(int) * B, hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(dsoff, soff.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(daoff, aoff.data(), sizeof(long long) * (B + 1), hipMemcpyHostToDevice, c.stream));
    for (int r = 0; r < nrows; r++)
        HIP_OK(hipMemcpyAsync(dseq + (size_t)r * L, hi.rows[r], L, hipMemcpyHostToDevice, c.stream));
    if (u) HIP_OK(hipMemcpyAsync(du, u, sizeof(double) * L, hipMemcpyHostToDevice, c.stream));
    d.model = dm; d.parent = dpar; d.c0 = dc0; d.c1 = dc1; d.age = dage; d.map = dmap;
    d.spr = dspr; d.bstart = dbs; d.blen = dbl; d.soff = dsoff; d.aoff = daoff; d.seq = dseq;
    d.u = du;

    // work buffers (variant rows sized after the count is known)
    size_t wk = 3 * bytes_of<int>(L) + 2 * bytes_of<int>(Stot) + 2 * bytes_of<int>((size_t)B * N) +
                bytes_of<int>(Stot) + bytes_of<int>((size_t)B * (T + 1)) +
                bytes_of<double>((size_t)B * T1 * T1) + 2 * bytes_of<double>((size_t)B * TC) +
                bytes_of<double>((size_t)B * 9 * T) + bytes_of<double>(Stot) +
                bytes_of<double>((size_t)Stot * 6) + bytes_of<int>(Stot) + 3 * bytes_of<double>(Stot) +
                bytes_of<int>(2 * B) + bytes_of<int>(Stot + B) + bytes_of<int>(Stot) +
                bytes_of<double>(Smax) + 2 * bytes_of<int>((size_t)B * N) + bytes_of<int>(B) +
                bytes_of<double>((size_t)B * N * 2) + bytes_of<double>(W) + bytes_of<int>(L) +
                bytes_of<double>(1) + 8192;

    (void)hipEventRecord(c.ev[0], c.stream);
    // classify + count variant columns
    if (c.work.reserve(wk)) return 1;
    c.work.used = 0;
    d.vflag = c.work.take<int>(L);
    d.cls = c.work.take<int>(L);
    d.vcol = c.work.take<int>(L);
    hipLaunchKernelGGL(k_classify, dim3((L + 255) / 256), dim3(256), 0, c.stream, d);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c.stream, d, c.nvar_dev);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(c.nvar_host, c.nvar_dev, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    const int nvar = *c.nvar_host;
    d.nvar = nvar;

    // make room for the variant rows, keeping the classification results
    size_t need = wk + bytes_of<double>((size_t)std::max(nvar, 1) * Smax);
    if (need > c.work.cap) {
        std::vector<int> keep(3 * (size_t)L);
        HIP_OK(hipMemcpy(keep.data(), d.vflag, sizeof(int) * 3 * L, hipMemcpyDeviceToHost));
        // vflag, cls and vcol are laid out contiguously at the arena start
        if (c.work.reserve(need)) return 1;
        c.work.used = 0;
        d.vflag = c.work.take<int>(L);
        d.cls = c.work.take<int>(L);
        d.vcol = c.work.take<int>(L);
        HIP_OK(hipMemcpy(d.vflag, keep.data(), sizeof(int) * L, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d.cls, keep.data() + bytes_of<int>(L) / sizeof(int), sizeof(int) * L, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d.vcol, keep.data() + 2 * (bytes_of<int>(L) / sizeof(int)), sizeof(int) * L, hipMemcpyHostToDevice));
    }
    d.snode = c.work.take<int>(Stot);
    d.stime = c.work.take<int>(Stot);
    d.nfirst = c.work.take<int>((size_t)B * N);
    d.nhi = c.work.take<int>((size_t)B * N);
    d.tperm = c.work.take<int>(Stot);
    d.tstart = c.work.take<int>((size_t)B * (T + 1));
    d.tm = c.work.take<double>((size_t)B * T1 * T1);
    d.tm2 = c.work.take<double>((size_t)B * TC);
    d.tm2s = c.work.take<double>((size_t)B * TC);
    d.tv = c.work.take<double>((size_t)B * 9 * T);
    d.einv = c.work.take<double>(Stot);
    d.eterm = c.work.take<double>((size_t)Stot * 6);
    d.determ = c.work.take<int>(Stot);
    d.determp = c.work.take<double>(Stot);
    d.recombrow = c.work.take<double>(Stot);
    d.recoalrow = c.work.take<double>(Stot);
    d.srcs = c.work.take<int>(2 * B);
    d.invptr = c.work.take<int>(Stot + B);
    d.invidx = c.work.take<int>(Stot);
    d.prior = c.work.take<double>(Smax);
    d.order = c.work.take<int>((size_t)B * N);
    d.roots = c.work.take<int>(B);
    d.nmut = c.work.take<double>((size_t)B * N * 2);
    d.alpha = c.work.take<double>(W);
    d.path = c.work.take<int>(L);
    d.lnl = c.work.take<double>(1);
    d.vemit = c.work.take<double>((size_t)std::max(nvar, 1) * Smax);

    (void)hipEventRecord(c.ev[1], c.stream);
    hipLaunchKernelGGL(k_prep, dim3(B), dim3(256), 0, c.stream, d);
    HIP_OK(hipGetLastError());
    if (prior)
        HIP_OK(hipMemcpyAsync(d.prior, prior, sizeof(double) * Sv[0], hipMemcpyHostToDevice, c.stream));
    (void)hipEventRecord(c.ev[2], c.stream);
    if (nvar > 0) hipLaunchKernelGGL(k_var_emit, dim3(nvar), dim3(64), 0, c.stream, d);
    HIP_OK(hipGetLastError());
    (void)hipEventRecord(c.ev[3], c.stream);

    // forward
    const size_t shf = sizeof(double) * (2 * (size_t)Smax + T1 * T1 + TC + 32) +
                       sizeof(int) * ((size_t)Smax + 33 + 32);
    if (shf > (size_t)c.prop.sharedMemPerBlock) {
        fprintf(stderr, "argweaver_amd: %d states need %zu B of LDS (device has %zu)\n", Smax,
                shf, (size_t)c.prop.sharedMemPerBlock);
        return 1;
    }
    int rc;
    if (Smax <= 256) rc = launch_forward<256, 1>(d, shf, c.stream);
    else if (Smax <= 512) rc = launch_forward<256, 2>(d, shf, c.stream);
    else if (Smax <= 1024) rc = launch_forward<256, 4>(d, shf, c.stream);
    else if (Smax <= 2048) rc = launch_forward<256, 8>(d, shf, c.stream);
    else if (Smax <= 4096) rc = launch_forward<256, 16>(d, shf, c.stream);
    else if (Smax <= 8192) rc = launch_forward<512, 16>(d, shf, c.stream);
    else {
        fprintf(stderr, "argweaver_amd: %d states per block exceeds 8192\n", Smax);
        return 1;
    }
    if (rc) return rc;
    (void)hipEventRecord(c.ev[4], c.stream);

    if (u) {
        const size_t sht = sizeof(double) * ((size_t)Smax + T1 * T1 + TC) +
                           sizeof(int) * (2 * (size_t)Smax + 32);
        HIP_OK(hipFuncSetAttribute((const void *)k_traceback64,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)sht));
        hipLaunchKernelGGL(k_traceback64, dim3(1), dim3(64), sht, c.stream, d);
        HIP_OK(hipGetLastError());
    }
    (void)hipEventRecord(c.ev[5], c.stream);
    if (path && u) HIP_OK(hipMemcpyAsync(path, d.path, sizeof(int) * L, hipMemcpyDeviceToHost, c.stream));
    if (alpha) HIP_OK(hipMemcpyAsync(alpha, d.alpha, sizeof(double) * W, hipMemcpyDeviceToHost, c.stream));
    double lnl_h = 0.0;
    HIP_OK(hipMemcpyAsync(&lnl_h, d.lnl, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    if (lnl) *lnl = lnl_h;
    if (stats) {
        float ms;
        (void)hipEventElapsedTime(&ms, c.ev[1], c.ev[2]); stats->ms_prep = ms;
        (void)hipEventElapsedTime(&ms, c.ev[2], c.ev[3]); stats->ms_emit = ms;
        (void)hipEventElapsedTime(&ms, c.ev[3], c.ev[4]); stats->ms_forward = ms;
        (void)hipEventElapsedTime(&ms, c.ev[4], c.ev[5]); stats->ms_traceback = ms;
        stats->W = W;
        stats->nvar = nvar;
    }
    return 0;
}

}  // namespace aw
