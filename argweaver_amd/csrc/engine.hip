// engine.hip -- gfx950 kernels of the threading HMM and their host driver.
//
// Everything runs on one HIP stream over a Batch of independent instances
// whose inputs (local trees, sequences, uniforms) are resident in HBM.
//
//   k_classify_b   one lane per column: masked / invariant / variant
//   k_scan_b       one workgroup per instance: variant-column numbering
//   k_prep_b       one workgroup per local-tree block: state space, lineage
//                  counts, TransMatrix vectors and tables, the switch matrix
//                  from the previous block, emission terms, first-column
//                  prior, and the per-state coefficient planes of the
//                  lane-run layout (engine.h)
//   k_var_emit_b   one wave per variant column: Felsenstein partials and the
//                  emission row over the block's states
//   k_fw_all       one wave per instance over all its columns, every
//                  instance of a batch in one launch (a switch on R picks
//                  the instantiation).  Lane l owns states l*R .. l*R+R-1
//                  (node-major).  The same-branch part of the transition is
//                  a per-branch prefix and suffix sum (in-lane scan plus a
//                  DPP wave-shift carry chain across the few lanes a branch
//                  spans); the other-branch part goes through the per-time
//                  sums (FP64 LDS atomics into 16 lane-group copies, a
//                  lane-pair reduction, readlane broadcast) and the (a,b)
//                  table, the result reaching each state by ds_bpermute; a
//                  column costs O(S + T^2) and rows are written coalesced.
//   k_lnl_b        per-instance sum of log normalisers (FP64)
//   k_tb_spec      one wave per option of a column tile, every tile in one
//                  launch: FP64 weights and CDF (DPP wave scan), rows
//                  staged through LDS in double-buffered batches.  A tile
//                  before the last runs a speculative chain (a state
//                  sampled from a column past its end and a warm-up) and
//                  alternative chains from the neighbouring states of that
//                  chain's branch; once the true state at the tile's end is
//                  known the option that started from it is exact
//   k_tb_rerun     a tile no option fits, re-sampled from its true state
//                  until it coalesces with an option's path (identical from
//                  there on)
//
// The forward kernel is a template over the table type: double is the
// drop-in arithmetic (the reference's forward table to ~1e-15), float keeps
// the table and the recursion in FP32 with per-column rescaling while the
// normalisers, the log-likelihood and the traceback CDF stay FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
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

// Read-only data indexed by wave-uniform values (block bounds, uniforms)
// goes through the constant address space so it is fetched with scalar loads.
#define AW_CONST __attribute__((address_space(4)))
template <class X>
__device__ inline AW_CONST const X *cp(const X *p)
{
    return (AW_CONST const X *)p;
}

// LDS operations of one wave are performed in order, so inside a
// single-wave workgroup only the compiler's order matters; every LDS access
// of a kernel goes through one dynamic shared array, so accesses that may
// overlap are never reordered.  This marks the places that rely on it.
__device__ inline void lds_order() {}

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
// cross-lane moves (DPP).  A lane whose source is outside the wave (wave
// shifts) or outside its 16-lane row (row shifts) receives 0.

constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114,
              DPP_ROW_SHR8 = 0x118, DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138,
              DPP_QUAD_SWAP1 = 0xB1;   // quad_perm [1,0,3,2]: lanes 2k and 2k+1 swap

template <int CTRL>
__device__ __forceinline__ int dpp_i(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_r(float x)
{
    return __int_as_float(dpp_i<CTRL>(__float_as_int(x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_r(double x)
{
    const long long v = __double_as_longlong(x);
    const unsigned lo = (unsigned)dpp_i<CTRL>((int)v);
    const unsigned hi = (unsigned)dpp_i<CTRL>((int)(v >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double readlane_d(double x, int l)
{
    const long long v = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)v, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// inclusive prefix sum over the 64 lanes: Hillis-Steele inside each 16-lane
// row, then the totals of the rows before
__device__ __forceinline__ double wave_incl_scan(double v)
{
    v += dpp_r<DPP_ROW_SHR1>(v);
    v += dpp_r<DPP_ROW_SHR2>(v);
    v += dpp_r<DPP_ROW_SHR4>(v);
    v += dpp_r<DPP_ROW_SHR8>(v);
    const double t0 = readlane_d(v, 15), t1 = readlane_d(v, 31), t2 = readlane_d(v, 47);
    const int row = threadIdx.x >> 4;
    double off = 0.0;
    if (row >= 1) off = t0;
    if (row >= 2) off += t1;
    if (row >= 3) off += t2;
    return v + off;
}

//===========================================================================
// column classification

__global__ void k_classify_b(const DevInst *__restrict__ D, const long long *coloff, int ninst)
{
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= coloff[ninst]) return;
    int i = find_owner(coloff, ninst, g);
    const DevInst &d = D[i];
    if (d.hostprep) return;
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
    if (d.hostprep) {   // the whole workgroup leaves: the host knows the count
        if (tid == 0) nvar_out[blockIdx.x] = 0;
        return;
    }
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
    int wsum[8];
    int hgt[MAXN], dep[MAXN];
    double rates[2 * MAXT + 1], Cs[2 * MAXT + 2], lnBt[MAXT], eexp[MAXT];
    double red[3 * 4];
    // per-time factors of the state planes
    double eB[MAXT], eN[MAXT], eE2[MAXT], low[MAXT];
    EmitTime et[MAXT];
};

// Lineage counts (count_lineages) with one thread per node and integer LDS
// atomics: the single-thread loop is a chain of dependent LDS updates.
// The counts must be zero on entry; the caller synchronises after.
__device__ inline void lineages_par(const TreeView &t, int T, Lineages &L, int tid, int NT)
{
    for (int i = tid; i < t.n; i += NT) {
        const int p = t.parent[i];
        const int page = (p == -1) ? T - 2 : t.age[p];
        for (int j = t.age[i]; j < page; j++) {
            atomicAdd(&L.nb[j], 1);
            atomicAdd(&L.nr[j], 1);
            atomicAdd(&L.nc[j], 1);
        }
        atomicAdd(&L.nr[page], 1);
        atomicAdd(&L.nc[page], 1);
        if (p == -1) atomicAdd(&L.nb[page], 1);
    }
}

// the node index (first state and top time of every node's branch,
// states.cpp:53-75) over the workgroup, one node per thread (blockDim.x ==
// MAXN == 256): a shuffle scan of the branch state counts in each wave,
// then the totals of the waves before (wsum[4], one per call); the caller
// synchronises before ni is read
__device__ inline void node_index_par(const TreeView &t, int T, NodeIndex &ni, int *wsum, int tid)
{
    int len = 0;
    if (tid < t.n) {
        ni.hi[tid] = t.top(tid, T);
        len = ni.hi[tid] - t.age[tid] + 1;
    }
    const int lane = tid & 63;
    int v = len;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == 63) wsum[tid >> 6] = v;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < (tid >> 6); w++) base += wsum[w];
    if (tid < t.n) ni.first[tid] = base + v - len;
}

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

// The coefficient planes of one block, written in the run's precision:
// FP32 runs keep only the FP32 planes (the same values as an FP64 plane
// rounded once)
struct PlaneOut {
    AW_GLOBAL double *p64;
    AW_GLOBAL float *p32;
    int RW;
    __device__ void put(int plane, int p, double v) const
    {
        if (p32) p32[(size_t)plane * RW + p] = (float)v;
        else p64[(size_t)plane * RW + p] = v;
    }
};

// Three sums over a 256-thread workgroup in a fixed order (a butterfly in
// each wave, then the four wave sums in wave order), so that every run adds
// the same terms in the same order; every thread gets the totals.
__device__ inline void block_sums3(double (&v)[3], double *red)
{
#pragma unroll
    for (int x = 0; x < 3; x++)
        for (int off = 32; off > 0; off >>= 1) v[x] += __shfl_xor(v[x], off);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int x = 0; x < 3; x++) red[4 * x + w] = v[x];
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 3; x++) v[x] = ((red[4 * x] + red[4 * x + 1]) + red[4 * x + 2]) + red[4 * x + 3];
}

// At most 95 VGPRs (five waves per SIMD, no spill): 16.5 -> 14.7 ms on the
// 256-instance batch against the 119 the compiler picks unbounded (four).
__global__ __launch_bounds__(256, 5) void k_prep_b(const DevInst *__restrict__ D,
                                                const int *__restrict__ blkoff, int ninst)
{
    __shared__ PrepShared sh;
    const int inst = find_owner(blkoff, ninst, (int)blockIdx.x);
    const DevInst &d = D[inst];
    if (d.hostprep) return;   // uniform over the workgroup
    const int b = blockIdx.x - blkoff[inst];
    const int tid = threadIdx.x, NT = blockDim.x;
    const int T = d.T, N = d.N, T1 = T - 1;
    const int R = d.R, RW = 64 * R;

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
    // roots, zeroed lineage counts
    for (int i = tid; i < N; i += NT) {
        if (sh.par[i] == -1) sh.root = i;
        if (b > 0 && sh.ppar[i] == -1) sh.proot = i;
    }
    for (int i = tid; i < MAXT; i += NT) {
        sh.L.nb[i] = sh.L.nr[i] = sh.L.nc[i] = 0;
        sh.PL.nb[i] = sh.PL.nr[i] = sh.PL.nc[i] = 0;
    }
    __syncthreads();
    t.root = sh.root;
    pt.root = (b > 0) ? sh.proot : -1;
    lineages_par(t, T, sh.L, tid, NT);
    if (b > 0) lineages_par(pt, T, sh.PL, tid, NT);
    node_index_par(t, T, sh.ni, sh.wsum, tid);
    if (b > 0) node_index_par(pt, T, sh.pni, sh.wsum + 4, tid);
    if (tid == 0) {
        sh.L.nb[T - 1] = 1;
        sh.PL.nb[T - 1] = 1;
    }
    // node levels for the level-synchronous pruning of the emissions: depth
    // (root 0) by walking up to the root, height (leaves 0) by walking up
    // from every leaf and raising each ancestor's height to its distance
    // (integer LDS atomics); every walk is bounded by N
    for (int i = tid; i < N; i += NT) sh.hgt[i] = 0;
    __syncthreads();
    for (int j = tid; j < N; j += NT) {
        int dd = 0;
        for (int p = sh.par[j]; p != -1 && dd < N; p = sh.par[p]) dd++;
        sh.dep[j] = dd;
        if (sh.ch0[j] == -1) {
            int k = 1;
            for (int p = sh.par[j]; p != -1 && k <= N; p = sh.par[p], k++) atomicMax(&sh.hgt[p], k);
        }
    }
    // the block's tree lengths, fixed-order sums over the workgroup (one
    // node per thread): get_treelen of this tree (calc_trans) and of the
    // tree before (the switch), and the emissions' branch-length sum with
    // its minimum branch length (emit.cpp:778-819)
    double lens[3] = {0.0, 0.0, 0.0};
    {
        const double *tm_ = sh.M.times;
        const double mintime = tm_[1] * .1;
        if (tid < N) {
            const int p = sh.par[tid];
            if (p != -1) {
                lens[0] = tm_[sh.ag[p]] - tm_[sh.ag[tid]];
                lens[1] = fmax(tm_[sh.ag[p]] - tm_[sh.ag[tid]], mintime);
            }
            const int pp = (b > 0) ? sh.ppar[tid] : -1;
            if (pp != -1) lens[2] = tm_[sh.pag[pp]] - tm_[sh.pag[tid]];
        }
        block_sums3(lens, sh.red);   // also completes the levels
    }
    for (int i = tid; i < N; i += NT) gp(d.order)[(size_t)b * N + i] = sh.hgt[i] | (sh.dep[i] << 16);
    __syncthreads();
    // TransMatrix vectors (calc_trans) split over the workgroup with the
    // sequential version's expressions: the rate prefix in time order on
    // one thread, the per-time terms one thread each, the running logadd of
    // lnB as a prefix sum of exponentials.  Meanwhile one thread computes
    // the switch context's sums (calc_recoal_sums and friends).
    if (tid == 0) sh.mainlen = lens[1];
    if (tid == 128) gp(d.roots)[b] = sh.root;
    if (tid == 192 && b > 0) switch_ctx_sums(pt, sh.M, sh.PL, spr, lens[2], sh.cx);
    {
        const Model &M = sh.M;
        double *rates = sh.rates + 1, *C = sh.Cs + 2;
        if (tid == 64) rates[-1] = 0.0;
        for (int i = tid; i < 2 * T - 1; i += NT)
            rates[i] = M.csteps[i] * sh.L.nb[i / 2] / (2.0 * M.pop[i / 2]);
        __syncthreads();
        if (tid == 64) {
            C[-2] = 0.0;
            C[-1] = 0.0;
            double c = 0.0;
            for (int b2 = 0; b2 < 2 * T - 1; b2++) {
                c = c + rates[b2];
                C[b2] = c;
            }
        }
        __syncthreads();
        TransVec &tv = sh.tv;
        const int ri = t.age[t.root];
        const double root_age = M.times[ri];
        const double len = lens[0];
        for (int bb = tid; bb < T; bb += NT) {
            tv.D[bb] = tv.E[bb] = tv.lnB[bb] = tv.lnE2[bb] = tv.lnNegG1[bb] = 0.0;
            tv.G2[bb] = tv.G3[bb] = tv.lnG4[bb] = tv.norecombs[bb] = 0.0;
            if (bb >= T - 1) continue;
            const int nb = sh.L.nb[bb], nr = sh.L.nr[bb], nc = sh.L.nc[bb];
            double len2 = len + M.times[bb];
            double len2_b;
            if (bb > ri) {
                len2 += M.times[bb] - root_age;
                len2_b = len2 + M.steps[bb];
            } else {
                len2_b = len2 + M.steps[ri];
            }
            sh.lnBt[bb] = C[2 * bb - 1] + log(M.steps[bb] * (nb + 1.0) / (nr + 1.0));
            double e2 = -C[2 * bb - 2] + (bb < T - 2 ? log(1 - exp(-rates[2 * bb] - rates[2 * bb - 1])) : 0.0);
            tv.lnE2[bb] = e2;
            tv.lnG4[bb] = e2;
            tv.lnNegG1[bb] = C[2 * bb - 1] +
                log(-M.steps[bb] * ((nb / (nr + 1.0 + int(bb < ri))) - (nb + 1.0) / (nr + 1.0)));
            double q = (bb < T - 2 ? 1.0 - exp(-rates[2 * bb]) : 1.0);
            tv.G2[bb] = q * M.steps[bb] * (nb + 1.0) / (nr + 1.0);
            tv.G3[bb] = q * M.steps[bb] * (nb / (nr + 1.0 + int(bb < ri)));
            tv.D[bb] = (1.0 - exp(-M.rho * len2)) / len2_b;
            tv.E[bb] = 1.0 / nc;
            tv.norecombs[bb] = exp(-fmax(M.rho * len2, M.rho));
        }
        // the switch context's per-time recoalescence terms
        if (b > 0)
            for (int a = tid - 128; a >= 0 && a < T; a += NT)
                sh.cx.recoals[a] = sw_recoal(M, sh.PL, spr.rt, spr.ct, a, sh.cx.rpa);
        __syncthreads();
        // lnB[b] = logadd(lnB[b-1], term_b) (trans.cpp:26-115) as
        // tmax + log(sum of exp(term_k - tmax), k <= b): the same value up to
        // rounding; the running logadd on one thread when the terms span
        // more than exp's range
        double tmax = -INFINITY, tmin = INFINITY;
        for (int k = 0; k < T - 1; k++) {
            tmax = fmax(tmax, sh.lnBt[k]);
            tmin = fmin(tmin, sh.lnBt[k]);
        }
        if (tmax - tmin < 600.0) {
            for (int k = tid; k < T - 1; k += NT) sh.eexp[k] = exp(sh.lnBt[k] - tmax);
            __syncthreads();
            for (int bb = tid; bb < T - 1; bb += NT) {
                double sum = 0.0;
                for (int k = 0; k <= bb; k++) sum += sh.eexp[k];
                tv.lnB[bb] = tmax + log(sum);
            }
        } else if (tid == 0) {
            double lb = sh.lnBt[0];
            tv.lnB[0] = lb;
            for (int bb = 1; bb < T - 1; bb++) {
                lb = logadd(lb, sh.lnBt[bb]);
                tv.lnB[bb] = lb;
            }
        }
    }
    __syncthreads();
    t.root = sh.root;
    pt.root = (b > 0) ? sh.proot : -1;
    const Model &M = sh.M;
    const TransVec &tv = sh.tv;
    const int so = cp(d.soff)[b], S = cp(d.soff)[b + 1] - so;
    const PlaneOut coef{d.fp32 ? nullptr : gp(d.lcoef) + (size_t)b * NCOEF * RW,
                        d.fp32 ? gp(d.lcoef32) + (size_t)b * NCOEF * RW : nullptr, RW};
    auto info = gp(d.linfo) + (size_t)b * RW;
    auto det = gp(d.ldet) + (size_t)b * RW;

    // the factors of the planes that depend on the state's time only, each
    // the expression the per-state formulas below would evaluate
    for (int a = tid; a < T1; a += NT) {
        sh.eB[a] = exp(tv.lnB[a]);
        sh.eN[a] = exp(tv.lnNegG1[a]);
        sh.eE2[a] = exp(tv.lnE2[a]);
        sh.low[a] = a > 0 ? exp(tv.lnE2[a] + tv.lnB[a - 1]) : 0.0;
        emit_time_terms(t, M, sh.mainlen, a, sh.et[a]);
    }
    for (int n = tid; n < N; n += NT) {
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

    __syncthreads();   // per-time factors complete

    // Per-state planes.  For a target state (n, b) with c = age(n), the
    // reference's same-branch entry tmatrix2[a][k] = get_time(a,b,c,same) -
    // get_time(a,b,0,other) (sample_thread.cpp:210-226, trans.h:89-128)
    // factorises as
    //   a <  b:  D[a] (exp(lnB[a]) - exp(lnNegG1[a]) - B_{c-1}) * E[b] exp(lnE2[b])
    //   a >  b:  D[a] * E[b] (low_b + G2[b] - ct_b)
    //   a == b:  D[b] E[b] (low_b + G3[b] - ct_b) + norecombs[b]
    // with B_{c-1} = exp(lnB[c-1]), low_b = exp(lnE2[b] + lnB[b-1]) and
    // ct_b = exp(lnG4[b] + lnB[c-1]) (lnG4 = lnE2), so the branch-local
    // sum is a prefix sum (a < b) and a suffix sum (a > b) along the branch.
    for (int p = tid; p < RW; p += NT) {
        const int s = (p & 63) * R + (p >> 6);
        double cf = 0, cd = 0, cg = 0, ck = 0, cj = 0, ce = 0, pr = 0;
        int inf = INFO_FIRST | INFO_LAST | 31;   // padding: time bin 31, whose TF is 0
        if (s < S) {
            const int n = node_of_state(sh.ni, N, s);
            const int a = sh.ag[n] + s - sh.ni.first[n];
            const int c = sh.ag[n], h = sh.ni.hi[n];
            const double Bc = c > 0 ? sh.eB[c - 1] : 0.0;
            const double low = sh.low[a];
            const double ct = c > 0 ? exp(tv.lnG4[a] + tv.lnB[c - 1]) : 0.0;
            cf = tv.D[a] * (sh.eB[a] - sh.eN[a] - Bc);
            cd = tv.D[a];
            cg = tv.E[a] * sh.eE2[a];
            ck = tv.E[a] * (low + tv.G2[a] - ct);
            // the diagonal (stay) entry dominates every state's value: it is
            // get_time(a, a, c, true) - get_time(a, a, 0, false), the
            // reference's tmatrix2 entry, in get_time's own expressions (its
            // low and c_term are low and ct above)
            const double term1 = tv.D[a] * tv.E[a];
            cj = (term1 * ((2 * (low + tv.G3[a])) - ct) + tv.norecombs[a]) -
                 term1 * (low + tv.G3[a]);
            EmitTerms e;
            emit_terms(t, M, sh.et[a], n, a, e);
            ce = e.inv;
            for (int x = 0; x < 3; x++) {
                gp(d.eterm)[6 * (size_t)(so + s) + x] = e.mut[x];
                gp(d.eterm)[6 * (size_t)(so + s) + 3 + x] = e.nomut[x];
            }
            if (b == 0) pr = prior_time(a, sh.L, M);
            inf = a | (a == c ? INFO_FIRST : 0) | (a == h ? INFO_LAST : 0) | INFO_VALID |
                  (c << 11) | (n << 16);
        }
        coef.put(CF, p, cf);
        coef.put(CD, p, cd);
        coef.put(CG, p, cg);
        coef.put(CK, p, ck);
        coef.put(CJ, p, cj);
        coef.put(CE, p, ce);
        coef.put(CRR, p, 0.0);
        coef.put(CCR, p, 0.0);
        coef.put(CDP, p, 0.0);
        det[p] = -1;
        info[p] = inf;
        if (b == 0) gp(d.lprior)[p] = pr;
    }

    const int TT = T1 * T1;
    for (int x = tid; x < TT; x += NT) {
        int a = x / T1, bb = x % T1;
        gp(d.tm)[(size_t)b * TT + x] = tv.get_time(a, bb, 0, false);
    }
    for (int x = tid; x < d.TC; x += NT) {
        int c = 0, base = 0;
        while (base + (T1 - c) * (T1 - c) <= x) {
            base += (T1 - c) * (T1 - c);
            c++;
        }
        int r = x - base, w = T1 - c;
        int a = c + r / w, bb = c + r % w;
        gp(d.tm2s)[(size_t)b * d.TC + x] = tv.get_time(a, bb, c, true);
    }

    if (b == 0) return;
    __syncthreads();   // planes cleared before the switch entries are set

    // switch matrix from block b-1 (calc_transition_probs_switch), entries
    // of the previous block's states at their lane-run positions
    const int so1 = cp(d.soff)[b - 1], S1 = so - so1;
    for (int j = tid; j < S1; j += NT) {
        int n1 = node_of_state(sh.pni, N, j);
        int t1 = sh.pag[n1] + j - sh.pni.first[n1];
        int tgt = sw_determ(pt, t, sh.ni, spr, sh.map, n1, t1);
        const int p = lane_pos(j, R);
        det[p] = tgt;
        coef.put(CDP, p, tgt >= 0 ? sw_determprob(pt, M, sh.PL, spr, sh.cx, n1, t1) : 0.0);
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
            coef.put(CRR, lane_pos(stay, R), sw_recomb_recoal(
                pt, M, sh.PL, spr, spr.ct, spr.rn, spr.rt, sh.cx.rpa, sh.cx.last_len));
        if (esc != -1)
            coef.put(CRR, lane_pos(esc, R), sw_recomb_recoal(
                pt, M, sh.PL, spr, spr.ct, spr.rn, spr.rt, spr.rt, sh.cx.last_len));
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
            coef.put(CCR, lane_pos(k, R), sw_recomb_recoal(pt, M, sh.PL, spr, t2, spr.cn, spr.ct,
                                                           sh.cx.rpa, sh.cx.last_len));
        }
    }
}

//===========================================================================
// emissions of variant columns, written as lane-run rows

// Felsenstein pruning (pruning) by one wave, level by level: the inner
// partials of all nodes of one height at once (children first), then the
// outer partials of all nodes of one depth (parents first).  Every node's
// arithmetic is the sequential version's, so the results are identical;
// one wave's LDS operations run in order, so a level sees the one before.
__device__ inline void pruning_wave(const TreeView &t, const int *lvl, const char *bases,
                                    const double *muts, const double *nomuts, double *inner,
                                    double *outer)
{
    const int lane = threadIdx.x;
    int hmax = 0, dmax = 0;
    for (int j = lane; j < t.n; j += 64) {
        hmax = max(hmax, lvl[j] & 0xffff);
        dmax = max(dmax, lvl[j] >> 16);
    }
    for (int off = 32; off > 0; off >>= 1) {
        hmax = max(hmax, __shfl_xor(hmax, off));
        dmax = max(dmax, __shfl_xor(dmax, off));
    }
    for (int h = 0; h <= hmax; h++) {
        for (int j = lane; j < t.n; j += 64) {
            if ((lvl[j] & 0xffff) != h) continue;
            if (t.c0[j] == -1) {
                leaf_partial(bases[j], inner + 4 * j);
                continue;
            }
            const int c1 = t.c0[j], c2 = t.c1[j];
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
        __syncthreads();
    }
    for (int dd = 0; dd <= dmax; dd++) {
        for (int j = lane; j < t.n; j += 64) {
            if ((lvl[j] >> 16) != dd) continue;
            if (j == t.root) {
                for (int a = 0; a < 4; a++) outer[4 * j + a] = 1.0;
                continue;
            }
            const int sib = t.sibling(j);
            const int parent = t.parent[j];
            for (int a = 0; a < 4; a++) {
                double p1 = 0.0, p2 = 0.0;
                for (int b = 0; b < 4; b++) {
                    const double ms = (a == b) ? nomuts[sib] : muts[sib];
                    p1 += inner[4 * sib + b] * ms;
                    if (parent != t.root) {
                        const double mp = (a == b) ? nomuts[parent] : muts[parent];
                        p2 += outer[4 * parent + b] * mp;
                    }
                }
                outer[4 * j + a] = (parent != t.root) ? p1 * p2 : p1;
            }
        }
        __syncthreads();
    }
}

// LDS of k_var_emit_b for trees of up to n nodes: the partials and branch
// probabilities (FP64), the tree arrays and levels, the leaf bases.  Sized
// by the batch's largest tree, so that small trees leave room for more
// waves on a CU.
inline size_t var_emit_lds(int n)
{
    return sizeof(double) * (10 * (size_t)n + 4) + sizeof(int) * 5 * (size_t)n + n;
}

__global__ __launch_bounds__(64) void k_var_emit_b(const DevInst *__restrict__ D,
                                                   const int *__restrict__ voff, int ninst,
                                                   int nmax)
{
    extern __shared__ double ve_dyn[];
    double *inner = ve_dyn, *outer = inner + 4 * nmax;
    double *muts = outer + 4 * nmax, *nomuts = muts + nmax, *in2 = nomuts + nmax;
    int *par = (int *)(in2 + 4), *ch0 = par + nmax, *ch1 = ch0 + nmax, *ag = ch1 + nmax,
        *lvl = ag + nmax;
    char *bases = (char *)(lvl + nmax);
    const int inst = find_owner(voff, ninst, (int)blockIdx.x);
    const DevInst &d = D[inst];
    if (d.hostprep) return;   // uniform over the wave
    const int v = blockIdx.x - voff[inst];
    const int lane = threadIdx.x;
    const int N = d.N, R = d.R, RW = 64 * R;
    const int col = gp(d.vcol)[v];
    const int b = find_owner(cp(d.bstart), d.B, col);
    const int nleaves = (N + 1) / 2;
    for (int i = lane; i < N; i += 64) {
        size_t x = (size_t)b * N + i;
        par[i] = gp(d.parent)[x];
        ch0[i] = gp(d.c0)[x];
        ch1[i] = gp(d.c1)[x];
        ag[i] = gp(d.age)[x];
        lvl[i] = gp(d.order)[x];
        muts[i] = gp(d.nmut)[2 * x];
        nomuts[i] = gp(d.nmut)[2 * x + 1];
        if (i < nleaves) bases[i] = gp(d.seq)[(size_t)i * d.L + col];
    }
    __syncthreads();
    const int root = gp(d.roots)[b];
    {
        TreeView t{N, root, par, ch0, ch1, ag};
        pruning_wave(t, lvl, bases, muts, nomuts, inner, outer);
        if (lane == 0) leaf_partial(gp(d.seq)[(size_t)nleaves * d.L + col], in2);
    }
    __syncthreads();
    const int so = cp(d.soff)[b], S = cp(d.soff)[b + 1] - so;
    auto row = gp(d.vemit) + (size_t)v * RW;
    for (int p = lane; p < RW; p += 64) {
        const int k = (p & 63) * R + (p >> 6);
        double e = 0.0;
        if (k < S) {
            int n = gp(d.snode)[so + k];
            const double *et = d.eterm + 6 * (size_t)(so + k);
            e = emit_variant(in2, inner + 4 * n, outer + 4 * n, n == root, et, et + 3);
        }
        row[p] = e;
    }
}

//===========================================================================
// forward recursion, one wave per instance (arghmm_forward_alg,
// sample_thread.cpp:394-460; arghmm_forward_block :186-296;
// arghmm_forward_switch :345-389)

// NT1 is a compile-time bound on the T-1 time bins (19 for the usual 20
// time points), so the per-time loops are fully unrolled without padding.
//
// Per block, each lane keeps in registers the factor planes of its R
// states, two 0/1 factors per state (not the first / not the last state of
// its branch: they cut the prefix and suffix sums), and whether the lane
// holds a branch start / end (the carry stops there).  Per state it also
// keeps its time, which gives both the lane holding the other-branch term
// of its time and where its value goes in the per-time partial sums;
// padding states use time bin 31, whose term is 0, so they come out 0 with
// no mask.  The 0/1 factors are floats in both precisions: FP64 runs apply
// them as selects, which keeps their registers at one per state.
template <class Real, int R, int NT1>
struct FwLane {
    Real cf[R], cd[R], cg[R], ck[R], cj[R], ce[R];
    float nf[R], nl[R];
    int boff[R];     // 4 * time: ds_bpermute address of the lane holding TF[time]
    Real m1, m2;
    Real tmc[NT1];   // lane b < T-1: column b of the (a,b) table (FP32)
};

// x times a 0/1 factor m (a select in FP64: the same value)
template <class Real>
__device__ __forceinline__ Real keep(Real x, float m)
{
    if constexpr (sizeof(Real) == 4) return x * m;
    else return (m != 0.f) ? x : Real(0);
}

// The (a,b) table lives in registers (lane b holds column b) for FP32 and
// in LDS for FP64, where registers are short.
template <class Real>
constexpr bool tm_in_lds()
{
    return sizeof(Real) == 8;
}

// Per-time sums are accumulated in GCOPY copies of the 32 time bins, one
// copy per group of 64 / GCOPY consecutive lanes, so that the LDS atomics of
// one instruction seldom meet at one address (colliding atomics are
// serialised).  Bin a, copy c sits at GP[a * GCOPY + c].  The sums are FP64
// in both precisions: ds_add_f64 is much faster than ds_add_f32 here.
constexpr int GCOPY = 16;

template <class Real>
__device__ __forceinline__ const AW_GLOBAL Real *planes(const DevInst &d, int b, int RW)
{
    if (sizeof(Real) == 4)
        return (const AW_GLOBAL Real *)(gp(d.lcoef32) + (size_t)b * NCOEF * RW);
    return (const AW_GLOBAL Real *)(gp(d.lcoef) + (size_t)b * NCOEF * RW);
}

template <class Real, int R, int NT1>
__device__ __forceinline__ void fw_load_block(const DevInst &d, int b, FwLane<Real, R, NT1> &fl,
                                              Real *TMS)
{
    constexpr int RW = 64 * R;
    const int lane = threadIdx.x, T1 = d.T - 1;
    // uniform bases, per-lane index: loads use the scalar-base address form
    const AW_GLOBAL Real *coef = planes<Real>(d, b, RW);
    const AW_GLOBAL int *info = gp(d.linfo) + (size_t)b * RW;
    int ti[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int x = r * 64 + lane;
        fl.cf[r] = coef[CF * RW + x];
        fl.cd[r] = coef[CD * RW + x];
        fl.cg[r] = coef[CG * RW + x];
        fl.ck[r] = coef[CK * RW + x];
        fl.cj[r] = coef[CJ * RW + x];
        fl.ce[r] = coef[CE * RW + x];
        ti[r] = info[x];
    }
    bool first = false, last = false;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int t = ti[r] & INFO_TIME;
        fl.nf[r] = (ti[r] & INFO_FIRST) ? 0.f : 1.f;
        fl.nl[r] = (ti[r] & INFO_LAST) ? 0.f : 1.f;
        fl.boff[r] = t * 4;
        first |= (ti[r] & INFO_FIRST) != 0;
        last |= (ti[r] & INFO_LAST) != 0;
    }
    fl.m1 = first ? Real(0) : Real(1);
    fl.m2 = last ? Real(0) : Real(1);
    // Loaded values go straight into their registers: selecting or copying
    // a value whose load is still in flight makes the compiler wait for it
    // at once.  Indices are clamped instead; entries a >= T-1 meet a zero
    // per-time sum and lanes >= T-1 compute terms nobody uses.
    const AW_GLOBAL double *tmb = gp(d.tm) + (size_t)b * T1 * T1;
    const int lc = min(lane, T1 - 1);
    if (tm_in_lds<Real>()) {
        // TMS[a * 32 + b] = P(a -> b); lanes 32..63 duplicate lanes 0..31
        double t[NT1];
#pragma unroll
        for (int a = 0; a < NT1; a++) t[a] = tmb[min(a, T1 - 1) * T1 + lc];
#pragma unroll
        for (int a = 0; a < NT1; a++)
            if (lane < 32) TMS[a * 32 + lane] = (Real)t[a];
    } else {
#pragma unroll
        for (int a = 0; a < NT1; a++) fl.tmc[a] = (Real)tmb[min(a, T1 - 1) * T1 + lc];
    }
}

// variant row v, clamped to the last row (an instance without variant
// columns has one allocated row that is never used)
template <int R>
__device__ __forceinline__ void vrow_load(const DevInst &d, int v, double (&ev)[R])
{
    constexpr int RW = 64 * R;
    const AW_GLOBAL double *row = gp(d.vemit) + (size_t)max(min(v, d.nvar - 1), 0) * RW;
#pragma unroll
    for (int r = 0; r < R; r++) ev[r] = row[r * 64 + threadIdx.x];
}

// 1 / x: FP32 tables use the hardware reciprocal (within 1 ulp; the
// normaliser itself, and so the log-likelihood, is kept exactly), FP64
// tables the correctly rounded division
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

template <class Real>
__device__ __forceinline__ void lds_add(Real *p, Real v)
{
    (void)atomicAdd(p, v);
}

__device__ __forceinline__ float readlane_r(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double readlane_r(double x, int l)
{
    return readlane_d(x, l);
}

// x of the lane addr / 4 (ds_bpermute)
__device__ __forceinline__ float bperm_r(float x, int addr)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x)));
}
__device__ __forceinline__ double bperm_r(double x, int addr)
{
    const long long v = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(v >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// variant rows staged per LDS refill (at most 16 KB of them), and rows per
// load batch of a refill (bounded by the registers they take)
template <class Real, int R>
constexpr int fw_vst()
{
    return (16384 / (64 * R * (int)sizeof(Real))) < 8 ? (16384 / (64 * R * (int)sizeof(Real))) : 8;
}
template <class Real, int R>
constexpr int fw_vch()
{
    return (32 / R) < 2 ? 2 : (32 / R) > fw_vst<Real, R>() ? fw_vst<Real, R>() : (32 / R);
}

// LDS of the forward kernel, in Reals: partial sums, the previous column in
// state order, the switch targets (plus one dummy slot per lane), staged
// variant rows, the (a,b) table
template <class Real, int R>
constexpr int fw_lds_reals()
{
    return 32 * GCOPY * (int)(sizeof(double) / sizeof(Real)) + 64 * R + (64 * R + 64) +
           fw_vst<Real, R>() * 64 * R + 31 * 32;
}

// A forward tile: columns [c0, c1) of an instance computed by one wave.
// mode 0 starts from the prior at column 0.  mode 1 starts at c0 from a
// uniform vector and stores nothing before cs (a warm-up: the filter
// forgets its start).  mode 2 starts at c0 from the column held in boundary
// slot bin (written by the tile before).  A tile with bout >= 0 also
// computes column c1 into boundary slot bout, for the check of the next
// tile's start.
struct FwTile {
    int inst;
    int c0, cs, c1;
    int mode;
    int bout, bin;
};
constexpr int BND_STRIDE = 64 * 32;   // Reals per boundary slot (the largest row)

// wave-wide maximum
__device__ __forceinline__ double wave_max(double x)
{
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off));
    return x;
}

// columns between the checks of a re-run tile against its stored columns
constexpr int FW_RECHECK = 128;

template <class Real, int R, int NT1>
__device__ __forceinline__ void fw_instance(const DevInst d, Real *sm, const FwTile tile,
                                            AW_GLOBAL Real *bnd, AW_GLOBAL double *bnorm,
                                            double tol, AW_GLOBAL int *early)
{
    constexpr int RW = 64 * R;
    constexpr int VST = fw_vst<Real, R>();
    constexpr int VCH = fw_vch<Real, R>();
    const int lane = threadIdx.x;
    const int T1 = d.T - 1, L = d.L;
    // carry steps: a branch has at most T-1 states, so it reaches back at
    // most ceil((T-2)/R) lanes from the lane holding its last state
    const int K = (T1 - 1 + R - 1) / R;
    double *GP = (double *)sm;       // 32 x GCOPY: per-time partial sums
    // byte offset of this lane's copy within a time bin's GCOPY copies
    const int gslot = (lane / (64 / GCOPY)) * (int)sizeof(double);
    // the half of bin lane>>1 that this lane reduces, and its read rotation
    double *gq = GP + (lane >> 1) * GCOPY + (lane & 1) * (GCOPY / 2);
    const int grot = (lane >> 1) & (GCOPY / 2 - 1);
    Real *XS = (Real *)(GP + 32 * GCOPY);   // RW: previous column in state order (switch)
    Real *NW = XS + RW;              // RW + 64: switch targets, then a dummy slot per lane
    Real *VS = NW + RW + 64;         // VST rows: staged variant-column emissions
    Real *TMS = VS + VST * RW;       // 31 x 32: the (a,b) table (FP64)
    AW_GLOBAL Real *alpha = (AW_GLOBAL Real *)d.alpha;
    auto norms = gp(d.norms);

    const int c0 = tile.c0, cs = tile.cs;
    // the last column computed: c1 - 1, or c1 into the boundary slot
    const int clast = tile.bout >= 0 ? tile.c1 : tile.c1 - 1;
    FwLane<Real, R, NT1> fl;
    int b = find_owner(cp(d.bstart), d.B, c0);
    int bend = cp(d.bstart)[b] + cp(d.blen)[b];
    fw_load_block<Real, R, NT1>(d, b, fl, TMS);
#pragma unroll
    for (int k = 0; k < 32 * GCOPY / 64; k++) GP[k * 64 + lane] = 0.0;
    // Column classes of a window of 64 columns, one per lane, read with
    // readlane; the next window's are loaded a window ahead.  vmcnt counts
    // the table stores too, so any wait on a load inside the column loop
    // would also wait for the stores of the column before: the windows are
    // an outer loop, and the one wait per window is taken there.
    const AW_GLOBAL int *clsg = gp(d.cls);
    int wb = c0 & ~63;
    int clsv = clsg[min(wb + lane, L - 1)];
    int clsn = clsg[min(wb + 64 + lane, L - 1)];
    // Variant rows (row index = the column's class) are consumed in order
    // and staged in LDS, VST at a time, so that a refill costs a few waits
    // and the columns read their row from LDS.
    int vbase = 0, vend = 0;

    Real x[R], tfs[R], v[R];
    Real sc = Real(1);   // the previous column is x * sc

    // A re-run tile (mode 2) stops once a column it computes agrees with
    // the column stored by its speculative run (relative tolerance tol):
    // the filter is contractive, so the stored columns after it, the
    // normalisers and the boundary column stay within tolerance.
    bool rejoined = false;

    // per-time sums, normaliser, stored row, other-branch term of the next
    // column
    auto finish = [&](int col, bool first) {
#pragma unroll
        for (int r = 0; r < R; r++)
            lds_add((double *)((char *)GP + fl.boff[r] * (2 * GCOPY) + gslot), (double)v[r]);
        // lane l takes the copies of bin l>>1 in its half (l&1), leaving 0
        // (one exchange each; the LDS operations of one wave run in order).
        // Step k reads copy (k + l/2) mod 8 of the half: the lanes of one
        // step then spread over the banks (copy k of every bin would sit in
        // two banks), which also fixes each lane's summation order
        double q[GCOPY / 2];
#pragma unroll
        for (int k = 0; k < GCOPY / 2; k++)
            q[k] = __longlong_as_double((long long)__hip_atomic_exchange(
                (unsigned long long *)(gq + ((k + grot) & (GCOPY / 2 - 1))), 0ull,
                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
#pragma unroll
        for (int w = 1; w < GCOPY / 2; w *= 2)
#pragma unroll
            for (int k = 0; k < GCOPY / 2; k += 2 * w) q[k] += q[k + w];
        const Real g = (Real)(q[0] + dpp_r<DPP_QUAD_SWAP1>(q[0]));   // lanes 2a, 2a+1: G[a]
        // four interleaved partial sums keep the dependency chains short
        Real nq[4] = {Real(0), Real(0), Real(0), Real(0)};
        Real tq[4] = {Real(0), Real(0), Real(0), Real(0)};
#pragma unroll
        for (int a = 0; a < NT1; a++) {
            const Real ga = readlane_r(g, 2 * a);
            nq[a & 3] += ga;
            tq[a & 3] += ga * (tm_in_lds<Real>() ? TMS[a * 32 + (lane & 31)] : fl.tmc[a]);
        }
        const Real norm = (nq[0] + nq[1]) + (nq[2] + nq[3]);
        const Real t = (tq[0] + tq[1]) + (tq[2] + tq[3]);
        // column 0 is stored unnormalised (scale 1); a restarted tile's first
        // column is the boundary column as stored by the tile before
        const Real nsc = (col == 0 || (first && tile.mode == 2)) ? Real(1) : recip(norm);
        if (col >= cs && col < tile.c1) {
            AW_GLOBAL Real *row = alpha + (size_t)col * RW;
            if (tile.mode == 2 && !first && (col - tile.c0) % FW_RECHECK == 0) {
                double dmax = 0.0, amax = 0.0;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const double nv = (double)(v[r] * nsc), ov = (double)row[r * 64 + lane];
                    dmax = fmax(dmax, fabs(nv - ov));
                    amax = fmax(amax, fabs(nv));
                }
                dmax = wave_max(dmax);
                amax = wave_max(amax);
                rejoined = dmax <= tol * amax;
            }
#pragma unroll
            for (int r = 0; r < R; r++) row[r * 64 + lane] = v[r] * nsc;
            if (lane == 0) {
                double nv = (col == 0) ? ((L == 1) ? (double)norm : 1.0) : (double)norm;
                if (first && tile.mode == 2) nv = bnorm[tile.bin];
                norms[col] = nv;
            }
        } else if (col == tile.c1) {
            AW_GLOBAL Real *row = bnd + (size_t)tile.bout * BND_STRIDE;
#pragma unroll
            for (int r = 0; r < R; r++) row[r * 64 + lane] = v[r] * nsc;
            if (lane == 0) bnorm[tile.bout] = (double)norm;
        }
        // TF[b] lives in lane b; lanes >= T-1 (padding bin 31) hold 0
        const Real tv = (lane < T1) ? t * nsc : Real(0);
#pragma unroll
        for (int r = 0; r < R; r++) {
            tfs[r] = bperm_r(tv, fl.boff[r]);
            x[r] = v[r];
        }
        sc = nsc;
    };

    // first column.  mode 0: the unnormalised prior, not multiplied by the
    // emission (calc_state_priors, trans.cpp:785-831); padding positions
    // hold 0.  mode 1: 1 on every state.  mode 2: the boundary column.
    if (tile.mode == 0) {
        const AW_GLOBAL double *pr = gp(d.lprior);
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = (Real)pr[r * 64 + lane];
    } else if (tile.mode == 1) {
        const int S = cp(d.soff)[b + 1] - cp(d.soff)[b];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = (lane * R + r < S) ? Real(1) : Real(0);
    } else {
        const AW_GLOBAL Real *src = bnd + (size_t)tile.bin * BND_STRIDE;
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[r * 64 + lane];
    }
    finish(c0, true);

    // an empty asm reading the register makes the compiler complete its
    // load here rather than at a use inside the column loop
    asm volatile("" : "+v"(clsv));
    int col = c0 + 1;
    while (col <= clast && !rejoined) {
        if (col >= wb + 64) {
            wb += 64;
            clsv = clsn;
            asm volatile("" : "+v"(clsv));
            clsn = clsg[min(wb + 64 + lane, L - 1)];
        }
        const int wend = min(wb + 63, clast);
    for (; col <= wend; col++) {
        const int cls = __builtin_amdgcn_readlane(clsv, col & 63);
        const bool var = cls >= 0;
        Real ev[R];
        if (var) {
            if (cls >= vend) {
                vbase = cls;
                vend = min(cls + VST, d.nvar);
#pragma unroll
                for (int k0 = 0; k0 < VST; k0 += VCH) {
                    // VST need not be a multiple of VCH
                    double buf[VCH][R];
#pragma unroll
                    for (int k = 0; k < VCH; k++)
                        if (k0 + k < VST) vrow_load<R>(d, vbase + k0 + k, buf[k]);
#pragma unroll
                    for (int k = 0; k < VCH; k++)
#pragma unroll
                        for (int r = 0; r < R; r++)
                            if (k0 + k < VST) VS[(k0 + k) * RW + r * 64 + lane] = (Real)buf[k][r];
                }
            }
#pragma unroll
            for (int r = 0; r < R; r++) ev[r] = VS[(cls - vbase) * RW + r * 64 + lane];
        }

        if (col == bend) {
            // switch column: deterministic transitions scattered into NW
            // (the recombination and recoalescence sources excluded), then
            // the two dense rows.  Every load of the column is issued before
            // the first use, branch-free, so the column waits once.
            const int nb = b + 1;
            const int so0 = cp(d.soff)[b], so1 = cp(d.soff)[nb];
            const int S1 = so1 - so0, S2 = cp(d.soff)[nb + 1] - so1;
            const int rs = cp(d.srcs)[2 * nb], cs = cp(d.srcs)[2 * nb + 1];
            const AW_GLOBAL Real *swn = planes<Real>(d, nb, RW);
            const AW_GLOBAL int *detn = gp(d.ldet) + (size_t)nb * RW;
            int dt[R];
            Real dpv[R], rr[R], cr[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int x = r * 64 + lane;
                dt[r] = detn[x];
                dpv[r] = swn[CDP * RW + x];
                rr[r] = swn[CRR * RW + x];
                cr[r] = swn[CCR * RW + x];
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                XS[lane * R + r] = x[r];   // padding slots (s >= S1) are never read
                NW[r * 64 + lane] = Real(0);
            }
            b = nb;
            bend += cp(d.blen)[nb];
            fw_load_block<Real, R, NT1>(d, b, fl, TMS);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int s = lane * R + r;
                const bool ok = s < S1 && dt[r] >= 0 && s != rs && s != cs;
                lds_add(&NW[ok ? dt[r] : RW + lane], ok ? x[r] * dpv[r] : Real(0));
            }
            const Real prs = (rs != -1) ? XS[max(rs, 0)] * sc : Real(0);
            const Real pcs = (cs != -1) ? XS[max(cs, 0)] * sc : Real(0);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int s = lane * R + r;
                const Real e = var ? ev[r] : (cls == -1) ? fl.ce[r] : Real(1);
                Real acc = (s < S2) ? NW[s] * sc : Real(0);
                if (rs != -1 && rr[r] > Real(0)) acc += prs * rr[r];
                if (cs != -1 && cr[r] > Real(0)) acc += pcs * cr[r];
                v[r] = (s < S2) ? acc * e : Real(0);
            }
        } else {
            // same-branch prefix sums (younger states) with the carry from
            // the lanes before, suffix sums (older states) with the carry
            // from the lanes after
            Real P[R], Q[R];
            Real p = Real(0);
#pragma unroll
            for (int r = 0; r < R; r++) {
                P[r] = keep(p, fl.nf[r]);
                p = fl.cf[r] * x[r] + P[r];
            }
            Real q = Real(0);
#pragma unroll
            for (int r = R - 1; r >= 0; r--) {
                Q[r] = keep(q, fl.nl[r]);
                q = fl.cd[r] * x[r] + Q[r];
            }
            Real c1 = Real(0), c2 = Real(0);
            for (int k = 0; k < K; k++) {
                c1 = dpp_r<DPP_WAVE_SHR1>(fl.m1 * c1 + p);
                c2 = dpp_r<DPP_WAVE_SHL1>(fl.m2 * c2 + q);
            }
            // the carry reaches the slots before the lane's first branch
            // start (prefix) / after its last branch end (suffix)
            Real f1 = Real(1);
#pragma unroll
            for (int r = 0; r < R; r++) {
                f1 = keep(f1, fl.nf[r]);
                P[r] += c1 * f1;
            }
            Real f2 = Real(1);
#pragma unroll
            for (int r = R - 1; r >= 0; r--) {
                f2 = keep(f2, fl.nl[r]);
                Q[r] += c2 * f2;
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                const Real e = var ? ev[r] : (cls == -1) ? fl.ce[r] : Real(1);
                const Real same = (fl.cg[r] * P[r] + fl.ck[r] * Q[r]) + fl.cj[r] * x[r];
                v[r] = e * (tfs[r] + sc * same);
            }
        }
        finish(col, false);
        if (rejoined) break;
    }
    }
    // the column after the one that rejoined (never 0), else 0
    if (early && lane == 0) *early = rejoined ? col + 1 : 0;
}

// One launch covers every instance of a batch, whatever its states-per-lane
// count: waves of different R run side by side from the start, with no
// dependence on how streams map to hardware queues.
// REDO marks the launches of re-run tiles (a separate symbol, so that
// profiles report the main forward launch on its own).  BIG: the instances
// with 32 states per lane (blocks of 1025-2048 states) run in a symbol of
// their own, launched over the same tiles, so that their register demand
// (spilled to AGPRs and scratch) does not weigh on the others.
template <class Real, int NT1, bool REDO, bool BIG>
__global__ __launch_bounds__(64) void k_fw_all(const DevInst *__restrict__ D,
                                               const FwTile *__restrict__ tiles,
                                               Real *__restrict__ bnd, double *__restrict__ bnorm,
                                               double tol, int *__restrict__ early)
{
    extern __shared__ __align__(16) double fw_dyn[];
    const FwTile t = tiles[blockIdx.x];
    const DevInst &d = D[t.inst];
    Real *sm = (Real *)fw_dyn;
    AW_GLOBAL Real *bg = (AW_GLOBAL Real *)bnd;
    AW_GLOBAL double *ng = (AW_GLOBAL double *)bnorm;
    AW_GLOBAL int *eg = early ? (AW_GLOBAL int *)(early + blockIdx.x) : nullptr;
    if ((d.R > 16) != BIG) return;   // the other symbol's tile
    if constexpr (BIG) {
        fw_instance<Real, 32, NT1>(d, sm, t, bg, ng, tol, eg);
    } else {
        switch (d.R) {
        case 1: fw_instance<Real, 1, NT1>(d, sm, t, bg, ng, tol, eg); break;
        case 2: fw_instance<Real, 2, NT1>(d, sm, t, bg, ng, tol, eg); break;
        case 4: fw_instance<Real, 4, NT1>(d, sm, t, bg, ng, tol, eg); break;
        case 6: fw_instance<Real, 6, NT1>(d, sm, t, bg, ng, tol, eg); break;
        case 8: fw_instance<Real, 8, NT1>(d, sm, t, bg, ng, tol, eg); break;
        case 12: fw_instance<Real, 12, NT1>(d, sm, t, bg, ng, tol, eg); break;
        default: fw_instance<Real, 16, NT1>(d, sm, t, bg, ng, tol, eg); break;
        }
    }
}

// Check of a speculative forward tile's start: the boundary column computed
// by the tile before (slot bin) against the tile's own stored column c0,
// both normalised.  flag = 1 when they differ by more than tol relative to
// the column's largest entry, or the normalisers by more than tol.
template <class Real>
__global__ __launch_bounds__(64) void k_fw_check(const DevInst *__restrict__ D,
                                                 const FwTile *__restrict__ tiles,
                                                 const int *__restrict__ idx,
                                                 const Real *__restrict__ bnd,
                                                 const double *__restrict__ bnorm, double tol,
                                                 int *__restrict__ flags)
{
    const FwTile t = tiles[idx[blockIdx.x]];
    const DevInst &d = D[t.inst];
    const int RW = 64 * d.R, lane = threadIdx.x;
    const Real *a = (const Real *)d.alpha + (size_t)t.cs * RW;
    const Real *q = bnd + (size_t)t.bin * BND_STRIDE;
    double dmax = 0.0, amax = 0.0;
    for (int x = lane; x < RW; x += 64) {
        dmax = fmax(dmax, fabs((double)a[x] - (double)q[x]));
        amax = fmax(amax, fabs((double)q[x]));
    }
    for (int off = 32; off > 0; off >>= 1) {
        dmax = fmax(dmax, __shfl_xor(dmax, off));
        amax = fmax(amax, __shfl_xor(amax, off));
    }
    if (lane == 0) {
        const double n0 = gp(d.norms)[t.cs], n1 = bnorm[t.bin];
        const bool bad = !(dmax <= tol * amax) || !(fabs(n0 - n1) <= tol * fabs(n1));
        flags[blockIdx.x] = bad ? 1 : 0;
    }
}

// log-likelihood: sum of log column normalisers
__global__ __launch_bounds__(256) void k_lnl_b(const DevInst *__restrict__ D,
                                               const int *__restrict__ list)
{
    __shared__ double part[256];
    const DevInst &d = D[list[blockIdx.x]];
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

//===========================================================================
// stochastic traceback (stochastic_traceback, sample_thread.cpp:522-569;
// sample_hmm_posterior_step :506-519; sample, common.h:272-290)

struct Chunk {
    int inst, c0, c1;        // columns [c0, c1) of instance inst
    int cw;                  // column where the speculative chain starts (>= c1)
    int first, last;         // first / last tile of its instance
    long long coff;          // the tile's alternative paths in the option buffer
};

// Options of a traceback tile.  Option 0 is the speculative chain: a state
// sampled from column cw alone, a warm-up back to column c1 (not stored),
// then the tile's columns into the path.  Options 1 .. K-1 are chains of the
// same tile started at column c1 from the other states of the branch that
// option 0 holds there (the most probable first, by the forward column):
// two backward chains on one branch at neighbouring times can run side by
// side for tens of thousands of columns with the same uniforms, and one of
// them is then the true chain.  Each option records the state it starts
// from at c1 and the state it reaches at c0; once the true state at c1 is
// known, an option that started from it is the tile's exact path.
constexpr int TB_KMAX = 4;

template <int R>
struct TbLane {
    int ti[R];
    int node[R];   // node of the state (info_node)
    int off1[R];   // a * (T-1): other-branch entry P(a -> b) at off1 + b
    int off2[R];   // same-branch entry P(a -> b) at off2 + b
    int dt[R];     // deterministic switch target into the next block
    double dp[R];  // and its weight
};

// LDS tables of block b: the (a,b) table followed by the (c,a,b) table, in
// the run's precision, copied with every load of a lane in flight at once
// (TCP per round)
template <class Real, int R>
__device__ __forceinline__ void tb_load_block(const DevInst &d, int b, TbLane<R> &tl, Real *tab)
{
    constexpr int RW = 64 * R;
    const int lane = threadIdx.x, T1 = d.T - 1, TT = T1 * T1;
    const AW_GLOBAL int *info = gp(d.linfo) + (size_t)b * RW;
    // switch into block b+1 from this block's states (the last block
    // re-reads its own entries, never used)
    const int bn = min(b + 1, d.B - 1);
    const AW_GLOBAL int *detn = gp(d.ldet) + (size_t)bn * RW;
    const AW_GLOBAL Real *dpn = planes<Real>(d, bn, RW) + CDP * RW;
    const AW_GLOBAL double *tm = gp(d.tm) + (size_t)b * TT;
    const AW_GLOBAL double *tm2 = gp(d.tm2s) + (size_t)b * d.TC;
    const int n = TT + d.TC;
    constexpr int TCP = sizeof(Real) == 4 ? 48 : 24;
    for (int base = 0; base < n; base += 64 * TCP) {
        double t[TCP];
#pragma unroll
        for (int k = 0; k < TCP; k++) {
            const int x = min(base + k * 64 + lane, n - 1);
            const AW_GLOBAL double *src = (x < TT) ? tm + x : tm2 + (x - TT);
            t[k] = *src;
        }
        if (base == 0) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                tl.ti[r] = info[r * 64 + lane];
                tl.dt[r] = detn[r * 64 + lane];
                tl.dp[r] = dpn[r * 64 + lane];
            }
        }
        // entries past the end repeat the last one (same index, same value):
        // a per-entry condition would become a branch per store
#pragma unroll
        for (int k = 0; k < TCP; k++) tab[min(base + k * 64 + lane, n - 1)] = (Real)t[k];
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        // padding slots carry time bin 31: clamp so their (zero-weighted)
        // table reads stay inside the tables
        const int a = min(tl.ti[r] & INFO_TIME, T1 - 1), c = info_age(tl.ti[r]);
        tl.node[r] = info_node(tl.ti[r]);
        tl.off1[r] = a * T1;
        tl.off2[r] = TT + tc_base(c, T1) + (a - c) * (T1 - c) - c;
    }
}

// First index k of the lane-run weights (lane l holds states l*R..l*R+R-1)
// whose cumulative sum reaches u * total; S-1 if none does.  Returns k and
// the info word of state k in kinfo.  pre[r] = w[0] + ... + w[r] of the
// lane.  Branch-free: the owner lane counts the prefixes below the pick.
template <int R>
__device__ __forceinline__ int lane_sample(const double (&pre)[R], const int (&ti)[R], int S,
                                           double u, int &kinfo)
{
    const int lane = threadIdx.x;
    const double loc = pre[R - 1];
    const double inc = wave_incl_scan(loc);
    const double excl = dpp_r<DPP_WAVE_SHR1>(inc);
    const double total = readlane_d(inc, 63);
    const double pick = u * total;
    const bool has = lane * R < S;
    const unsigned long long m = __ballot(has && excl + loc >= pick);
    const int owner = m ? (int)__ffsll((long long)m) - 1 : (S - 1) / R;
    const int nloc = min(R, S - lane * R);
    const double thr = pick - excl;
    int cnt = 0;
#pragma unroll
    for (int r = 0; r < R; r++) cnt += (pre[r] < thr) ? 1 : 0;
    const int rsel = m ? min(cnt, nloc - 1) : nloc - 1;
    // select ti[rsel] with masks (a compare-and-branch form becomes a
    // dynamically indexed scratch array)
    int inf = 0;
#pragma unroll
    for (int r = 0; r < R; r++) inf |= ti[r] & -(int)(r == rsel);
    kinfo = __builtin_amdgcn_readlane(inf, owner);
    return __builtin_amdgcn_readlane(lane * R + rsel, owner);
}

// forward-table rows staged per traceback batch: about TB_ROWREG registers
// of a lane hold the next batch.  Fewer rows per batch mean less LDS per
// tile: at 48 an FP32 tile with 20 time points takes about 35 KB, so four
// tiles fit on a CU (three at 64); at 32 the batches of R = 12 and 16 are
// too short to hide the row loads.
constexpr int TB_ROWREG = 48;
template <class Real, int R>
constexpr int tb_batch()
{
    return (TB_ROWREG / (R * (int)(sizeof(Real) / 4))) > 16 ? 16
         : (TB_ROWREG / (R * (int)(sizeof(Real) / 4))) < 2 ? 2
                                                           : (TB_ROWREG / (R * (int)(sizeof(Real) / 4)));
}
// host mirror of tb_batch
inline int tb_batch_host(int R, size_t rsz)
{
    return std::max(2, std::min(16, TB_ROWREG / (R * (int)(rsz / 4))));
}

template <class Real, int R>
__device__ __forceinline__ void row_load(AW_GLOBAL const Real *alpha, int col, Real (&rv)[R])
{
    constexpr int RW = 64 * R;
    AW_GLOBAL const Real *row = alpha + (size_t)col * RW;
#pragma unroll
    for (int r = 0; r < R; r++) rv[r] = row[r * 64 + threadIdx.x];
}

// The state of a backward chain: the state k sampled at the column last
// visited, its info word, and the block of that column.
struct Chain {
    int k, info, blk;
};

// Stored paths a re-run compares its samples with (option 0 in the path,
// the alternatives in the option buffer), each indexed by column; null for
// an option not compared.
struct Cmp {
    const AW_GLOBAL int *p[TB_KMAX];
};

// Sample columns hi, hi-1, .., lo given the chain's state at column hi+1,
// storing the samples of columns <= whi through dst (dst[c] for column c;
// nothing when dst is null).  The walk stops at the first column whose
// sample equals a compared stored path o (the chains have coalesced: the
// rest is that path) and returns it with opt = o; else it returns -1 with
// the chain at column lo.  The uniform of column col is u[L-1-col].
//
// Rows come through LDS in batches of BT columns, double-buffered: the
// loads of the next batch are issued when a batch starts and written to
// LDS when it ends, so the column loop itself never waits on HBM (vmcnt
// orders loads behind stores, so a per-column prefetch would not hide it).
// The uniforms and the stored paths come 64 columns at a time, one per
// lane, read with readlane.
template <class Real, int R>
__device__ __forceinline__ int tb_range(const DevInst &d, Real *tab, Real *rows, int hi, int lo,
                                        Chain &chn, AW_GLOBAL int *dst, int whi, const Cmp &cmp,
                                        int &opt)
{
    constexpr int RW = 64 * R;
    constexpr int BT = tb_batch<Real, R>();
    const int lane = threadIdx.x, L = d.L;
    AW_GLOBAL const Real *alpha = (AW_GLOBAL const Real *)d.alpha;
    const AW_GLOBAL double *ug = gp(d.u);
    opt = -1;
    if (hi < lo) return -1;
    int knext = chn.k, kinfo = chn.info, kblk = chn.blk;
    int bc = find_owner(cp(d.bstart), d.B, hi);
    int bc_start = cp(d.bstart)[bc];
    int S = cp(d.soff)[bc + 1] - cp(d.soff)[bc];
    TbLane<R> tl;
    tb_load_block<Real, R>(d, bc, tl, tab);
    // window of 64 columns w0, w0-1, .., w0-63: lane j holds column w0-j
    // (its uniform, the stored paths, and the sampled state, which is
    // stored once per window: per-column stores would make every later load
    // wait for them, vmcnt being in order)
    int w0 = hi;
    double uw = ug[min(L - 1 - (w0 - lane), L - 1)];
    // stored paths are read within the walked columns [lo, hi] only
    int pw[TB_KMAX];
#pragma unroll
    for (int o = 0; o < TB_KMAX; o++) pw[o] = cmp.p[o] ? cmp.p[o][max(w0 - lane, lo)] : -1;
    int pnew = 0;
    // store the sampled states of the window's columns w0 .. clo
    auto flush = [&](int clo) {
        const int c = w0 - lane;
        if (dst && c >= clo && c <= min(w0, whi)) dst[c] = pnew;
    };
    Real nxt[BT][R];
#pragma unroll
    for (int k = 0; k < BT; k++) row_load<Real, R>(alpha, max(hi - k, lo), nxt[k]);
#pragma unroll
    for (int k = 0; k < BT; k++)
#pragma unroll
        for (int r = 0; r < R; r++) rows[k * RW + r * 64 + lane] = nxt[k][r];
    int buf = 0;
    for (int top = hi; top >= lo; top -= BT) {
        const int bot = max(top - BT + 1, lo);
#pragma unroll
        for (int k = 0; k < BT; k++) row_load<Real, R>(alpha, max(top - BT - k, lo), nxt[k]);
        const Real *rb = rows + buf * BT * RW;
        for (int col = top; col >= bot; col--) {
            if (w0 - col >= 64) {
                flush(col + 1);
                w0 = col;
                uw = ug[min(L - 1 - (w0 - lane), L - 1)];
#pragma unroll
                for (int o = 0; o < TB_KMAX; o++)
                    if (cmp.p[o]) pw[o] = cmp.p[o][max(w0 - lane, lo)];
            }
            const double uval = readlane_d(uw, w0 - col);
            const Real *row = rb + (top - col) * RW;
            // every load first, then the weights: per-slot conditions around
            // a load would give each load its own wait
            Real rv[R];
#pragma unroll
            for (int r = 0; r < R; r++) rv[r] = row[r * 64 + lane];
            Real tr[R];
            if (bc != kblk) {
                // col is the last column of block bc: switch into kblk = bc+1
                // (TransMatrixSwitch::get, trans.h:201-209)
                const int rs = cp(d.srcs)[2 * kblk], cs = cp(d.srcs)[2 * kblk + 1];
                const int pk = lane_pos(knext, R);
                const AW_GLOBAL Real *cn = planes<Real>(d, kblk, RW);
                const double rrk = cn[CRR * RW + pk], crk = cn[CCR * RW + pk];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int s = lane * R + r;
                    tr[r] = (Real)((s == cs) ? crk : (s == rs) ? rrk : (tl.dt[r] == knext ? tl.dp[r] : 0.0));
                }
            } else {
                const int nk = info_node(kinfo), bk = kinfo & INFO_TIME;
#pragma unroll
                for (int r = 0; r < R; r++) tr[r] = tab[(tl.node[r] == nk ? tl.off2[r] : tl.off1[r]) + bk];
            }
            // padding slots hold 0 in the table, so their weights are 0; the
            // product is in the table's precision, the cumulative sums FP64
            double pre[R];
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < R; r++) {
                acc += (double)(rv[r] * tr[r]);
                pre[r] = acc;
            }
            knext = lane_sample<R>(pre, tl.ti, S, uval, kinfo);
            kblk = bc;
            int hit = -1;
#pragma unroll
            for (int o = TB_KMAX - 1; o >= 0; o--)
                if (cmp.p[o] && __builtin_amdgcn_readlane(pw[o], w0 - col) == knext) hit = o;
            if (hit >= 0) {
                flush(col + 1);
                chn = Chain{knext, kinfo, kblk};
                opt = hit;
                return col;
            }
            pnew = (lane == w0 - col) ? knext : pnew;
            if (col - 1 < bc_start && col - 1 >= lo) {
                bc--;
                bc_start = cp(d.bstart)[bc];
                S = cp(d.soff)[bc + 1] - cp(d.soff)[bc];
                tb_load_block<Real, R>(d, bc, tl, tab);
            }
        }
        buf ^= 1;
        Real *wb = rows + buf * BT * RW;
#pragma unroll
        for (int k = 0; k < BT; k++)
#pragma unroll
            for (int r = 0; r < R; r++) wb[k * RW + r * 64 + lane] = nxt[k][r];
        lds_order();
    }
    flush(lo);
    chn = Chain{knext, kinfo, kblk};
    return -1;
}

// state sampled from column col alone (the last column, or the speculative
// start of a tile)
template <class Real, int R>
__device__ __forceinline__ Chain tb_column(const DevInst &d, int col, double uval)
{
    constexpr int RW = 64 * R;
    AW_GLOBAL const Real *alpha = (AW_GLOBAL const Real *)d.alpha;
    const int b = find_owner(cp(d.bstart), d.B, col);
    const int S = cp(d.soff)[b + 1] - cp(d.soff)[b];
    const AW_GLOBAL int *info = gp(d.linfo) + (size_t)b * RW;
    Real rv[R];
    row_load<Real, R>(alpha, col, rv);
    int ti[R];
#pragma unroll
    for (int r = 0; r < R; r++) ti[r] = info[r * 64 + threadIdx.x];
    double pre[R];
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        acc += (double)rv[r];   // padding slots hold 0
        pre[r] = acc;
    }
    Chain c;
    c.k = lane_sample<R>(pre, ti, S, uval, c.info);
    c.blk = b;
    return c;
}

// The alternative start of option o >= 1 at column col: among the other
// states of the branch of state k0 (block b), the one of rank o - 1 by the
// forward column (ties by state index); -1 when the branch has too few.
// Lane l looks at state k0 - 31 + l: a branch has fewer than 32 states.
template <class Real, int R>
__device__ __forceinline__ int tb_alternative(const DevInst &d, int col, int b, int k0, int o)
{
    constexpr int RW = 64 * R;
    const int lane = threadIdx.x;
    const int S = cp(d.soff)[b + 1] - cp(d.soff)[b];
    const AW_GLOBAL int *info = gp(d.linfo) + (size_t)b * RW;
    AW_GLOBAL const Real *row = (AW_GLOBAL const Real *)d.alpha + (size_t)col * RW;
    const int node = info_node(info[lane_pos(k0, R)]);
    const int x = k0 - 31 + lane;
    const bool ok = x >= 0 && x < S && x != k0 && info_node(info[lane_pos(min(max(x, 0), S - 1), R)]) == node;
    const double v = ok ? (double)row[lane_pos(min(max(x, 0), S - 1), R)] : -1.0;
    // rank among the candidates: those with a larger value, or an equal
    // value and a smaller index
    int rank = 0;
    for (int j = 0; j < 64; j++) {
        const double vj = readlane_d(v, j);
        rank += (vj > v || (vj == v && j < lane)) ? 1 : 0;
    }
    const unsigned long long m = __ballot(ok && rank == o - 1);
    if (!m) return -1;
    return __builtin_amdgcn_readlane(x, (int)__ffsll((long long)m) - 1);
}

// The option buffer column c of option o >= 1 of tile ch (dst[c] / src[c]).
__device__ __forceinline__ AW_GLOBAL int *opt_path(int *obuf, const Chunk &ch, int o)
{
    return (AW_GLOBAL int *)obuf + ch.coff + (long long)(o - 1) * (ch.c1 - ch.c0) - ch.c0;
}

// One option of a tile (the speculative launch).  ostart / oend receive the
// option's state at column c1 (-1: no such option) and at column c0.
template <class Real, int R>
__device__ __forceinline__ void tb_option(const DevInst d, const Chunk &ch, int o, int K, int *obuf,
                                          int *ostart, int *oend, void *lds)
{
    Real *tab = (Real *)lds;
    Real *rows = tab + (d.T - 1) * (d.T - 1) + d.TC;
    constexpr int RW = 64 * R;
    const int L = d.L;
    AW_GLOBAL int *path = gp(d.path);
    const Cmp none{{nullptr, nullptr, nullptr, nullptr}};
    int dummy;
    Chain c;
    if (ch.last) {
        // the last tile is exact: option 0 only
        if (o > 0) {
            if (threadIdx.x == 0) *ostart = *oend = -1;
            return;
        }
        if (d.last_state >= 0) {
            // conditioned on the last column's state
            c.blk = d.B - 1;
            c.k = d.last_state;
            c.info = gp(d.linfo)[(size_t)c.blk * RW + lane_pos(c.k, R)];
        } else {
            c = tb_column<Real, R>(d, L - 1, cp(d.u)[0]);
        }
        if (threadIdx.x == 0) path[L - 1] = c.k;
        tb_range<Real, R>(d, tab, rows, L - 2, ch.c0, c, path, L - 2, none, dummy);
        if (threadIdx.x == 0) {
            *ostart = -1;
            *oend = c.k;
        }
        return;
    }
    // the speculative chain's warm-up, the same in every option
    c = tb_column<Real, R>(d, ch.cw, cp(d.u)[L - 1 - ch.cw]);
    tb_range<Real, R>(d, tab, rows, ch.cw - 1, ch.c1, c, nullptr, -1, none, dummy);
    AW_GLOBAL int *dst = path;
    if (o > 0) {
        const int k = tb_alternative<Real, R>(d, ch.c1, c.blk, c.k, o);
        if (k < 0) {
            if (threadIdx.x == 0) *ostart = *oend = -1;
            return;
        }
        c.k = k;
        c.info = gp(d.linfo)[(size_t)c.blk * RW + lane_pos(k, R)];
        dst = opt_path(obuf, ch, o);
    }
    const int start = c.k;
    tb_range<Real, R>(d, tab, rows, ch.c1 - 1, ch.c0, c, dst, ch.c1 - 1, none, dummy);
    if (threadIdx.x == 0) {
        *ostart = start;
        *oend = c.k;
    }
    (void)K;
}

// Re-run of a tile from its true state at column c1 until it coalesces with
// one of its options' stored paths; the path then follows that option (an
// alternative's columns are copied in).  res: the column of coalescence (-1
// for none), the option, and the state at column c0.
template <class Real, int R>
__device__ __forceinline__ void tb_rerun(const DevInst d, const Chunk &ch, int k1, int K, int *obuf,
                                         const int *ostart, int *res, void *lds)
{
    Real *tab = (Real *)lds;
    Real *rows = tab + (d.T - 1) * (d.T - 1) + d.TC;
    constexpr int RW = 64 * R;
    AW_GLOBAL int *path = gp(d.path);
    Cmp cmp{{path, nullptr, nullptr, nullptr}};
#pragma unroll
    for (int o = 1; o < TB_KMAX; o++)
        if (o < K && ostart[o] >= 0) cmp.p[o] = opt_path(obuf, ch, o);
    Chain c;
    c.blk = find_owner(cp(d.bstart), d.B, ch.c1);
    c.k = k1;
    c.info = gp(d.linfo)[(size_t)c.blk * RW + lane_pos(k1, R)];
    int hit;
    const int col = tb_range<Real, R>(d, tab, rows, ch.c1 - 1, ch.c0, c, path, ch.c1 - 1, cmp, hit);
    int end = c.k;
    const int o = (col >= 0) ? hit : -1;
    if (col >= 0) {
        if (o > 0) {
            // the rest of the tile is the alternative's
            const AW_GLOBAL int *src = opt_path(obuf, ch, o);
            for (int cc = ch.c0 + (int)threadIdx.x; cc <= col; cc += 64) path[cc] = src[cc];
            end = src[ch.c0];
        } else {
            end = path[ch.c0];
        }
    }
    if (threadIdx.x == 0) {
        res[0] = col;
        res[1] = o;
        res[2] = end;
    }
}

// Launch of the speculative options of the tiles list[0..nlist): block
// (jj, o) of tile list[jj], option o sits at 8 * (K * (jj / 8) + o) + jj % 8,
// so the options of a tile run on one XCD (blocks go round-robin over the
// eight) and share its L2 for their rows.
// BIG: the tiles of instances with 32 states per lane (a symbol of their own,
// as for the forward)
template <class Real, bool BIG>
__global__ __launch_bounds__(64) void k_tb_spec(const DevInst *__restrict__ D,
                                                const Chunk *__restrict__ chunks,
                                                const int *__restrict__ list, int nlist, int K,
                                                int *__restrict__ obuf, int *__restrict__ ostart,
                                                int *__restrict__ oend)
{
    extern __shared__ double tb_dyn[];
    const int bx = blockIdx.x;
    const int jj = (bx / (8 * K)) * 8 + bx % 8, o = (bx / 8) % K;
    if (jj >= nlist) return;
    const int j = list[jj];
    const Chunk ch = chunks[j];
    const DevInst &d = D[ch.inst];
    int *os = ostart + (size_t)j * K + o, *oe = oend + (size_t)j * K + o;
    if ((d.R > 16) != BIG) return;   // the other symbol's tile
    if constexpr (BIG) {
        tb_option<Real, 32>(d, ch, o, K, obuf, os, oe, tb_dyn);
    } else {
        switch (d.R) {
        case 1: tb_option<Real, 1>(d, ch, o, K, obuf, os, oe, tb_dyn); break;
        case 2: tb_option<Real, 2>(d, ch, o, K, obuf, os, oe, tb_dyn); break;
        case 4: tb_option<Real, 4>(d, ch, o, K, obuf, os, oe, tb_dyn); break;
        case 6: tb_option<Real, 6>(d, ch, o, K, obuf, os, oe, tb_dyn); break;
        case 8: tb_option<Real, 8>(d, ch, o, K, obuf, os, oe, tb_dyn); break;
        case 12: tb_option<Real, 12>(d, ch, o, K, obuf, os, oe, tb_dyn); break;
        default: tb_option<Real, 16>(d, ch, o, K, obuf, os, oe, tb_dyn); break;
        }
    }
}

// Re-runs: block x re-samples tile idx[x] from state k1[x] at its column c1.
template <class Real, bool BIG>
__global__ __launch_bounds__(64) void k_tb_rerun(const DevInst *__restrict__ D,
                                                 const Chunk *__restrict__ chunks,
                                                 const int *__restrict__ idx,
                                                 const int *__restrict__ k1, int K,
                                                 int *__restrict__ obuf,
                                                 const int *__restrict__ ostart,
                                                 int *__restrict__ res)
{
    extern __shared__ double tb_dyn[];
    const int j = idx[blockIdx.x];
    const Chunk ch = chunks[j];
    const DevInst &d = D[ch.inst];
    const int *os = ostart + (size_t)j * K;
    int *rs = res + 3 * blockIdx.x;
    const int k = k1[blockIdx.x];
    if ((d.R > 16) != BIG) return;   // the other symbol's tile
    if constexpr (BIG) {
        tb_rerun<Real, 32>(d, ch, k, K, obuf, os, rs, tb_dyn);
    } else {
        switch (d.R) {
        case 1: tb_rerun<Real, 1>(d, ch, k, K, obuf, os, rs, tb_dyn); break;
        case 2: tb_rerun<Real, 2>(d, ch, k, K, obuf, os, rs, tb_dyn); break;
        case 4: tb_rerun<Real, 4>(d, ch, k, K, obuf, os, rs, tb_dyn); break;
        case 6: tb_rerun<Real, 6>(d, ch, k, K, obuf, os, rs, tb_dyn); break;
        case 8: tb_rerun<Real, 8>(d, ch, k, K, obuf, os, rs, tb_dyn); break;
        case 12: tb_rerun<Real, 12>(d, ch, k, K, obuf, os, rs, tb_dyn); break;
        default: tb_rerun<Real, 16>(d, ch, k, K, obuf, os, rs, tb_dyn); break;
        }
    }
}

// Adopted alternatives: block x copies option opt[x] of tile idx[x] into the
// path over the tile's columns.
__global__ __launch_bounds__(256) void k_tb_adopt(const DevInst *__restrict__ D,
                                                  const Chunk *__restrict__ chunks,
                                                  const int *__restrict__ idx,
                                                  const int *__restrict__ opt,
                                                  int *__restrict__ obuf)
{
    const Chunk ch = chunks[idx[blockIdx.x]];
    AW_GLOBAL int *path = gp(D[ch.inst].path);
    const AW_GLOBAL int *src = opt_path(obuf, ch, opt[blockIdx.x]);
    for (int c = ch.c0 + (int)threadIdx.x; c < ch.c1; c += 256) path[c] = src[c];
}

//===========================================================================
// host driver

namespace {

// supported states-per-lane counts (one kernel instantiation each)
constexpr int NRV = 8;
constexpr int RVALS[NRV] = {1, 2, 4, 6, 8, 12, 16, 32};

// Device-wide state, initialised once per process.
struct Ctx {
    bool init = false;
    int dev = 0;
    hipDeviceProp_t prop{};
};

Ctx &ctx()
{
    static Ctx c;
    return c;
}

int ctx_init()
{
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    Ctx &c = ctx();
    if (c.init) return 0;
    HIP_OK(hipGetDevice(&c.dev));
    HIP_OK(hipGetDeviceProperties(&c.prop, c.dev));
    c.init = true;
    return 0;
}

// The upload, run, tuning and copy-out methods of a Batch hold a lock for
// their whole device work.  By default it is one process-wide lock, so
// batches driven from different host threads run one after another; a
// batch set concurrent (the server's) holds a lock of its own and runs
// beside the others.  Concurrent batches are safe because no stream ever
// waits on another stream's event (streams are joined by the host), so no
// hardware queue can hold a wait ahead of the work it waits for, and the
// kernels' dynamic-LDS limits only rise (set_lds_limit).  Recursive because
// run() uploads a batch that is not yet uploaded.
std::recursive_mutex &device_mutex()
{
    static std::recursive_mutex mu;
    return mu;
}
#define DEVICE_LOCK std::lock_guard<std::recursive_mutex> device_lock_(device_mutex())

// The batch's streams: main carries every stage; side carries only the
// traceback of the instances whose forward is final (their options,
// re-runs and adoptions, driven by a helper host thread) while the forward
// re-runs of others continue on main.  The two are joined by the host
// (stream synchronisation), never by a stream waiting on an event.
struct Streams {
    bool init = false;
    hipStream_t main = nullptr, side = nullptr;
    hipEvent_t ev[10];
    int create()
    {
        if (init) return 0;
        HIP_OK(hipStreamCreateWithFlags(&main, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        for (auto &e : ev) HIP_OK(hipEventCreate(&e));
        init = true;
        return 0;
    }
    void destroy()
    {
        if (!init) return;
        (void)hipStreamSynchronize(main);
        (void)hipStreamSynchronize(side);
        for (auto &e : ev) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(main);
        (void)hipStreamDestroy(side);
        init = false;
    }
};

// A caller-provided device region that buffers are carved from instead of
// hipMalloc (the no-allocation entry point); reset() before each use.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void reset() { used = 0; }
};

// grow-only device buffer; with an arena it takes its memory from the arena
// (and never frees it)
struct DevBuf {
    char *p = nullptr;
    size_t cap = 0;
    Arena *arena = nullptr;
    int ensure(size_t n)
    {
        if (n <= cap) return 0;
        if (arena) {
            const size_t off = (arena->used + 255) & ~size_t(255);
            if (off + n > arena->cap) {
                fprintf(stderr,
                        "argweaver_amd: the instance needs more than the %zu bytes of the "
                        "no-allocation workspace\n",
                        arena->cap);
                return 1;
            }
            p = arena->base + off;
            cap = n;
            arena->used = off + n;
            return 0;
        }
        if (p) HIP_OK(hipFree(p));
        p = nullptr;
        cap = 0;
        HIP_OK(hipMalloc(&p, n));
        cap = n;
        return 0;
    }
    void release()
    {
        if (p && !arena) (void)hipFree(p);
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

// smallest supported states-per-lane count covering Smax states
int lanes_per_state_r(int Smax)
{
    for (int R : RVALS)
        if (64 * R >= Smax) return R;
    return -1;
}

// LDS bytes / size of Real of the forward kernel for R states per lane
int fw_lds_reals_r(bool fp32, int R)
{
    switch (R) {
    case 1: return fp32 ? fw_lds_reals<float, 1>() : fw_lds_reals<double, 1>();
    case 2: return fp32 ? fw_lds_reals<float, 2>() : fw_lds_reals<double, 2>();
    case 4: return fp32 ? fw_lds_reals<float, 4>() : fw_lds_reals<double, 4>();
    case 6: return fp32 ? fw_lds_reals<float, 6>() : fw_lds_reals<double, 6>();
    case 8: return fp32 ? fw_lds_reals<float, 8>() : fw_lds_reals<double, 8>();
    case 12: return fp32 ? fw_lds_reals<float, 12>() : fw_lds_reals<double, 12>();
    case 16: return fp32 ? fw_lds_reals<float, 16>() : fw_lds_reals<double, 16>();
    default: return fp32 ? fw_lds_reals<float, 32>() : fw_lds_reals<double, 32>();
    }
}

// The dynamic-LDS limit of a kernel only ever rises, under one lock: a
// batch on another host thread that set a larger limit for the same kernel
// is never undercut between its attribute call and its launch.
int set_lds_limit(const void *kern, size_t sh)
{
    static std::mutex mu;
    static std::map<const void *, size_t> cur;
    std::lock_guard<std::mutex> lk(mu);
    size_t &c = cur[kern];
    if (sh <= c) return 0;
    HIP_OK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    c = sh;
    return 0;
}

template <class Real, bool BIG = false>
int launch_fw(const DevInst *dD, const FwTile *tiles, int cnt, void *bnd, double *bnorm,
              size_t sh, hipStream_t st, int T1max, double tol = 0.0, int *early = nullptr)
{
    auto kern = early ? ((T1max <= 19) ? k_fw_all<Real, 19, true, BIG> : k_fw_all<Real, MAXT - 1, true, BIG>)
                      : ((T1max <= 19) ? k_fw_all<Real, 19, false, BIG> : k_fw_all<Real, MAXT - 1, false, BIG>);
    if (set_lds_limit((const void *)kern, sh)) return 1;
    hipLaunchKernelGGL(kern, dim3(cnt), dim3(64), sh, st, dD, tiles, (Real *)bnd, bnorm, tol, early);
    HIP_OK(hipGetLastError());
    return 0;
}

// launch of a traceback kernel with sh bytes of dynamic LDS
template <class Kern, class... Args>
int launch_lds(Kern kern, int grid, size_t sh, hipStream_t st, Args... args)
{
    if (set_lds_limit((const void *)kern, sh)) return 1;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), sh, st, args...);
    HIP_OK(hipGetLastError());
    return 0;
}

}  // namespace

struct Batch::Impl {
    std::vector<HostInstance> hosts;
    std::vector<std::vector<double>> us, priors;
    std::vector<std::vector<int>> Sv;
    std::vector<std::vector<int>> soff;
    std::vector<DevInst> h;
    std::vector<long long> W;
    std::vector<double> lnl;
    std::vector<int> nvar;
    std::vector<Chunk> chunks;
    std::vector<int> tlist0, tlist1;   // host tile lists of the option launches
    DevBuf in, work, tab, aux, vbuf, tiles, fwb;
    Arena arena;   // base == nullptr: buffers come from hipMalloc
    std::vector<std::vector<float>> staged32;   // FP32 planes of host-prepared instances
    Streams st;
    DevInst *dD = nullptr;
    long long *d_coloff = nullptr;
    int *d_blkoff = nullptr, *d_voff = nullptr, *d_nvar = nullptr, *d_list = nullptr;
    // traceback options: start / end states [nch * K], alternative paths,
    // re-run lists and results, adopted alternatives
    int *d_ostart = nullptr, *d_oend = nullptr, *d_obuf = nullptr;
    int *d_ridx = nullptr, *d_rk1 = nullptr, *d_rres = nullptr, *d_aidx = nullptr, *d_aopt = nullptr;
    int *d_tlist = nullptr;   // [2 nch] tile lists of the two option launches
    int tb_opts = 1;
    // instances with 32 states per lane run in kernel symbols of their own;
    // the LDS bytes of those symbols' forward and traceback launches
    bool any_big = false;
    size_t shf_big = 0, sht_big = 0;
    Chunk *d_chunks = nullptr;
    FwTile *d_fwt = nullptr, *d_fwre = nullptr;
    char *d_bnd = nullptr;
    double *d_bnorm = nullptr;
    int *d_fidx = nullptr, *d_fflag = nullptr;
    bool uploaded = false;
    Precision prec_done = Precision::FP64;
    bool own_lock = false;     // set_concurrent: this batch's own lock
    std::recursive_mutex mu;

    ~Impl()
    {
        st.destroy();
        in.release();
        work.release();
        tab.release();
        aux.release();
        vbuf.release();
        tiles.release();
        fwb.release();
    }
};

// the lock of a batch's device work (see device_mutex)
#define BATCH_LOCK \
    std::lock_guard<std::recursive_mutex> batch_lock_(impl_->own_lock ? impl_->mu : device_mutex())

Batch::Batch() : impl_(new Impl) {}
void Batch::set_concurrent(bool on) { impl_->own_lock = on; }
Batch::~Batch() { delete impl_; }
int Batch::size() const { return (int)impl_->hosts.size(); }

void Batch::use_workspace(void *base, size_t cap)
{
    Impl &m = *impl_;
    m.arena.base = static_cast<char *>(base);
    m.arena.cap = cap;
    m.arena.reset();
    for (DevBuf *b : {&m.in, &m.work, &m.tab, &m.aux, &m.vbuf, &m.tiles, &m.fwb}) {
        b->release();
        b->arena = &m.arena;
    }
}

int Batch::init_streams()
{
    if (ctx_init()) return 1;
    return impl_->st.create();
}

void Batch::clear()
{
    Impl &m = *impl_;
    if (m.arena.base) {
        // a workspace batch carves its buffers again from the start
        m.arena.reset();
        for (DevBuf *b : {&m.in, &m.work, &m.tab, &m.aux, &m.vbuf, &m.tiles, &m.fwb}) {
            b->p = nullptr;
            b->cap = 0;
        }
    }
    m.hosts.clear();
    m.us.clear();
    m.priors.clear();
    m.Sv.clear();
    m.soff.clear();
    m.h.clear();
    m.W.clear();
    m.lnl.clear();
    m.nvar.clear();
    m.chunks.clear();
    m.uploaded = false;
}
double Batch::lnl(int i) const { return impl_->lnl[i]; }
long long Batch::W(int i) const { return impl_->W[i]; }

int Batch::add(const HostInstance &hi, const double *u, const double *prior)
{
    return add(HostInstance(hi), u, prior);
}

int Batch::add(HostInstance &&hin, const double *u, const double *prior)
{
    Impl &m = *impl_;
    // upload() releases the host rows of the instances it copied, so a batch
    // that was uploaded takes no more instances until clear()
    if (m.uploaded) {
        fprintf(stderr, "argweaver_amd: Batch::add after upload (clear() the batch first)\n");
        return -1;
    }
    std::string why;
    if (instance_supported(hin, &why)) {
        fprintf(stderr, "argweaver_amd: Batch::add refused the instance: %s\n", why.c_str());
        return -1;
    }
    m.hosts.push_back(std::move(hin));
    const HostInstance &hi = m.hosts.back();
    m.us.emplace_back();
    if (u) m.us.back().assign(u, u + hi.ft.seqlen);
    m.priors.emplace_back();
    std::vector<int> Sv;
    long long W = 0;
    if (hi.prep) {
        Sv = hi.prep->S;
        for (int b = 0; b < hi.ft.nblocks; b++) W += (long long)Sv[b] * hi.ft.len[b];
    } else {
        W = instance_states(hi.ft, hi.model.T, &Sv);
    }
    if (prior) m.priors.back().assign(prior, prior + Sv[0]);
    m.W.push_back(W);
    m.Sv.push_back(Sv);
    m.uploaded = false;
    return (int)m.hosts.size() - 1;
}

// Inputs: trees, sequences, uniforms and the model of every instance.
int Batch::upload()
{
    BATCH_LOCK;
    if (ctx_init()) return 1;
    Impl &m = *impl_;
    if (m.st.create()) return 1;
    Streams &c = m.st;
    const int n = (int)m.hosts.size();
    m.h.assign(n, DevInst{});
    m.soff.assign(n, {});
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
                so.assign(B + 1, 0);
                int Smax = 1;
                for (int b = 0; b < B; b++) {
                    so[b + 1] = so[b] + m.Sv[i][b];
                    Smax = std::max(Smax, m.Sv[i][b]);
                }
                const int R = lanes_per_state_r(Smax);
                if (R < 0) {
                    fprintf(stderr, "argweaver_amd: %d states per block exceed the %d of one wave\n",
                            Smax, 64 * RVALS[NRV - 1]);
                    return 1;
                }
                d.hostprep = hi.prep ? 1 : 0;
                d.last_state = hi.last_state;
                d.T = T; d.N = N; d.B = B; d.L = L; d.nrows = (int)hi.rows.size();
                d.Smax = Smax; d.TC = tc_base(T - 1, T - 1); d.W = m.W[i]; d.R = R;
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
            d.seq = p.take<char>((size_t)d.nrows * L);
            d.u = p.take<double>(L);
            if (pass == 1) {
                hipStream_t s = c.main;
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
                for (int r = 0; r < d.nrows; r++)
                    HIP_OK(hipMemcpyAsync((char *)d.seq + (size_t)r * L, hi.rows[r], L, hipMemcpyHostToDevice, s));
                if (!m.us[i].empty())
                    HIP_OK(hipMemcpyAsync((void *)d.u, m.us[i].data(), sizeof(double) * L, hipMemcpyHostToDevice, s));
            }
        }
        if (pass == 0 && m.in.ensure(p.used + 4096)) return 1;
    }
    HIP_OK(hipStreamSynchronize(c.main));
    // host rows may be released now
    for (auto &hi : m.hosts) hi.rows.clear();
    m.uploaded = true;
    return 0;
}

// FP32 traceback column cost (us) by states per lane, measured on one
// instance per R (20 time points) on MI355X.
inline double tb_cost_model(int R)
{
    return R <= 1 ? 0.50 : R == 2 ? 0.52 : R == 4 ? 0.63 : R == 6 ? 0.75 : R == 8 ? 0.86 : R == 12 ? 1.05 : 1.25;
}

// Traceback tiles of every instance: about tile_cols columns (0: one per
// instance; scaled by the column cost of the instance's R), tail_cols for
// the instances with tail[inst] set, the instances with the most states per
// lane first, K options per tile.  Copies the tile list to the device on the
// main stream; returns the LDS size of the tile kernels (sht).
static int plan_tiles(Batch::Impl &m, bool fp32, int tile_cols, int warm, int K, size_t &sht,
                      const std::vector<char> *tail = nullptr, int tail_cols = 0)
{
    const int n = (int)m.h.size();
    const size_t rsz = fp32 ? sizeof(float) : sizeof(double);
    std::vector<Chunk> &chs = m.chunks;
    chs.clear();
    m.sht_big = 0;
    // The tail instances' tiles run when the rest of the device is idle:
    // tail_cols long, or longer when that would give more option waves
    // than the device has SIMDs (each wave a SIMD to itself), since a
    // second round of waves costs more than longer tiles do
    if (tail && tail_cols > 0) {
        double cols = 0;
        for (int i = 0; i < n; i++)
            if ((*tail)[i]) cols += (double)m.h[i].L * tb_cost_model(m.h[i].R) / tb_cost_model(12);
        const double simds = 4.0 * std::max(1, ctx().prop.multiProcessorCount);
        tail_cols = std::max(tail_cols, (int)std::ceil(cols * K / simds));
    }
    // tiles of the instances with the most states per lane (the slowest
    // columns) are dispatched first
    std::vector<int> byr(n);
    std::iota(byr.begin(), byr.end(), 0);
    std::stable_sort(byr.begin(), byr.end(), [&](int a, int b) { return m.h[a].R > m.h[b].R; });
    sht = 0;
    long long ocols = 0;   // columns of the tiles that have alternatives
    for (int i : byr) {
        const DevInst &d = m.h[i];
        // tile_cols columns at R = 12's column cost; instances with cheaper
        // columns get proportionally longer tiles (fewer tiles, the same
        // time per tile)
        const int tc = (tail && (*tail)[i] && tail_cols > 0) ? tail_cols : tile_cols;
        const int tile =
            tc > 0 ? std::max(64, (int)(tc * tb_cost_model(12) / tb_cost_model(d.R))) : d.L;
        const int nt = (d.L + tile - 1) / tile;
        for (int t = 0; t < nt; t++) {
            Chunk k;
            k.inst = i;
            k.c0 = t * tile;
            k.c1 = std::min(d.L, (t + 1) * tile);
            k.first = (t == 0);
            k.last = (t == nt - 1);
            k.cw = k.last ? d.L - 1 : std::min(d.L - 1, k.c1 + std::max(warm, 0));
            k.coff = k.last ? 0 : (long long)(K - 1) * ocols;
            if (!k.last) ocols += k.c1 - k.c0;
            chs.push_back(k);
        }
        // (a,b) and (c,a,b) tables, then two row batches
        const int T1 = d.T - 1;
        const int bt = tb_batch_host(d.R, rsz);
        size_t &sh = d.R > 16 ? m.sht_big : sht;
        sh = std::max(sh, rsz * ((size_t)T1 * T1 + d.TC) + rsz * 2 * bt * 64 * (size_t)d.R);
    }
    const int nch = (int)chs.size();
    m.tb_opts = K;
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.tiles.p : nullptr;
        m.d_chunks = p.take<Chunk>(nch);
        m.d_ostart = p.take<int>((size_t)nch * K);
        m.d_oend = p.take<int>((size_t)nch * K);
        m.d_ridx = p.take<int>(2 * (size_t)nch);
        m.d_rk1 = p.take<int>(2 * (size_t)nch);
        m.d_rres = p.take<int>(6 * (size_t)nch);
        m.d_aidx = p.take<int>(2 * (size_t)nch);
        m.d_aopt = p.take<int>(2 * (size_t)nch);
        m.d_tlist = p.take<int>(2 * (size_t)nch);
        m.d_obuf = p.take<int>((size_t)std::max(1LL, (long long)(K - 1) * ocols));
        if (pass == 0 && m.tiles.ensure(p.used + 4096)) return 1;
    }
    HIP_OK(hipMemcpyAsync(m.d_chunks, chs.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice,
                          m.st.main));
    return 0;
}

// The speculative launch on stream s: every option of the tiles of the
// instances with take[inst] set (all when take is null); slot 0 or 1 of the
// device tile lists.
static int launch_options(Batch::Impl &m, bool fp32, size_t sht, const std::vector<char> *take,
                          int slot, hipStream_t s)
{
    const int nch = (int)m.chunks.size(), K = m.tb_opts;
    std::vector<int> &list = slot ? m.tlist1 : m.tlist0;
    list.clear();
    // the tiles of the 32-states-per-lane instances after the others (a
    // launch of their own symbol)
    for (int big = 0; big < 2; big++)
        for (int j = 0; j < nch; j++)
            if ((!take || (*take)[m.chunks[j].inst]) && (m.h[m.chunks[j].inst].R > 16) == (big == 1))
                list.push_back(j);
    const int nl = (int)list.size();
    if (!nl) return 0;
    int nn = 0;
    while (nn < nl && m.h[m.chunks[list[nn]].inst].R <= 16) nn++;
    int *dl = m.d_tlist + (size_t)slot * nch;
    HIP_OK(hipMemcpyAsync(dl, list.data(), sizeof(int) * nl, hipMemcpyHostToDevice, s));
    if (nn > 0) {
        const int grid = (nn + 7) / 8 * 8 * K;
        if (fp32 ? launch_lds(k_tb_spec<float, false>, grid, sht, s, (const DevInst *)m.dD,
                              (const Chunk *)m.d_chunks, (const int *)dl, nn, K, m.d_obuf,
                              m.d_ostart, m.d_oend)
                 : launch_lds(k_tb_spec<double, false>, grid, sht, s, (const DevInst *)m.dD,
                              (const Chunk *)m.d_chunks, (const int *)dl, nn, K, m.d_obuf,
                              m.d_ostart, m.d_oend))
            return 1;
    }
    if (nl > nn) {
        const int nb = nl - nn, grid = (nb + 7) / 8 * 8 * K;
        if (fp32 ? launch_lds(k_tb_spec<float, true>, grid, m.sht_big, s, (const DevInst *)m.dD,
                              (const Chunk *)m.d_chunks, (const int *)(dl + nn), nb, K, m.d_obuf,
                              m.d_ostart, m.d_oend)
                 : launch_lds(k_tb_spec<double, true>, grid, m.sht_big, s, (const DevInst *)m.dD,
                              (const Chunk *)m.d_chunks, (const int *)(dl + nn), nb, K, m.d_obuf,
                              m.d_ostart, m.d_oend))
            return 1;
    }
    return 0;
}

// Resolution of the tiles, right to left within each instance.  A tile
// whose true state at c1 is the start of one of its options takes that
// option's path (option 0 is in place; an alternative is copied in at the
// end) and passes that option's end state on as the true state of the
// tile before.  The first tile that no option fits is re-run from its true
// state, and in the same pass every tile before it whose predicted start
// (the end of the next tile's option 0, or of the option its own
// prediction fits) fits none of its options is re-run from that
// prediction.  A re-run replaces option 0 by the chain from its start (the
// samples until it coalesces with a stored option, then that option), so a
// right prediction leaves the tile decided for the next pass.
// take: the instances to resolve (all when null); slot 0 or 1 of the
// device lists, so that two resolutions of disjoint instances can run at
// once on two streams.
static int resolve_tiles(Batch::Impl &m, bool fp32, size_t sht, RunStats &rs, hipStream_t s,
                         const std::vector<char> *take, int slot)
{
    const std::vector<Chunk> &chs = m.chunks;
    const int nch = (int)chs.size(), K = m.tb_opts;
    int *ridx = m.d_ridx + (size_t)slot * nch, *rk1 = m.d_rk1 + (size_t)slot * nch;
    int *rres = m.d_rres + (size_t)slot * 3 * nch;
    int *aidx = m.d_aidx + (size_t)slot * nch, *aopt = m.d_aopt + (size_t)slot * nch;
    std::vector<int> os((size_t)nch * K), oe((size_t)nch * K);
    HIP_OK(hipMemcpyAsync(os.data(), m.d_ostart, sizeof(int) * os.size(), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(oe.data(), m.d_oend, sizeof(int) * oe.size(), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    auto match = [&](int j, int b) {
        for (int o = 0; o < K; o++)
            if (os[(size_t)j * K + o] == b) return o;
        return -1;
    };
    // per instance: its first tile, the rightmost undecided tile and the
    // true state at that tile's column c1
    struct Walk {
        int f, r, b;
    };
    std::vector<Walk> ws;
    for (int j = 0; j < nch; j++)
        if (chs[j].last && !chs[j].first && (!take || (*take)[chs[j].inst])) {
            int f = j;
            while (!chs[f].first) f--;
            ws.push_back(Walk{f, j - 1, oe[(size_t)j * K]});
        }
    std::vector<int> aj, ao, pj, pk, res;
    while (true) {
        pj.clear();
        pk.clear();
        for (Walk &w : ws) {
            while (w.r >= w.f) {
                const int o = match(w.r, w.b);
                if (o < 0) break;
                if (o > 0) {
                    aj.push_back(w.r);
                    ao.push_back(o);
                }
                w.b = oe[(size_t)w.r * K + o];
                w.r--;
            }
            if (w.r < w.f) continue;
            pj.push_back(w.r);
            pk.push_back(w.b);
            int g = oe[(size_t)w.r * K];
            for (int j = w.r - 1; j >= w.f; j--) {
                const int o = match(j, g);
                if (o >= 0) {
                    g = oe[(size_t)j * K + o];
                    continue;
                }
                pj.push_back(j);
                pk.push_back(g);
                g = oe[(size_t)j * K];
            }
        }
        if (pj.empty()) break;
        const int np = (int)pj.size();
        HIP_OK(hipMemcpyAsync(ridx, pj.data(), sizeof(int) * np, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(rk1, pk.data(), sizeof(int) * np, hipMemcpyHostToDevice, s));
        if (fp32 ? launch_lds(k_tb_rerun<float, false>, np, sht, s, (const DevInst *)m.dD,
                              (const Chunk *)m.d_chunks, (const int *)ridx,
                              (const int *)rk1, K, m.d_obuf, (const int *)m.d_ostart, rres)
                 : launch_lds(k_tb_rerun<double, false>, np, sht, s, (const DevInst *)m.dD,
                              (const Chunk *)m.d_chunks, (const int *)ridx,
                              (const int *)rk1, K, m.d_obuf, (const int *)m.d_ostart, rres))
            return 1;
        if (m.any_big &&
            (fp32 ? launch_lds(k_tb_rerun<float, true>, np, m.sht_big, s, (const DevInst *)m.dD,
                               (const Chunk *)m.d_chunks, (const int *)ridx,
                               (const int *)rk1, K, m.d_obuf, (const int *)m.d_ostart, rres)
                  : launch_lds(k_tb_rerun<double, true>, np, m.sht_big, s, (const DevInst *)m.dD,
                               (const Chunk *)m.d_chunks, (const int *)ridx,
                               (const int *)rk1, K, m.d_obuf, (const int *)m.d_ostart, rres)))
            return 1;
        res.resize(3 * (size_t)np);
        HIP_OK(hipMemcpyAsync(res.data(), rres, sizeof(int) * res.size(), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        rs.tb_passes++;
        for (int x = 0; x < np; x++) {
            const int j = pj[x];
            rs.tb_rerun_cols += chs[j].c1 - (res[3 * x] < 0 ? chs[j].c0 : res[3 * x]);
            os[(size_t)j * K] = pk[x];
            oe[(size_t)j * K] = res[3 * x + 2];
        }
    }
    if (!aj.empty()) {
        const int na = (int)aj.size();
        HIP_OK(hipMemcpyAsync(aidx, aj.data(), sizeof(int) * na, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(aopt, ao.data(), sizeof(int) * na, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_tb_adopt, dim3(na), dim3(256), 0, s, (const DevInst *)m.dD,
                           (const Chunk *)m.d_chunks, (const int *)aidx, (const int *)aopt,
                           m.d_obuf);
        HIP_OK(hipGetLastError());
    }
    rs.tb_adopted = (int)aj.size();
    return 0;
}

// The traceback of every instance whose forward table is resident: the
// options of every tile in one launch, then the resolution (re-runs where
// no option fits).  resolve = false leaves the speculative option-0 paths
// (diagnostics).
static int traceback_stage(Batch::Impl &m, bool fp32, const BatchOptions &opt, RunStats &rs)
{
    size_t sht;
    const int K = std::max(1, std::min(opt.tile_opts, TB_KMAX));
    if (plan_tiles(m, fp32, opt.tile_cols, opt.tile_warm, K, sht)) return 1;
    if (launch_options(m, fp32, sht, nullptr, 0, m.st.main)) return 1;
    return opt.fixup ? resolve_tiles(m, fp32, sht, rs, m.st.main, nullptr, 0) : 0;
}

int Batch::tune_tiles(Precision prec, const int *cand, int ncand, double *ms_out, int warm,
                      int opts)
{
    BATCH_LOCK;
    BatchOptions opt;
    opt.prec = prec;
    opt.traceback = false;
    if (run(opt)) return -1;
    Impl &m = *impl_;
    Streams &c = m.st;
    const bool fp32 = prec == Precision::FP32;
    int best = cand[0];
    float best_ms = 0;
    for (int i = 0; i < ncand; i++) {
        RunStats rs;
        (void)hipEventRecord(c.ev[6], c.main);
        BatchOptions topt;
        topt.prec = prec;
        topt.tile_cols = cand[i];
        topt.tile_warm = warm;
        topt.tile_opts = opts;
        if (traceback_stage(m, fp32, topt, rs)) return -1;
        (void)hipEventRecord(c.ev[7], c.main);
        HIP_OK(hipStreamSynchronize(c.main));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c.ev[6], c.ev[7]);
        if (ms_out) ms_out[i] = ms;
        if (i == 0 || ms < best_ms) {
            best_ms = ms;
            best = cand[i];
        }
    }
    return best;
}

int Batch::tune_forward(const int *tiles, const int *warms, int ncand, double tol, double *ms_out)
{
    BATCH_LOCK;
    int best = -1;
    double best_ms = 0;
    for (int i = 0; i < ncand; i++) {
        BatchOptions opt;
        opt.prec = Precision::FP32;
        opt.traceback = false;
        opt.fw_tile = tiles[i];
        opt.fw_warm = warms[i] > 0 ? warms[i] : opt.fw_warm;
        if (tol > 0) opt.fw_tol = tol;
        RunStats rs;
        if (run(opt, &rs)) return -1;
        if (ms_out) ms_out[i] = rs.ms_forward;
        if (best < 0 || rs.ms_forward < best_ms) {
            best_ms = rs.ms_forward;
            best = i;
        }
    }
    return best;
}

// FP32 forward column cost (us) by states per lane: medians of the tiles of
// the 256-instance batch (20 time points) on MI355X.  Small R pays for a
// longer carry chain of the same-branch scan, so the cost is nearly flat up
// to R = 6.
constexpr double FW_CREF = 0.94;
inline double fw_cost_model(int R)
{
    return R <= 2 ? 0.94 : R == 4 ? 0.90 : R == 6 ? 0.98 : R == 8 ? 1.09 : R == 12 ? 1.34 : 1.60;
}

// The forward of every instance.  FP32 runs may split long instances into
// speculative tiles (BatchOptions::fw_tile); each tile's start is then
// checked against the boundary column of the tile before, and tiles that
// fail are re-run from that column until every check passes (a re-run
// changes the boundary column of the next tile, which is checked again).
// AWAMD_TRACE_TILES=1 in the environment: the forward re-runs are printed
// to stderr (diagnostics)
static bool trace_tiles()
{
    static const bool on = [] {
        const char *e = getenv("AWAMD_TRACE_TILES");
        return e && e[0] == '1';
    }();
    return on;
}

static int forward_stage(Batch::Impl &m, bool fp32, const BatchOptions &opt,
                         const std::vector<int> &order, size_t shf, int t1max, RunStats &rs,
                         const std::function<int(const std::vector<char> &)> &first_check)
{
    Streams &c = m.st;
    hipStream_t s = c.main;
    const size_t rsz = fp32 ? sizeof(float) : sizeof(double);
    std::vector<FwTile> tl;
    std::vector<int> next;   // tile index of the following tile of its instance, -1 for none
    int nslot = 0;
    const int C0 = (fp32 && opt.fw_tile > 0) ? opt.fw_tile : 0;
    const int W = std::max(opt.fw_warm, 1);
    for (int i : order) {
        const int L = m.h[i].L;
        // tiles of C0 columns at a column cost of FW_CREF us; an instance
        // whose columns cost more gets proportionally shorter tiles, so that
        // the tiles of every instance take about the same time
        const int C = C0 ? std::max(W + W / 4, (int)(C0 * FW_CREF / fw_cost_model(m.h[i].R))) : 0;
        if (!C || L <= C || C <= W) {
            tl.push_back(FwTile{i, 0, 0, L, 0, -1, -1});
            next.push_back(-1);
            continue;
        }
        // tile 0: [0, C); then tiles of C - W stored columns after a warm-up
        // of W columns
        const int F = C - W;
        int s0 = 0, e0 = C, prev = -1;
        while (true) {
            const bool last = e0 >= L;
            const int slot = last ? -1 : nslot++;
            FwTile t;
            t.inst = i;
            t.cs = s0;
            t.c1 = std::min(e0, L);
            t.c0 = s0 == 0 ? 0 : s0 - W;
            t.mode = s0 == 0 ? 0 : 1;
            t.bin = prev;
            t.bout = slot;
            if (!next.empty() && prev >= 0) next.back() = (int)tl.size();
            tl.push_back(t);
            next.push_back(-1);
            if (last) break;
            prev = slot;
            s0 = e0;
            e0 = s0 + F;
        }
    }
    const int nt = (int)tl.size();
    rs.fw_tiles = nt;
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.fwb.p : nullptr;
        m.d_fwt = p.take<FwTile>(nt);
        m.d_fwre = p.take<FwTile>(nt);
        m.d_bnd = p.take<char>((size_t)std::max(nslot, 1) * BND_STRIDE * rsz);
        m.d_bnorm = p.take<double>(std::max(nslot, 1));
        m.d_fidx = p.take<int>(nt);
        m.d_fflag = p.take<int>(nt);
        if (pass == 0 && m.fwb.ensure(p.used + 4096)) return 1;
    }
    HIP_OK(hipMemcpyAsync(m.d_fwt, tl.data(), sizeof(FwTile) * nt, hipMemcpyHostToDevice, s));
    const auto t_fw = std::chrono::steady_clock::now();
    // the 32-states-per-lane tiles run on the side stream, beside the
    // others: the host joins the main stream first (preparation, emissions
    // and the tile list done), and the side stream before the checks
    if (m.any_big) {
        HIP_OK(hipStreamSynchronize(s));
        if (fp32 ? launch_fw<float, true>(m.dD, m.d_fwt, nt, m.d_bnd, m.d_bnorm, m.shf_big, c.side,
                                          t1max)
                 : launch_fw<double, true>(m.dD, m.d_fwt, nt, m.d_bnd, m.d_bnorm, m.shf_big, c.side,
                                           t1max))
            return 1;
    }
    (void)hipEventRecord(c.ev[6], s);
    if (fp32 ? launch_fw<float>(m.dD, m.d_fwt, nt, m.d_bnd, m.d_bnorm, shf, s, t1max)
             : launch_fw<double>(m.dD, m.d_fwt, nt, m.d_bnd, m.d_bnorm, shf, s, t1max))
        return 1;
    (void)hipEventRecord(c.ev[7], s);
    if (m.any_big) HIP_OK(hipStreamSynchronize(c.side));
    // tiles whose start is to be checked: those with a boundary slot in
    std::vector<int> chk;
    for (int k = 0; k < nt; k++)
        if (tl[k].bin >= 0) chk.push_back(k);
    std::vector<int> flags;
    bool first = true;
    while (!chk.empty()) {
        const int nc = (int)chk.size();
        HIP_OK(hipMemcpyAsync(m.d_fidx, chk.data(), sizeof(int) * nc, hipMemcpyHostToDevice, s));
        if (fp32)
            hipLaunchKernelGGL(k_fw_check<float>, dim3(nc), dim3(64), 0, s, m.dD, m.d_fwt, m.d_fidx,
                               (const float *)m.d_bnd, m.d_bnorm, opt.fw_tol, m.d_fflag);
        else
            hipLaunchKernelGGL(k_fw_check<double>, dim3(nc), dim3(64), 0, s, m.dD, m.d_fwt, m.d_fidx,
                               (const double *)m.d_bnd, m.d_bnorm, opt.fw_tol, m.d_fflag);
        HIP_OK(hipGetLastError());
        flags.resize(nc);
        HIP_OK(hipMemcpyAsync(flags.data(), m.d_fflag, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        std::vector<int> redo;
        for (int x = 0; x < nc; x++)
            if (flags[x]) redo.push_back(chk[x]);
        if (first) {
            // the instances that will be re-run; the others are final
            first = false;
            std::vector<char> busy(m.h.size(), 0);
            for (int k : redo) busy[tl[k].inst] = 1;
            if (first_check(busy)) return 1;
        }
        if (redo.empty()) break;
        // re-run from the boundary column, then check the tiles after them
        std::vector<FwTile> sub;
        for (int k : redo) {
            FwTile &t = tl[k];
            t.mode = 2;
            t.c0 = t.cs;
            sub.push_back(t);
        }
        rs.fw_reruns += (int)redo.size();
        if (trace_tiles()) {
            for (int k : redo)
                fprintf(stderr, "fw re-run: inst %d (L %d, R %d) tile [%d, %d)\n", tl[k].inst,
                        m.h[tl[k].inst].L, m.h[tl[k].inst].R, tl[k].cs, tl[k].c1);
        }
        // the tile list keeps every tile's index for the checks; the re-run
        // tiles are launched from a list of their own
        const int nr = (int)sub.size();
        HIP_OK(hipMemcpyAsync(m.d_fwt, tl.data(), sizeof(FwTile) * nt, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(m.d_fwre, sub.data(), sizeof(FwTile) * nr, hipMemcpyHostToDevice, s));
        if (fp32 ? launch_fw<float>(m.dD, m.d_fwre, nr, m.d_bnd, m.d_bnorm, shf, s, t1max,
                                    opt.fw_tol, m.d_fflag)
                 : launch_fw<double>(m.dD, m.d_fwre, nr, m.d_bnd, m.d_bnorm, shf, s, t1max,
                                     opt.fw_tol, m.d_fflag))
            return 1;
        if (m.any_big &&
            (fp32 ? launch_fw<float, true>(m.dD, m.d_fwre, nr, m.d_bnd, m.d_bnorm, m.shf_big, s,
                                           t1max, opt.fw_tol, m.d_fflag)
                  : launch_fw<double, true>(m.dD, m.d_fwre, nr, m.d_bnd, m.d_bnorm, m.shf_big, s,
                                            t1max, opt.fw_tol, m.d_fflag)))
            return 1;
        // a tile that rejoined its stored columns kept its boundary column;
        // the tile after one that ran to its end is checked again
        std::vector<int> early(nr);
        HIP_OK(hipMemcpyAsync(early.data(), m.d_fflag, sizeof(int) * nr, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        chk.clear();
        for (int x = 0; x < nr; x++)
            if (!early[x] && next[redo[x]] >= 0) chk.push_back(next[redo[x]]);
        if (trace_tiles()) {
            for (int x = 0; x < nr; x++)
                fprintf(stderr, "  inst %d tile [%d, %d): %s %d\n", sub[x].inst, sub[x].cs,
                        sub[x].c1, early[x] ? "rejoined at" : "ran to", early[x] ? early[x] - 1 : sub[x].c1);
            fprintf(stderr, "fw re-run round: %d tiles, %.1f ms since the forward was queued\n", nr,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_fw).count());
        }
    }
    if (first) return first_check(std::vector<char>(m.h.size(), 0));
    return 0;
}

int Batch::run(const BatchOptions &opt, RunStats *st)
{
    BATCH_LOCK;
    if (ctx_init()) return 1;
    Impl &m = *impl_;
    if (!m.uploaded && upload()) return 1;
    Streams &c = m.st;
    const int n = (int)m.hosts.size();
    hipStream_t s = c.main;
    const bool fp32 = opt.prec == Precision::FP32;
    const size_t rsz = fp32 ? sizeof(float) : sizeof(double);

    std::vector<long long> coloff(n + 1, 0);
    std::vector<int> blkoff(n + 1, 0);
    for (int i = 0; i < n; i++) {
        coloff[i + 1] = coloff[i] + m.h[i].L;
        blkoff[i + 1] = blkoff[i] + m.h[i].B;
    }
    // per-instance work arrays
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.work.p : nullptr;
        for (int i = 0; i < n; i++) {
            DevInst &d = m.h[i];
            const int B = d.B, N = d.N, L = d.L, T1 = d.T - 1, RW = 64 * d.R;
            const int Stot = m.soff[i][B];
            d.vflag = p.take<int>(L);
            d.cls = p.take<int>(L);
            d.vcol = p.take<int>(L);
            d.snode = p.take<int>(Stot);
            d.stime = p.take<int>(Stot);
            d.eterm = p.take<double>((size_t)Stot * 6);
            d.order = p.take<int>((size_t)B * N);
            d.roots = p.take<int>(B);
            d.nmut = p.take<double>((size_t)B * N * 2);
            d.tm = p.take<double>((size_t)B * T1 * T1);
            d.tm2s = p.take<double>((size_t)B * d.TC);
            d.srcs = p.take<int>(2 * (size_t)B);
            d.linfo = p.take<int>((size_t)B * RW);
            // coefficient planes in the run's precision only
            d.fp32 = fp32 ? 1 : 0;
            d.lcoef = fp32 ? nullptr : p.take<double>((size_t)B * NCOEF * RW);
            d.lcoef32 = fp32 ? p.take<float>((size_t)B * NCOEF * RW) : nullptr;
            d.ldet = p.take<int>((size_t)B * RW);
            d.lprior = p.take<double>(RW);
            d.norms = p.take<double>(L);
            d.path = p.take<int>(L);
            d.lnl = p.take<double>(1);
        }
        if (pass == 0 && m.work.ensure(p.used + 4096)) return 1;
    }
    // forward tables
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.tab.p : nullptr;
        for (int i = 0; i < n; i++) {
            DevInst &d = m.h[i];
            d.alpha = p.take<char>((size_t)d.L * 64 * d.R * rsz);
        }
        if (pass == 0 && m.tab.ensure(p.used + 4096)) return 1;
    }
    // auxiliary arrays
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.aux.p : nullptr;
        m.dD = p.take<DevInst>(n);
        m.d_coloff = p.take<long long>(n + 1);
        m.d_blkoff = p.take<int>(n + 1);
        m.d_voff = p.take<int>(n + 1);
        m.d_nvar = p.take<int>(n);
        m.d_list = p.take<int>(n);
        if (pass == 0 && m.aux.ensure(p.used + 4096)) return 1;
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
    for (int i = 0; i < n; i++)
        if (m.hosts[i].prep) m.nvar[i] = m.hosts[i].prep->nvar;

    // variant rows, allocated now that the counts are known
    std::vector<int> voff(n + 1, 0);
    for (int pass = 0; pass < 2; pass++) {
        Plan p;
        p.base = pass ? m.vbuf.p : nullptr;
        for (int i = 0; i < n; i++) {
            m.h[i].nvar = m.nvar[i];
            m.h[i].vemit = p.take<double>((size_t)std::max(m.nvar[i], 1) * 64 * m.h[i].R);
            voff[i + 1] = voff[i] + m.nvar[i];
        }
        if (pass == 0 && m.vbuf.ensure(p.used + 4096)) return 1;
    }
    HIP_OK(hipMemcpyAsync(m.d_voff, voff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(m.dD, m.h.data(), sizeof(DevInst) * n, hipMemcpyHostToDevice, s));

    (void)hipEventRecord(c.ev[1], s);
    hipLaunchKernelGGL(k_prep_b, dim3(blkoff[n]), dim3(256), 0, s, m.dD, m.d_blkoff, n);
    HIP_OK(hipGetLastError());
    // caller-given first-column priors, in the lane-run layout
    std::vector<std::vector<double>> lpri(n);
    for (int i = 0; i < n; i++) {
        if (m.priors[i].empty()) continue;
        const int R = m.h[i].R;
        lpri[i].assign(64 * R, 0.0);
        for (size_t k = 0; k < m.priors[i].size(); k++) lpri[i][lane_pos((int)k, R)] = m.priors[i][k];
        HIP_OK(hipMemcpyAsync(m.h[i].lprior, lpri[i].data(), sizeof(double) * lpri[i].size(),
                              hipMemcpyHostToDevice, s));
    }
    // host-prepared instances: their per-block inputs, column classes and
    // variant rows (the device kernels above skipped them)
    m.staged32.clear();
    for (int i = 0; i < n; i++) {
        const HostPrep *hp = m.hosts[i].prep.get();
        if (!hp) continue;
        const DevInst &d = m.h[i];
        const int B = d.B, RW = 64 * d.R, T1 = d.T - 1;
        if (fp32) {
            m.staged32.emplace_back(hp->planes.begin(), hp->planes.end());
            HIP_OK(hipMemcpyAsync(d.lcoef32, m.staged32.back().data(),
                                  sizeof(float) * hp->planes.size(), hipMemcpyHostToDevice, s));
        } else {
            HIP_OK(hipMemcpyAsync(d.lcoef, hp->planes.data(), sizeof(double) * hp->planes.size(),
                                  hipMemcpyHostToDevice, s));
        }
        HIP_OK(hipMemcpyAsync(d.linfo, hp->info.data(), sizeof(int) * (size_t)B * RW,
                              hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(d.ldet, hp->det.data(), sizeof(int) * (size_t)B * RW,
                              hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(d.tm, hp->tm.data(), sizeof(double) * (size_t)B * T1 * T1,
                              hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(d.tm2s, hp->tm2s.data(), sizeof(double) * (size_t)B * d.TC,
                              hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(d.srcs, hp->srcs.data(), sizeof(int) * 2 * (size_t)B,
                              hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(d.lprior, hp->prior.data(), sizeof(double) * RW,
                              hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(d.cls, hp->cls.data(), sizeof(int) * d.L, hipMemcpyHostToDevice, s));
        if (hp->nvar > 0)
            HIP_OK(hipMemcpyAsync(d.vemit, hp->vemit.data(), sizeof(double) * hp->vemit.size(),
                                  hipMemcpyHostToDevice, s));
    }
    (void)hipEventRecord(c.ev[2], s);
    if (voff[n] > 0) {
        int nmax = 1;
        for (int i = 0; i < n; i++) nmax = std::max(nmax, m.h[i].N);
        hipLaunchKernelGGL(k_var_emit_b, dim3(voff[n]), dim3(64), var_emit_lds(nmax), s, m.dD,
                           m.d_voff, n, nmax);
    }
    HIP_OK(hipGetLastError());
    (void)hipEventRecord(c.ev[3], s);

    // forward: one wave per tile, all tiles in one launch, the longest
    // instances first
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const double ca = (double)m.h[a].L * m.h[a].R, cb = (double)m.h[b].L * m.h[b].R;
        return ca > cb;
    });
    HIP_OK(hipMemcpyAsync(m.d_list, order.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    size_t shf = 0;
    int t1max = 0;
    m.any_big = false;
    m.shf_big = 0;
    for (int i = 0; i < n; i++) {
        const DevInst &d = m.h[i];
        t1max = std::max(t1max, d.T - 1);
        size_t &sh = d.R > 16 ? m.shf_big : shf;
        sh = std::max(sh, rsz * (size_t)fw_lds_reals_r(fp32, d.R));
        m.any_big |= d.R > 16;
    }
    RunStats rs;
    // Once the first check of the forward tiles has shown which instances
    // are re-run, the traceback options of all the others go to the side
    // stream and run while the re-runs continue on the main stream.
    // A helper host thread then resolves those instances' tiles on the side
    // stream (its re-runs too), while this thread drives the forward re-runs
    // and then the re-run instances' traceback on the main stream.
    size_t sht = 0;
    std::vector<char> busy, done;
    const int K = std::max(1, std::min(opt.tile_opts, TB_KMAX));
    std::thread helper;
    RunStats rs_side;
    int rc_side = 0;
    auto early_tb = [&](const std::vector<char> &b) -> int {
        if (!opt.traceback || std::find(b.begin(), b.end(), 1) == b.end()) return 0;
        busy = b;
        if (plan_tiles(m, fp32, opt.tile_cols, opt.tile_warm, K, sht, &b, opt.tail_tile_cols))
            return 1;
        HIP_OK(hipStreamSynchronize(s));   // the tile list is on the device
        done.assign(n, 0);
        for (int i = 0; i < n; i++) done[i] = !b[i];
        if (launch_options(m, fp32, sht, &done, 0, c.side)) return 1;
        if (opt.fixup)
            helper = std::thread([&] {
                rc_side = resolve_tiles(m, fp32, sht, rs_side, c.side, &done, 1);
                (void)hipEventRecord(c.ev[8], c.side);
            });
        return 0;
    };
    // the helper is joined on every path out of run()
    struct Joiner {
        std::thread &t;
        ~Joiner()
        {
            if (t.joinable()) t.join();
        }
    } joiner{helper};
    if (forward_stage(m, fp32, opt, order, shf, t1max, rs, early_tb)) return 1;
    (void)hipEventRecord(c.ev[4], s);
    hipLaunchKernelGGL(k_lnl_b, dim3(n), dim3(256), 0, s, m.dD, m.d_list);
    HIP_OK(hipGetLastError());
    // traceback tiles: the options of every tile, then their resolution
    if (opt.traceback) {
        if (busy.empty()) {
            if (traceback_stage(m, fp32, opt, rs)) return 1;
        } else {
            if (launch_options(m, fp32, sht, &busy, 1, s)) return 1;
            if (opt.fixup && resolve_tiles(m, fp32, sht, rs, s, &busy, 0)) return 1;
            (void)hipEventRecord(c.ev[9], s);
            if (helper.joinable()) helper.join();
            HIP_OK(hipStreamSynchronize(c.side));
            if (rc_side) return 1;
            rs.tb_passes = std::max(rs.tb_passes, rs_side.tb_passes);
            rs.tb_rerun_cols += rs_side.tb_rerun_cols;
            rs.tb_adopted += rs_side.tb_adopted;
        }
    }
    (void)hipEventRecord(c.ev[5], s);
    for (int i = 0; i < n; i++)
        HIP_OK(hipMemcpyAsync(&m.lnl[i], m.h[i].lnl, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    m.prec_done = opt.prec;
    float ms;
    (void)hipEventElapsedTime(&ms, c.ev[1], c.ev[2]); rs.ms_prep = ms;
    (void)hipEventElapsedTime(&ms, c.ev[2], c.ev[3]); rs.ms_emit = ms;
    // forward: until the forward kernel ends; traceback: from there to the
    // end of the fix-up passes
    float fwd = 0;
    (void)hipEventElapsedTime(&fwd, c.ev[3], c.ev[4]);
    rs.ms_forward = fwd;
    (void)hipEventElapsedTime(&ms, c.ev[6], c.ev[7]);
    rs.ms_fw_main = ms;
    (void)hipEventElapsedTime(&ms, c.ev[3], c.ev[5]);
    rs.ms_traceback = opt.traceback ? std::max(0.f, ms - fwd) : 0.f;
    (void)hipEventElapsedTime(&ms, c.ev[0], c.ev[5]); rs.ms_total = ms;
    if (opt.traceback && opt.fixup && !busy.empty()) {
        // the two halves of the overlapped traceback, from the end of the
        // forward: the re-run instances' (main stream) and the others' (side)
        (void)hipEventElapsedTime(&ms, c.ev[4], c.ev[9]); rs.ms_tb_tail = ms;
        (void)hipEventElapsedTime(&ms, c.ev[4], c.ev[8]); rs.ms_tb_side = ms;
    }
    for (int i = 0; i < n; i++) {
        rs.W += m.W[i];
        rs.nvar += m.nvar[i];
    }
    if (st) *st = rs;
    return 0;
}

int Batch::copy_path(int i, int *path) const
{
    BATCH_LOCK;
    const Impl &m = *impl_;
    HIP_OK(hipMemcpy(path, m.h[i].path, sizeof(int) * m.h[i].L, hipMemcpyDeviceToHost));
    return 0;
}

int Batch::copy_norms(int i, double *norms) const
{
    BATCH_LOCK;
    const Impl &m = *impl_;
    HIP_OK(hipMemcpy(norms, m.h[i].norms, sizeof(double) * m.h[i].L, hipMemcpyDeviceToHost));
    return 0;
}

// block-major view (block b holds blocklen_b rows of S_b states, the
// reference's table layout) of the lane-run table
int Batch::copy_alpha(int i, double *alpha) const
{
    BATCH_LOCK;
    const Impl &m = *impl_;
    const DevInst &d = m.h[i];
    const FlatTrees &ft = m.hosts[i].ft;
    const int R = d.R, RW = 64 * R;
    const size_t cells = (size_t)d.L * RW;
    std::vector<double> tab(cells);
    if (m.prec_done == Precision::FP64) {
        HIP_OK(hipMemcpy(tab.data(), d.alpha, sizeof(double) * cells, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t32(cells);
        HIP_OK(hipMemcpy(t32.data(), d.alpha, sizeof(float) * cells, hipMemcpyDeviceToHost));
        for (size_t x = 0; x < cells; x++) tab[x] = t32[x];
    }
    std::vector<int> pos;
    long long o = 0;
    for (int b = 0; b < ft.nblocks; b++) {
        const int S = m.Sv[i][b];
        pos.resize(S);
        for (int k = 0; k < S; k++) pos[k] = lane_pos(k, R);
        for (int cc = 0; cc < ft.len[b]; cc++) {
            const double *row = &tab[(size_t)(ft.start[b] + cc) * RW];
            for (int k = 0; k < S; k++) alpha[o++] = row[pos[k]];
        }
    }
    return 0;
}

int Batch::copy_alpha_range(int i, int c0, int c1, double *alpha) const
{
    BATCH_LOCK;
    const Impl &m = *impl_;
    const DevInst &d = m.h[i];
    const FlatTrees &ft = m.hosts[i].ft;
    if (c0 < 0 || c1 > d.L || c0 >= c1) return 1;
    const int R = d.R, RW = 64 * R;
    const size_t cells = (size_t)(c1 - c0) * RW;
    std::vector<double> tab(cells);
    if (m.prec_done == Precision::FP64) {
        HIP_OK(hipMemcpy(tab.data(), (const double *)d.alpha + (size_t)c0 * RW,
                         sizeof(double) * cells, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t32(cells);
        HIP_OK(hipMemcpy(t32.data(), (const float *)d.alpha + (size_t)c0 * RW,
                         sizeof(float) * cells, hipMemcpyDeviceToHost));
        for (size_t x = 0; x < cells; x++) tab[x] = t32[x];
    }
    long long o = 0;
    int b = 0;
    for (int c = c0; c < c1; c++) {
        while (ft.start[b] + ft.len[b] <= c) b++;
        const int S = m.Sv[i][b];
        const double *row = &tab[(size_t)(c - c0) * RW];
        for (int k = 0; k < S; k++) alpha[o++] = row[lane_pos(k, R)];
    }
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

int instance_supported(const HostInstance &hi, std::string *msg)
{
    char buf[160];
    buf[0] = 0;
    const FlatTrees &ft = hi.ft;
    if (ft.nnodes > MAXN || hi.model.T > MAXT) {
        snprintf(buf, sizeof buf, "instance has %d nodes and %d time points (limits %d and %d)",
                 ft.nnodes, hi.model.T, MAXN, MAXT);
    } else if (!hi.prep) {
        std::vector<int> Sv;
        instance_states(ft, hi.model.T, &Sv);
        int Smax = 1;
        for (int S : Sv) Smax = std::max(Smax, S);
        if (lanes_per_state_r(Smax) < 0)
            snprintf(buf, sizeof buf, "a block has %d states (limit %d per block)", Smax,
                     64 * RVALS[NRV - 1]);
    }
    if (msg) *msg = buf;
    return buf[0] ? 1 : 0;
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

// one instance on the given batch (cleared first)
static int run_one(Batch &bt, const HostInstance &hi, const double *u, int *path, double *alpha,
                   double *lnl, RunStats *stats, Precision prec, const double *prior,
                   int tile_cols)
{
    bt.clear();
    if (bt.add(hi, u, prior) < 0) return 1;
    if (bt.upload()) return 1;
    BatchOptions opt;
    opt.prec = prec;
    opt.traceback = (u != nullptr);
    // long instances trace back in tiles (the thread is the same as with one
    // tile: fix-up passes re-sample until the chains coalesce)
    opt.tile_cols = (tile_cols == 0 && hi.ft.seqlen >= 16384) ? 4096 : tile_cols;
    // FP32 forwards of long instances run as tiles (FP64 keeps the
    // sequential forward: its contract is the reference's table)
    if (prec == Precision::FP32) {
        opt.fw_tile = 50000;
        opt.fw_warm = 20000;
        opt.fw_tol = 1e-4;
    }
    if (bt.run(opt, stats)) return 1;
    if (path && u && bt.copy_path(0, path)) return 1;
    if (alpha && bt.copy_alpha(0, alpha)) return 1;
    if (lnl) *lnl = bt.lnl(0);
    return 0;
}

// one instance = a batch of one; the batch (its device buffers and streams)
// is kept per host thread, so repeated calls (every re-threading of
// resample_arg) reuse the allocations
int run_instance(const HostInstance &hi, const double *u, int *path, double *alpha, double *lnl,
                 RunStats *stats, Precision prec, const double *prior, int tile_cols)
{
    // never destroyed: releasing HIP objects from exit-time destructors can
    // run after the runtime itself has shut down
    static thread_local Batch *btp = new Batch();
    return run_one(*btp, hi, u, path, alpha, lnl, stats, prec, prior, tile_cols);
}

// The no-allocation workspace: one fixed device array, of which every GPU
// the library runs on holds its own copy until the process ends.  Host code
// never uses the name `scratch` as an address (on the host it is a stand-in
// symbol, not device memory): the copy on each device is found once, with
// that device current, by hipGetSymbolAddress, and that pointer is what the
// batch's buffers are carved from.
static __device__ __align__(256) char scratch[kNoAllocBytes];

namespace {
struct NoAllocDevice {
    char *ptr = nullptr;     // device address of this GPU's scratch
    Batch *batch = nullptr;  // never destroyed (see run_instance)
};
}  // namespace

// Set-up of the current device, once: the scratch address and a batch whose
// buffers live in it, with its stream and events created here, so that the
// calls after it allocate nothing.
static int noalloc_device(NoAllocDevice *&out)
{
    static std::mutex mu;
    static std::map<int, NoAllocDevice> devs;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = devs.find(dev);
    if (it == devs.end()) {
        NoAllocDevice d;
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "argweaver_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
                    dev, hipGetErrorString(e));
            return 1;
        }
        d.ptr = static_cast<char *>(p);
        d.batch = new Batch();
        d.batch->use_workspace(d.ptr, kNoAllocBytes);
        if (d.batch->init_streams()) return 1;
        it = devs.emplace(dev, d).first;
    }
    out = &it->second;
    return 0;
}

int run_instance_noalloc(const HostInstance &hi, const double *u, int *path, double *alpha,
                         double *lnl, RunStats *stats, Precision prec)
{
    NoAllocDevice *d = nullptr;
    if (noalloc_device(d)) return 1;
    DEVICE_LOCK;   // calls on one GPU share its scratch: one at a time
    return run_one(*d->batch, hi, u, path, alpha, lnl, stats, prec, nullptr, 0);
}

}  // namespace aw
