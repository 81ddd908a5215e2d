// engine.h -- device side of the threading HMM: per-instance HBM layout,
// batches of instances, and the host entry points that run the kernels on
// the current HIP device.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include <memory>

#include "aw_math.h"
#include "internal_prep.h"
#include "local_trees.h"

namespace aw {

// Everything one threading instance keeps in HBM.  POD; an array of these
// lives in device memory and every kernel finds its instance through it.
//
// Lane-run layout.  The HMM kernels run one 64-lane wave per instance (or
// per traceback tile); lane l owns the R consecutive states l*R .. l*R+R-1
// of the node-major state order, so the states of one branch sit in one
// lane or in neighbouring lanes.  Every per-state array is stored
// transposed, position(s) = (s % R) * 64 + s / R, in rows of RW = 64*R
// entries: slot r of all 64 lanes is one contiguous 64-entry segment, so
// the loads and stores of a slot are coalesced.  R is fixed per instance
// (the smallest supported value with 64*R >= max states per block), so
// forward-table rows have a constant stride and column col starts at
// col * RW.
struct DevInst {
    int T, N, B, L, nrows, Smax, TC, R, nvar;
    long long W;              // column-states sum_b blocklen_b * S_b
    const Model *model;
    // trees
    const int *parent, *c0, *c1, *age, *map, *spr, *bstart, *blen;
    const int *soff;          // [B+1] state offsets (state-ordered arrays)
    const char *seq;          // [nrows*L], leaves then the new sequence
    const double *u;          // [L] uniforms, u[L-1-col] used at column col
    // column classes: -1 invariant, -2 masked, >=0 variant row index
    int *vflag, *cls, *vcol;  // [L]
    // prep outputs in state order (pruning and traceback lookups)
    int *snode, *stime;       // [Stot]
    double *eterm;            // [Stot*6]
    int *order, *roots;       // [B*N] node levels: height | depth << 16, [B] roots
    double *nmut;             // [B*N*2]
    double *tm, *tm2s;        // [B*(T-1)^2] P(a->b) other branch, [B*TC] same branch
    int *srcs;                // [2B] recombination / recoalescence source of the switch into b
    // prep outputs in the lane-run layout, one row of RW per block
    int *linfo;               // [B*RW] time | flags | age(node) << 11 | node << 16
    double *lcoef;            // [B*NCOEF*RW] coefficient planes (see Coef), FP64 runs
    float *lcoef32;           // [B*NCOEF*RW] the planes in FP32, FP32 runs
    int fp32;                 // the run is FP32
    int last_state;           // the state of the last column when given, else -1 (sampled)
    int hostprep;             // the per-block inputs come from the host (internal threading)
    int *ldet;                // [B*RW] deterministic switch target of block b-1's states
    double *lprior;           // [RW] first-column prior
    double *vemit;            // [nvar*RW] emissions of variant columns
    // outputs
    void *alpha;              // [L*RW] forward table, double or float
    double *norms;            // [L] column normalisers
    int *path;                // [L]
    double *lnl;              // [1]
};

// state flags in linfo
enum : int {
    INFO_TIME = 0xff,
    INFO_FIRST = 1 << 8,   // first (youngest) state of its branch
    INFO_LAST = 1 << 9,    // last (oldest) state of its branch
    INFO_VALID = 1 << 10,  // a real state (not row padding)
};
__host__ __device__ inline int info_age(int x) { return (x >> 11) & 31; }
__host__ __device__ inline int info_node(int x) { return x >> 16; }

// Coefficient planes of a block.  With the forward column x, the
// within-block recursion at state s = (node n, time b), c = age(n), is
//   sum_j x_j P(j -> s) = tf[b] + CG_s * P1_s + CK_s * Q_s + CJ_s * x_s
// where tf[b] = sum_a G[a] P_other(a -> b) over the per-time sums G,
// P1_s = sum of CF_j x_j over the states j of n younger than s and
// Q_s = sum of CD_j x_j over the states of n older than s.
enum : int {
    CF = 0, CD, CG, CK, CJ,   // same-branch factors
    CE,                       // emission of an invariant column
    CRR, CCR,                 // recombination / recoalescence rows of the switch into b
    CDP,                      // deterministic switch weight of block b-1's states
    NCOEF
};

// transposed position of state s in a lane-run row
__host__ __device__ inline int lane_pos(int s, int R) { return (s % R) * 64 + s / R; }

// offset of the (c, a, b) entry of the per-block same-branch table:
// sum_{x<c} (T1-x)^2, in closed form
__host__ __device__ inline int tc_base(int c, int T1)
{
    return c * T1 * T1 - T1 * c * (c - 1) + (c - 1) * c * (2 * c - 1) / 6;
}

enum class Precision { FP64 = 0, FP32 = 1 };

struct RunStats {
    double ms_prep = 0, ms_emit = 0, ms_forward = 0, ms_traceback = 0, ms_total = 0;
    long long W = 0;   // column-states
    int nvar = 0;
    int tb_passes = 0;            // re-run passes of the tiled traceback
    long long tb_rerun_cols = 0;  // columns re-sampled by those passes
    int tb_adopted = 0;           // tiles that took an alternative option
    double ms_fw_main = 0;        // the main forward launch (without re-runs and checks)
    int fw_tiles = 0;             // forward tiles launched
    int fw_reruns = 0;            // forward tiles re-run from their boundary column
    // with forward re-runs: from the end of the forward to the end of the
    // re-run instances' traceback, and to the end of the others' (which
    // started beside the re-runs; negative if it ended before the forward)
    double ms_tb_tail = 0, ms_tb_side = 0;
};

// One threading instance described on the host.
struct HostInstance {
    FlatTrees ft;
    std::vector<const char *> rows;   // nleaves rows in leaf order, then the new sequence
    Model model;
    // internal threading: the per-block inputs prepared on the host (the
    // device preparation, classification and emission kernels skip it)
    std::shared_ptr<const HostPrep> prep;
    // the last column's state when the traceback is conditioned on it (its
    // index in the last block's state order), else -1: the last column's
    // uniform (u[0]) is then not used
    int last_state = -1;
};

// Schedule of a batch run.  The forward runs one wave per instance over
// all of its columns, or, in FP32, one wave per forward tile (checked
// speculative starts); the traceback runs one wave per option of a column
// tile.
struct BatchOptions {
    Precision prec = Precision::FP64;
    int tile_cols = 0;       // traceback tile length (0: one tile per instance)
    int tile_warm = 256;     // warm-up columns of a speculative tile start
    int tile_opts = 3;       // options per traceback tile (1 .. 4): the
                             // speculative chain and its alternatives
    // traceback tile length of the instances whose FP32 forward tiles are
    // re-run: their traceback runs after the forward, when the rest of the
    // batch is done or on the other stream, so shorter tiles spread it over
    // more of the device (0: tile_cols)
    int tail_tile_cols = 4000;
    bool traceback = true;
    bool fixup = true;       // false: speculative option-0 paths only (diagnostics)
    // FP32 forward tiles: long instances are split into tiles (counting the
    // warm-up) of fw_tile columns at a column cost of 0.94 us, scaled by
    // the column cost of the instance's states per lane, that run
    // concurrently; a tile after the first starts fw_warm columns early
    // from a uniform vector, and its start is checked against the column
    // the tile before computes (relative tolerance fw_tol), a tile that
    // fails being re-run from that column.  0: one tile per instance (FP64
    // runs always use one).
    int fw_tile = 0;
    int fw_warm = 25000;
    double fw_tol = 1e-5;
};

// A set of independent threading instances resident in HBM.  upload()
// places the inputs on the device; run() executes prep, emissions, forward
// and traceback for all of them (the timed region of the benchmark);
// results are then copied out per instance.
class Batch {
public:
    Batch();
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;
    // Take every device buffer from [base, base + cap) instead of hipMalloc
    // (re-carved from the start on clear(); a run that needs more fails).
    void use_workspace(void *base, size_t cap);
    // create the batch's stream and events now (else the first upload does)
    int init_streams();
    // false (default): the batch's device work takes the process-wide lock
    // and never overlaps another batch's; true: it takes a lock of its own
    // and may run beside other batches (the server's drivers)
    void set_concurrent(bool on);
    // the instance's row pointers and uniforms must stay valid until upload()
    int add(const HostInstance &hi, const double *u, const double *prior = nullptr);
    // the same, taking the instance's arrays over instead of copying them
    int add(HostInstance &&hi, const double *u, const double *prior = nullptr);
    // drop the instances, keep the device buffers for the next set
    void clear();
    int upload();
    int run(const BatchOptions &opt, RunStats *st = nullptr);
    // Autotuner of the traceback schedule: runs the forward once, then times
    // the traceback alone for each candidate tile length (0: one tile per
    // instance) on the resident tables; returns the fastest (ms per
    // candidate in ms_out when given), -1 on error.  Leaves the batch with
    // the path of the last candidate.  warm / opts: the tiles' warm-up and
    // options (BatchOptions::tile_warm, tile_opts).
    int tune_tiles(Precision prec, const int *cand, int ncand, double *ms_out = nullptr,
                   int warm = BatchOptions().tile_warm, int opts = BatchOptions().tile_opts);
    // FP32 forward-tile autotuner: times the forward (with its checks and
    // re-runs) for each (tile, warm-up) candidate and returns the index of
    // the fastest (-1 on error)
    int tune_forward(const int *tiles, const int *warms, int ncand, double tol,
                     double *ms_out = nullptr);
    int copy_path(int i, int *path) const;
    int copy_alpha(int i, double *alpha) const;   // block-major FP64 view of the table
    // columns [c0, c1) of it (S of the column's block per column)
    int copy_alpha_range(int i, int c0, int c1, double *alpha) const;
    int copy_norms(int i, double *norms) const;
    double lnl(int i) const;
    long long W(int i) const;
    int size() const;

    struct Impl;

private:
    Impl *impl_;
};

// Run one instance (a batch of one).  u may be null (no traceback).
int run_instance(const HostInstance &hi, const double *u, int *path, double *alpha,
                 double *lnl, RunStats *stats, Precision prec = Precision::FP64,
                 const double *prior = nullptr, int tile_cols = 0);

// 0 when the instance fits the kernels' limits (nodes per tree, time points,
// states per block); else 1, with the reason in msg.
int instance_supported(const HostInstance &hi, std::string *msg = nullptr);

// Bytes of the fixed per-GPU workspace of run_instance_noalloc.
constexpr size_t kNoAllocBytes = size_t(64) << 20;

// run_instance (FP64 or FP32, no prior) on a fixed workspace of
// kNoAllocBytes per GPU instead of allocated buffers: after the first call
// on a device (which finds the workspace and creates the stream) it calls
// no device allocator.  Fails (1, with a message) for an instance that
// needs more.  Calls are serialised by the library's device lock.
int run_instance_noalloc(const HostInstance &hi, const double *u, int *path, double *alpha,
                         double *lnl, RunStats *stats, Precision prec);

// Number of column-states W and per-block state counts (host, O(B*N)).
long long instance_states(const FlatTrees &ft, int T, std::vector<int> *S);

// Device attributes and availability.
bool device_available();
const char *device_name();
int device_info(char *buf, int n);

}  // namespace aw
