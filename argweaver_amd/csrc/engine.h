// engine.h -- device side of the threading HMM: per-instance HBM layout,
// batches of instances, and the host entry points that run the kernels on
// the current HIP device.
#pragma once

#include <cstdint>
#include <vector>

#include "aw_math.h"
#include "local_trees.h"

namespace aw {

// Everything one threading instance keeps in HBM.  POD; an array of these
// lives in device memory and every kernel finds its instance through it.
// Per-block arrays are indexed through soff (states) and aoff (forward
// table).  Switch arrays for the switch into block b are indexed by the
// source block's state offset (determ/determp/invidx) or by the target
// block's (recombrow/recoalrow, invptr with one extra entry per block).
struct DevInst {
    int T, N, B, L, nrows, Smax, TC, RM, nvar;
    long long W;
    const Model *model;
    // trees
    const int *parent, *c0, *c1, *age, *map, *spr, *bstart, *blen;
    const int *soff;          // [B+1]
    const long long *aoff;    // [B+1]
    const char *seq;          // [nrows*L], leaves then the new sequence
    const double *u;          // [L] uniforms, u[L-1-col] used at column col
    // column classes
    int *vflag, *cls, *vcol;  // [L]
    // prep outputs (FP64)
    int *snode, *stime;       // [Stot]
    int *nfirst, *nhi;        // [B*N]
    int *tperm, *tstart;      // [Stot], [B*(T+1)]
    double *tm, *tm2, *tm2s;  // [B*(T-1)^2], [B*TC], [B*TC]
    double *tv;               // [B*9*T] transition vectors (D,E,...,norecombs)
    double *einv, *eterm;     // [Stot], [Stot*6]
    int *determ;              // [Stot]
    double *determp;          // [Stot]
    double *recombrow, *recoalrow;   // [Stot]
    int *srcs;                // [2B]
    int *invptr, *invidx;     // [Stot+B], [Stot]
    double *prior;            // [Smax]
    int *order;               // [B*N] postorder
    int *roots;               // [B]
    double *nmut;             // [B*N*2]
    double *vemit;            // [nvar*Smax]
    // FP32 path: tables and the time-grouped lane assignment of every block
    float *tm32, *tm2_32;     // [B*(T-1)^2], [B*TC]
    float *einv32;            // [Stot]
    int *lslot;               // [B*64*RM] state index per (lane, slot), -1 empty
    int *lkey;                // [B*64] time of the lane's states (T-1 when empty)
    // outputs
    double *alpha;            // [W] FP64 table
    float *alpha32;           // [W] FP32 table (normalised columns)
    double *norms;            // [L] column normalisers
    int *path;                // [L]
    double *lnl;              // [1]
};

// offset of the (c, a, b) entry of the per-block same-branch table
__host__ __device__ inline int tc_base(int c, int T1)
{
    int s = 0;
    for (int x = 0; x < c; x++) s += (T1 - x) * (T1 - x);
    return s;
}

enum class Precision { FP64 = 0, FP32 = 1 };

struct RunStats {
    double ms_prep = 0, ms_emit = 0, ms_forward = 0, ms_traceback = 0, ms_total = 0;
    long long W = 0;   // column-states
    int nvar = 0;
    int fw_passes = 0, tb_passes = 0;   // fix-up passes of the column-tiled schedule
    long long fw_rerun_cols = 0, tb_rerun_cols = 0;
};

// One threading instance described on the host.
struct HostInstance {
    FlatTrees ft;
    std::vector<const char *> rows;   // nleaves rows in leaf order, then the new sequence
    Model model;
};

// Schedule of a batch run.
struct BatchOptions {
    Precision prec = Precision::FP64;
    int chunk_cols = 0;      // FP32 column tile length (0: one tile per instance)
    bool traceback = true;
    bool fixup = true;       // false: speculative tiles only (diagnostics)
    // a re-run FP32 column agrees with the stored one when every entry is
    // within tol_rel * |stored| + tol_abs (columns sum to one)
    float tol_rel = 1e-6f, tol_abs = 1e-9f;
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
    // the instance's row pointers and uniforms must stay valid until upload()
    int add(const HostInstance &hi, const double *u, const double *prior = nullptr);
    int upload();
    int run(const BatchOptions &opt, RunStats *st = nullptr);
    int copy_path(int i, int *path) const;
    int copy_alpha(int i, double *alpha) const;   // block-major FP64 view of the table
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
                 const double *prior = nullptr, int chunk_cols = 0);

// Number of column-states W and per-block state counts (host, O(B*N)).
long long instance_states(const FlatTrees &ft, int T, std::vector<int> *S);

// Device attributes and availability.
bool device_available();
const char *device_name();
int device_info(char *buf, int n);

}  // namespace aw
