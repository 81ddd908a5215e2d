// engine.h -- device side of the threading HMM: per-instance HBM layout and
// the host entry points that run the kernels on the current HIP device.
#pragma once

#include <cstdint>
#include <vector>

#include "aw_math.h"
#include "local_trees.h"

namespace aw {

// Everything one threading instance keeps in HBM.  POD so it can be passed
// to kernels by value.  Per-block arrays are indexed through soff (states)
// and aoff (forward table).  Switch arrays for the switch into block b are
// indexed by the *source* block's state offset (determ/determp) or by the
// target block's (recombrow/recoalrow/invptr).
struct DevInst {
    int T, N, B, L, nrows, Smax, TC;   // TC: entries of the (c,a,b) table
    int nvar;
    const Model *model;
    // trees
    const int *parent, *c0, *c1, *age, *map, *spr, *bstart, *blen;
    const int *soff;          // [B+1]
    const long long *aoff;    // [B+1]
    const char *seq;          // [nrows*L], leaves then the new sequence
    // column classes
    int *vflag, *cls, *vcol;  // [L]
    // prep outputs
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
    double *prior;            // [S0]
    int *order;               // [B*N] postorder
    int *roots;               // [B]
    double *nmut;             // [B*N*2]
    double *vemit;            // [nvar*Smax]
    // forward / traceback
    double *alpha;            // [W] FP64 table (plain path)
    float *alpha32;           // [W] FP32 table (rescaled path)
    const double *u;          // [L] uniforms, u[L-1-col] used at column col
    int *path;                // [L]
    double *lnl;              // [1]
};

// offset of the (c, a, b) entry of the per-block same-branch table
__host__ __device__ inline int tc_base(int c, int T1)
{
    // sum_{x<c} (T1-x)^2
    int s = 0;
    for (int x = 0; x < c; x++) s += (T1 - x) * (T1 - x);
    return s;
}

enum class Precision { FP64 = 0, FP32 = 1 };

struct RunStats {
    double ms_prep = 0, ms_emit = 0, ms_forward = 0, ms_traceback = 0;
    long long W = 0;   // column-states of the instance
    int nvar = 0;
};

// One threading instance described on the host.
struct HostInstance {
    FlatTrees ft;
    std::vector<const char *> rows;   // nleaves rows in leaf order, then the new sequence
    Model model;
};

// Run the threading HMM for one instance on the current device.
//   u        -- L uniforms (u[0] for the last column), may be null (no traceback)
//   path     -- L ints out (state index within block), may be null
//   alpha    -- optional host copy of the forward table (block-major), W doubles
//   lnl      -- optional log of the forward mass
// Returns 0 on success, nonzero on a HIP error (message on stderr).
//   prior    -- optional first-column prior replacing the computed one
int run_instance(const HostInstance &hi, const double *u, int *path, double *alpha,
                 double *lnl, RunStats *stats, Precision prec = Precision::FP64,
                 const double *prior = nullptr);

// Number of column-states W and per-block state counts (host, O(B*N)).
long long instance_states(const FlatTrees &ft, int T, std::vector<int> *S);

// Device attributes and availability.
bool device_available();
const char *device_name();

}  // namespace aw
