/*
 * oracle.h -- CPU restatement of ARGweaver's threading HMM (external-branch
 * threading of one new leaf), written from the reference's algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in argweaver_amd/ links, loads or calls
 * this library.  It is used by tests/ (as the checker the HIP path is
 * compared against), by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.
 *
 * Every function names the reference file:line it restates.  The instance
 * is passed as flat arrays (the same arrays arghmm_new_trees takes), plus an
 * explicit per-tree node mapping (row 0 ignored), because mappings after
 * remove_arg_thread are not the identity-except-broken default.
 */
#ifndef AW_ORACLE_H
#define AW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int ntimes;
    const double *times;      /* [ntimes] */
    const double *popsizes;   /* [ntimes] */
    double rho, mu;
    int ntrees, nnodes, start_coord;
    const int *ptrees;        /* [ntrees*nnodes] parent pointers */
    const int *ages;          /* [ntrees*nnodes] time indices */
    const int *sprs;          /* [ntrees*4] recomb_node, recomb_time, coal_node, coal_time */
    const int *mappings;      /* [ntrees*nnodes]; NULL -> make_node_mapping default */
    const int *blocklens;     /* [ntrees] */
    int nseqs, seqlen;
    const char *const *seqs;  /* [nseqs][seqlen] */
    const int *seqids;        /* [nleaves] leaf -> sequence row */
    int new_chrom;            /* sequence row of the thread being added */
} orc_instance;

/* a1: time grid, model.h:21-32 (delta = 0.01) */
void orc_time_points(int ntimes, double maxtime, double delta, double *times);
/* a1: model.cpp:9-21 */
void orc_coal_time_steps(const double *times, int ntimes, double *coal_time_steps);

/* a3: number of states of every tree (states.cpp:78-100) */
void orc_nstates(const orc_instance *inst, int *nstates);
/* a3: state list of tree `tree` as (node,time) pairs, returns count */
int orc_states(const orc_instance *inst, int tree, int *states2);
/* a4: count_lineages (local_tree.cpp:34-69) */
void orc_lineages(const orc_instance *inst, int tree, int *nbranches, int *nrecombs, int *ncoals);
/* get_treelen (local_tree.cpp:136-155) */
double orc_treelen(const orc_instance *inst, int tree, int use_basal);

/* a5: the nine TransMatrix vectors of tree `tree` (trans.cpp:26-115),
 * out = [D,E,lnB,lnE2,lnNegG1,G2,G3,lnG4,norecombs] each [ntimes] */
void orc_trans_vectors(const orc_instance *inst, int tree, double *out);
/* a5: dense transition probabilities P(i->j) of tree `tree` (trans.h:64-128), [S*S] */
void orc_trans_dense(const orc_instance *inst, int tree, double *out);
/* a6: switch matrix between tree-1 and tree (trans.cpp:538-739).
 * determ[S1], determprob[S1], recombrow[S2], recoalrow[S2]; srcs = {recombsrc, recoalsrc} */
void orc_switch(const orc_instance *inst, int tree, int *determ, double *determprob,
                double *recombrow, double *recoalrow, int *srcs);
/* a7: state priors of tree 0 (trans.cpp:785-831) */
void orc_prior(const orc_instance *inst, double *prior);
/* a8: emissions of tree `tree` over its block, [blocklen*S] row-major (emit.cpp:650-869) */
void orc_emissions(const orc_instance *inst, int tree, double *emit);

/* a10: forward algorithm (sample_thread.cpp:394-460, 186-296, 345-389).
 * fw is block-major: block b occupies blocklen_b * S_b doubles.
 * lnl receives sum over columns 1.. of log(column normaliser) plus the log
 * of the prior's mass, i.e. log of the total forward mass. slow != 0 uses
 * the dense per-block matrix (sample_thread.cpp:303-338). */
void orc_forward(const orc_instance *inst, int slow, double *fw, double *lnl);
/* total number of doubles of the block-major forward table */
long long orc_forward_size(const orc_instance *inst);

/* a11: stochastic traceback (sample_thread.cpp:522-569) consuming one uniform
 * per column, u[0] for the last column, u[1] for the one before, ... each
 * u = rand()/RAND_MAX as frand() draws it (common.h:103-107). path[col] is
 * the state index within its block. */
void orc_traceback(const orc_instance *inst, const double *fw, const double *u, int *path);

/* a12: sample_recombinations (recomb.cpp:151-235), drawing from glibc rand().
 * Returns the number of recombinations (<= maxrecomb). */
int orc_sample_recombinations(const orc_instance *inst, const int *path,
                              int *recomb_pos, int *recomb_node, int *recomb_time,
                              int maxrecomb);

#ifdef __cplusplus
}
#endif
#endif
