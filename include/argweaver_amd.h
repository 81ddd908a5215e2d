/*
 * argweaver_amd.h -- C ABI of libargweaver_amd.so.
 *
 * The arghmm_* / get_* / delete_* entry points keep the names, argument
 * order, C types and ownership rules of the reference library's C ABI
 * (bound from Python by argweaver/argweaverc.py:29-343 through ctypes), so
 * a caller can load this library in place of libargweaver.so for the
 * threading path.  Each declaration cites the reference definition it
 * replaces (paths relative to the reference repository root).
 *
 * The awamd_* entry points are extensions with no counterpart in the
 * reference: explicit uniform streams, FP32 mode, batched and served
 * threading, and the synthetic workload generator.
 *
 * Errors: an invalid ARG aborts the process, as the reference's asserts
 * do.  Calls that run the HMM print to stderr and return NULL (or a
 * nonzero status) when no HIP device is visible, when an instance exceeds
 * the kernels' limits (256 nodes per tree, 32 time points, 1024 states per
 * block) or when a device call fails, where the reference would abort.
 * Not thread-safe; the traceback draws from the process-global rand().
 * Device work of all calls is serialised inside the library.
 */
#ifndef ARGWEAVER_AMD_H
#define ARGWEAVER_AMD_H

#ifdef __cplusplus
#include <cstddef>
extern "C" {
#else
#include <stdbool.h>
#include <stddef.h>
#endif

typedef struct LocalTrees LocalTrees;   /* opaque ARG handle */
typedef int intstate[2];                 /* (node, time) */

/*---------------------------------------------------------------------------
 * ARG handles
 */
/* src/argweaver/local_tree.cpp:1817-1824 */
LocalTrees *arghmm_new_trees(int **ptrees, int **ages, int **sprs, int *blocklens, int ntrees,
                             int nnodes, int start_coord);
/* src/argweaver/local_tree.cpp:1827-1832 */
LocalTrees *arghmm_copy_trees(LocalTrees *trees);
/* src/argweaver/local_tree.cpp:1835-1838 */
int get_local_trees_ntrees(LocalTrees *trees);
/* src/argweaver/local_tree.cpp:1841-1844 */
int get_local_trees_nnodes(LocalTrees *trees);
/* src/argweaver/local_tree.cpp:1847-1892 */
void get_local_trees_ptrees(LocalTrees *trees, int **ptrees, int **ages, int **sprs,
                            int *blocklens);
/* src/argweaver/local_tree.cpp:1945-1958 */
void get_local_trees_blocks(LocalTrees *trees, int *starts, int *ends);
/* src/argweaver/local_tree.cpp:1934-1942 */
void get_treelens(LocalTrees *trees, double *times, int ntimes, double *treelens);
/* src/argweaver/local_tree.cpp:1895-1898 */
void delete_local_trees(LocalTrees *trees);

/*---------------------------------------------------------------------------
 * state spaces
 */
/* src/argweaver/states.cpp:209-226 */
void arghmm_get_nstates(LocalTrees *trees, int ntimes, bool internal, int *nstates);
/* src/argweaver/states.cpp:230-252 */
intstate **get_state_spaces(LocalTrees *trees, int ntimes, bool internal);
/* src/argweaver/states.cpp:256-261 */
void delete_state_spaces(intstate **all_states, int ntrees);

/*---------------------------------------------------------------------------
 * threading HMM
 */
/* src/argweaver/sample_thread.cpp:887-926.  Row-per-column table, freed with
 * delete_forward_matrix.  internal=true is unsupported (returns NULL). */
double **arghmm_forward_alg(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                            double rho, double mu, char **seqs, int nseqs, int seqlen,
                            bool prior_given, double *prior, bool internal, bool slow);
/* src/argweaver/sample_thread.cpp:1037-1042 */
void delete_forward_matrix(double **mat, int nrows);
/* src/argweaver/sample_thread.cpp:1032-1035 */
void delete_double_matrix(double **mat, int nrows);
/* src/argweaver/sample_thread.cpp:930-977 (path conversion quirk kept) */
intstate *arghmm_sample_posterior(int **ptrees, int **ages, int **sprs, int *blocklens, int ntrees,
                                  int nnodes, double *times, int ntimes, double *popsizes,
                                  double rho, double mu, char **seqs, int nseqs, int seqlen,
                                  intstate *path);
/* src/argweaver/sample_thread.cpp:1010-1023: threads sequence nseqs-1 */
LocalTrees *arghmm_sample_thread(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                                 double rho, double mu, char **seqs, int nseqs, int seqlen);
/* src/argweaver/sample_arg.cpp:665-678 */
LocalTrees *arghmm_sample_arg_seq(double *times, int ntimes, double *popsizes, double rho,
                                  double mu, char **seqs, int nseqs, int seqlen);
/* src/argweaver/sample_arg.cpp:682-697 */
LocalTrees *arghmm_sample_arg_refine(double *times, int ntimes, double *popsizes, double rho,
                                     double mu, char **seqs, int nseqs, int seqlen, int niters,
                                     int nremove);
/* src/argweaver/sample_arg.cpp:701-718 */
LocalTrees *arghmm_resample_arg(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                                double rho, double mu, char **seqs, int nseqs, int seqlen,
                                int niters, int nremove);
/* src/argweaver/sample_thread.cpp:1026-1029 */
void delete_path(int *path);
/* src/argweaver/logging.cpp:126-127 */
void setLogLevel(int level);

/* ARG probability (host): the log-likelihood of the sequences given the
 * local trees, the log prior of the trees and SPRs, their sum, and the log
 * prior of the first tree.  src/argweaver/total_prob.cpp:317-326 */
double arghmm_likelihood(LocalTrees *trees, double *times, int ntimes, double mu, char **seqs,
                         int nseqs, int seqlen);
/* src/argweaver/total_prob.cpp:329-342: aborts, as the reference's does */
double arghmm_likelihood_parsimony(LocalTrees *trees, double *times, int ntimes, double mu,
                                   char **seqs, int nseqs, int seqlen);
/* src/argweaver/total_prob.cpp:345-352 */
double arghmm_prior_prob(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                         double rho);
/* src/argweaver/total_prob.cpp:355-366 */
double arghmm_tree_prior_prob(LocalTrees *trees, double *times, int ntimes, double *popsizes);
/* src/argweaver/total_prob.cpp:369-378 */
double arghmm_joint_prob(LocalTrees *trees, double *times, int ntimes, double *popsizes,
                         double mu, double rho, char **seqs, int nseqs, int seqlen);

/*---------------------------------------------------------------------------
 * extensions (no counterpart in the reference)
 */
/* Remove the leaf holding sequence `seqid` from the ARG (the step
 * resample_arg performs before re-threading). */
LocalTrees *awamd_remove_thread(LocalTrees *trees, int seqid);

/* Forward (and traceback when `uniforms` is non-NULL) for sequence
 * `new_chrom` against the ARG, with an explicit uniform stream
 * (uniforms[0] for the last column).  precision: 0 = FP64 table,
 * 1 = FP32 table with per-column rescaling (log-normalisers, logL and the
 * traceback CDF stay FP64).  alpha_out: block-major table as FP64
 * (W doubles, see awamd_num_column_states), may be NULL.  ms_out[4]
 * receives prep/emission/forward/traceback kernel times.  Returns 0 on
 * success. */
int awamd_thread_forward(LocalTrees *trees, const double *times, int ntimes,
                         const double *popsizes, double rho, double mu, char **seqs,
                         int new_chrom, const double *uniforms, int precision,
                         double *alpha_out, int *path_out, double *lnl_out, double *ms_out);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  awamd_thread_forward with every device buffer taken
 * from a fixed 64 MiB workspace that each GPU holds for the life of the
 * process, instead of allocated: after the first call on a device (which
 * locates the workspace and creates the library's stream) no call
 * allocates device memory.  The same kernels and schedule as
 * awamd_thread_forward, so the same results.  Returns 1, with a message on
 * stderr, for an instance whose buffers need more than the workspace
 * (roughly seqlen * 64R * 8 bytes of forward table plus the per-block
 * planes; config-1-sized instances fit).  Every call on a GPU shares its
 * workspace: the library serialises calls in one process, and results are
 * copied out before a call returns. */
int awamd_thread_forward_noalloc(LocalTrees *trees, const double *times, int ntimes,
                                 const double *popsizes, double rho, double mu, char **seqs,
                                 int new_chrom, const double *uniforms, int precision,
                                 double *alpha_out, int *path_out, double *lnl_out,
                                 double *ms_out);

/* Column-states W = sum_b blocklen_b * S_b and, optionally, S_b per block. */
long long awamd_num_column_states(LocalTrees *trees, int ntimes, int *nstates_per_block);
/* Node mapping of every block (nnodes ints per block, -1 rows for block 0). */
int awamd_get_mappings(LocalTrees *trees, int *mappings);
/* leaf -> sequence row; returns the number of leaves */
int awamd_get_seqids(LocalTrees *trees, int *seqids);
/* 1 when a HIP device is visible */
int awamd_device_available(void);
/* one-line description of the current device (name, CUs, LDS sizes) */
int awamd_device_info(char *buf, int n);

/* Host steps of a thread insertion, exposed for callers that run the HMM
 * elsewhere: sample the recombination points of a sampled path
 * (recomb.cpp:151-235, draws from rand(); returns the count, at most
 * maxrecomb stored) and splice the thread into the ARG
 * (thread.cpp:272-418).  Recombination nodes of -1 mean the new branch. */
int awamd_sample_recombinations(LocalTrees *trees, const double *times, int ntimes,
                                const double *popsizes, double rho, double mu, const int *path,
                                int maxrecomb, int *recomb_pos, int *recomb_node,
                                int *recomb_time);
LocalTrees *awamd_add_thread(LocalTrees *trees, int ntimes, const int *path, int seqid,
                             int nrecomb, const int *recomb_pos, const int *recomb_node,
                             const int *recomb_time);

/* Batches of independent threading instances.  awamd_batch_add stages an
 * instance (the ARG is flattened at once; sequence rows and uniforms must
 * stay valid until awamd_batch_upload, which copies every input to HBM;
 * returns -1, adding nothing, after an upload or for an instance over the
 * kernels' limits).
 * awamd_batch_run executes classification, block preparation, emissions,
 * forward and traceback for all instances.  precision 0: FP64 forward table
 * (the drop-in arithmetic); precision 1: FP32 table and recursion with
 * per-column rescaling, FP64 normalisers, log-likelihood and traceback CDF.
 * The forward runs one wave per instance; the traceback runs in tiles of
 * tile_cols columns (0: one tile per instance) with fix-up passes.
 * stats[9]: prep, emit, forward, traceback and total ms, W, variant columns,
 * traceback fix-up passes and re-sampled columns. */
void *awamd_batch_new(void);
void awamd_batch_free(void *batch);
int awamd_batch_add(void *batch, LocalTrees *trees, const double *times, int ntimes,
                    const double *popsizes, double rho, double mu, char **seqs, int new_chrom,
                    const double *uniforms);
int awamd_batch_upload(void *batch);
int awamd_batch_run(void *batch, int precision, int tile_cols, int traceback, double *stats);
/* as awamd_batch_run, with the traceback fix-up switch (0: speculative tiles
 * only, for diagnostics) */
int awamd_batch_run_opts(void *batch, int precision, int tile_cols, int traceback, int fixup,
                         double *stats);
/* as awamd_batch_run_opts, with FP32 forward tiles: long instances run as
 * concurrent tiles of fw_tile columns (0: off; FP64 runs ignore it) scaled
 * by 0.94 us over the column cost of the instance's states per lane (so
 * that all tiles take about as long); a tile after the first starts fw_warm
 * columns early
 * (<= 0: 25000) from a uniform vector, its first stored column is checked
 * against the column computed by the tile before at relative tolerance
 * fw_tol (<= 0: 1e-5), and a tile that fails is re-run from that column.
 * stats[12]: the nine of awamd_batch_run, then forward tiles, re-run tiles
 * and the ms of the main forward launch. */
int awamd_batch_run_ex(void *batch, int precision, int tile_cols, int traceback, int fixup,
                       int fw_tile, int fw_warm, double fw_tol, double *stats);
/* Traceback-schedule autotuner: runs the forward once, times the traceback
 * for each candidate tile length (0: one tile per instance) and returns the
 * fastest (-1 on error); ms_out[ncand] receives the timings when given. */
int awamd_batch_tune(void *batch, int precision, const int *cand, int ncand, double *ms_out);
/* FP32 forward-tile autotuner: times the forward (checks and re-runs
 * included) for each candidate (tiles[i] columns, warms[i] warm-up columns;
 * tiles[i] = 0: one tile per instance) at tolerance tol and returns the
 * index of the fastest (-1 on error); ms_out[ncand] receives the timings. */
int awamd_batch_tune_forward(void *batch, const int *tiles, const int *warms, int ncand,
                             double tol, double *ms_out);
int awamd_batch_path(void *batch, int i, int *path);
int awamd_batch_alpha(void *batch, int i, double *alpha);
int awamd_batch_norms(void *batch, int i, double *norms);
double awamd_batch_lnl(void *batch, int i);
long long awamd_batch_W(void *batch, int i);

/* Seeded DSMC workload: k haplotypes over seqlen (compressed) columns with
 * per-column rho and mu.  Fills seqs[k*seqlen] row-major and returns the
 * ARG of all k leaves (NULL on failure). */
LocalTrees *awamd_simulate(int k, int seqlen, const double *times, int ntimes,
                           const double *popsizes, double rho, double mu,
                           unsigned long long seed, char *seqs);

/* Continuous-batching server (no counterpart in the reference).  One
 * driver thread owns the device work: whenever the GPU is free it takes the
 * queued requests, in arrival order, while their estimated device footprint
 * fits hbm_budget_gb (at most max_batch; a request larger than the budget
 * runs alone) and runs them as one batch, so requests that arrive while a
 * batch runs share the next launch.  submit copies the sequence rows and
 * uniforms and returns an id (-1 for a time grid of more than 32 points,
 * which is not queued); wait blocks for it (at most timeout_s seconds
 * when timeout_s > 0) and fills path[seqlen], the log-likelihood and
 * times3 = {submitted, started, finished} in seconds on the server clock.
 * It returns 0 on success, -2 on time-out (the request stays pending, wait
 * again), -1 for an unknown id, 3 for a request refused at submission
 * (over the kernels' node, time-point or states-per-block limits; it
 * fails alone and never joins a batch), else the failed batch's status.  A
 * request's result does not depend on the batch it ran in.  FP32 servers
 * run requests longer than 50000 columns as forward tiles (as
 * awamd_batch_run_ex with fw_tile 50000, fw_warm 20000, fw_tol 1e-4). */
void *awamd_server_new(double hbm_budget_gb, int precision, int tile_cols, int max_batch);
void awamd_server_free(void *server);
int awamd_server_submit(void *server, LocalTrees *trees, const double *times, int ntimes,
                        const double *popsizes, double rho, double mu, char **seqs,
                        int new_chrom, const double *uniforms);
int awamd_server_wait(void *server, int id, double timeout_s, int *path, double *lnl,
                      double *times3);
double awamd_server_now(void *server);

#ifdef __cplusplus
}
#endif
#endif
