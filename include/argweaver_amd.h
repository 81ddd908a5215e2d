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
 * the kernels' limits (256 nodes per tree, 32 time points, 2048 states per
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
 * delete_forward_matrix.  internal=true: the trees are in subtree/maintree
 * form and the removed subtree is the thread (the caller's prior is not
 * used then; the reference's internal prior is). */
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
/* src/argweaver/sample_thread.cpp:981-1006: forward and traceback of the
 * subtree of an ARG whose trees are in subtree/maintree form (after
 * arghmm_remove_arg_thread_path); thread_path receives one internal state
 * index per column.  The ARG is not changed. */
void arghmm_sample_arg_thread_internal(LocalTrees *trees, double *times, int ntimes,
                                       double *popsizes, double rho, double mu, char **seqs,
                                       int nseqs, int seqlen, int *thread_path);
/* src/argweaver/sample_arg.cpp:722-739: niters re-threadings of a random
 * internal branch (resample_arg_all, :111-127) */
LocalTrees *arghmm_resample_all_arg(LocalTrees *trees, double *times, int ntimes,
                                    double *popsizes, double rho, double mu, char **seqs,
                                    int nseqs, int seqlen, int niters, double prob_path_switch);
/* src/argweaver/sample_arg.cpp:790-806: niters re-threadings along removal
 * paths weighted towards recombinations (resample_arg_climb, :211-230) */
LocalTrees *arghmm_resample_climb_arg(LocalTrees *trees, double *times, int ntimes,
                                      double *popsizes, double rho, double mu, char **seqs,
                                      int nseqs, int seqlen, int niters, double recomb_preference);
/* src/argweaver/sample_arg.cpp:828-842 (resample_arg_region, :429-527):
 * niters re-threadings of a uniformly drawn branch over the columns
 * [region_start, region_end), each conditioned on the states the branch
 * held at the region's ends (open at an end of the ARG) and accepted with
 * probability exp(npaths - npaths2); NULL on failure */
LocalTrees *arghmm_resample_arg_region(LocalTrees *trees, double *times, int ntimes,
                                       double *popsizes, double rho, double mu, char **seqs,
                                       int nseqs, int seqlen, int region_start, int region_end,
                                       int niters);
/* src/argweaver/sample_arg.cpp:743-765 (resample_arg_mcmc_all, :190-203):
 * per iteration a leaf's re-threading (probability 1/2) or the region
 * re-threading over windows of `window` columns every window / 2, niters2
 * re-threadings each */
LocalTrees *arghmm_resample_mcmc_arg(LocalTrees *trees, double *times, int ntimes,
                                     double *popsizes, double rho, double mu, char **seqs,
                                     int nseqs, int seqlen, int niters, int niters2, int window);
/* src/argweaver/sample_thread.cpp:1026-1029 */
void delete_path(int *path);
/* src/argweaver/logging.cpp:126-127 */
void setLogLevel(int level);

/* ARG probability (host): the log-likelihood of the sequences given the
 * local trees, the log prior of the trees and SPRs, their sum, and the log
 * prior of the first tree.  src/argweaver/total_prob.cpp:317-326 */
double arghmm_likelihood(LocalTrees *trees, double *times, int ntimes, double mu, char **seqs,
                         int nseqs, int seqlen);
/* log P(sequences | ARG) computed on the device (no counterpart in the
 * reference; the same terms as arghmm_likelihood, calc_arg_likelihood,
 * total_prob.cpp:19-45, summed in another order: within 1e-11 relative).
 * NAN without a HIP device or for an ARG over the kernels' limits. */
double awamd_arg_likelihood(LocalTrees *trees, const double *times, int ntimes, double mu,
                            char **seqs, int nseqs, int seqlen);
/* log P(ARG) computed on the device (no counterpart in the reference; the
 * terms of arghmm_prior_prob, calc_arg_prior, total_prob.cpp:262-299, one
 * per local tree, summed in another order: within 1e-11 relative).  NAN
 * without a HIP device. */
double awamd_arg_prior(LocalTrees *trees, const double *times, int ntimes,
                       const double *popsizes, double rho);
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
 * dense matrices, their brute-force checks and the generic HMM (host; not on
 * the threading path, kept so the C ABI is complete).  Matrices are arrays
 * of separately allocated rows, freed by the matching delete_* call.
 */
/* src/argweaver/trans.cpp:1190-1208: log P(i -> j) within a block; the tree
 * length, time steps and lineage counts are recomputed from the tree */
double **new_transition_probs(int nnodes, int *ptree, int *ages, double treelen,
                              intstate *states, int nstates, int ntimes, double *times,
                              double *time_steps, int *nbranches, int *nrecombs, int *ncoals,
                              double *popsizes, double rho);
/* src/argweaver/trans.cpp:1211-1244: log P(i -> j) across the SPR between
 * last_ptree and ptree (states of the two trees in node-major order) */
double **new_transition_probs_switch(int *ptree, int *last_ptree, int nnodes, int recomb_name,
                                     int recomb_time, int coal_name, int coal_time,
                                     int *ages_index, int *last_ages_index, double treelen,
                                     double last_treelen, intstate *states1, int nstates1,
                                     intstate *states2, int nstates2, int ntimes, double *times,
                                     double *time_steps, int *nbranches, int *nrecombs,
                                     int *ncoals, double *popsizes, double rho);
/* src/argweaver/trans.cpp:1248-1251 */
void delete_transition_probs(double **transmat, int nstates);
/* src/argweaver/trans.cpp:1254-1271 (assert_transmat, :856-943) */
bool arghmm_assert_transmat(int nnodes, int *ptree, int *ages, int ntimes, double *times,
                            double *popsizes, double rho);
/* src/argweaver/trans.cpp:1275-1315 (assert_transmat_switch, :947-1032) */
bool arghmm_assert_transmat_switch(int nnodes, int *ptree, int *ages, int recomb_node,
                                   int recomb_time, int coal_node, int coal_time, int ntimes,
                                   double *times, double *popsizes, double rho);
/* src/argweaver/trans.cpp:1314-1343 (assert_transmat_internal, :1035-1121):
 * a branch drawn with irand lifted above the root, then the within-block
 * matrix of the internal states checked entry by entry */
bool arghmm_assert_transmat_internal(int nnodes, int *ptree, int *ages, int ntimes,
                                     double *times, double *popsizes, double rho);
/* src/argweaver/trans.cpp:1346-1379 (assert_transmat_switch_internal,
 * :1124-1186): a branch removed along a path from the first tree (trees is
 * left in subtree/maintree form), then every switch matrix checked */
bool arghmm_assert_transmat_switch_internal(LocalTrees *trees, int ntimes, double *times,
                                            double *popsizes, double rho);
/* src/argweaver/emit.cpp:1288-1303: emit[seqlen][nstates] for the new
 * sequence seqs[nleaves] against one tree */
double **new_emissions(intstate *states, int nstates, int *ptree, int nnodes, int *ages_index,
                       char **seqs, int nseqs, int seqlen, double *times, int ntimes, double mu);
/* src/argweaver/emit.cpp:1306-1309 */
void delete_emissions(double **emit, int seqlen);
/* src/argweaver/emit.cpp:1312-1337 (assert_emissions, :1222-1250) */
bool arghmm_assert_emit(LocalTrees *trees, int ntimes, double *times, double mu, char **seqs,
                        int nseqs, int seqlen);
/* src/argweaver/emit.cpp:1340-1380 (assert_emissions_internal,
 * :1253-1280): a uniform removal path applied to a copy of trees, then the
 * internal-state emissions of every block against the full likelihood */
bool arghmm_assert_emit_internal(LocalTrees *trees, int ntimes, double *times, double mu,
                                 char **seqs, int nseqs, int seqlen);
/* src/argweaver/hmm.cpp:27-44: log-space forward; fw[0] is the caller's */
void forward_alg(int n, int nstates, double **trans, double **emit, double **fw);
/* src/argweaver/hmm.cpp:48-64: log-space backward; bw[n-1] is the caller's */
void backward_alg(int n, int nstates, double **trans, double **emit, double **bw);
/* src/argweaver/hmm.cpp:67-84: path[n-1] is the caller's; draws rand() */
void sample_hmm_posterior(int n, int nstates, double **trans, double **fw, int *path);
/* src/argweaver/est_popsize.cpp:140-147: one tree every `step` bases;
 * popsizes[0..ntimes-2] */
void arghmm_est_popsizes_trees(LocalTrees *trees, double *times, int ntimes, int step,
                               double *popsizes);

/*---------------------------------------------------------------------------
 * internal-branch re-threading: removal paths and the branch surgery (host)
 */
/* src/argweaver/thread.cpp:1681-1684: a removal path (one node per local
 * tree) from `node` of the first tree, forward through the SPRs */
void arghmm_sample_arg_removal_path(LocalTrees *trees, int node, int *path);
/* src/argweaver/thread.cpp:1686-1690: a removal path through `node` of the
 * tree at position pos, forward and backward */
void arghmm_sample_arg_removal_path2(LocalTrees *trees, int node, int pos, int *path);
/* src/argweaver/thread.cpp:1692-1696: the path of a leaf */
void arghmm_sample_arg_removal_leaf_path(LocalTrees *trees, int node, int *path);
/* src/argweaver/thread.cpp:1698-1703: preferring branches that carry
 * recombinations */
void arghmm_sample_arg_removal_path_recomb(LocalTrees *trees, double recomb_preference,
                                           int *path);
/* src/argweaver/thread.cpp:1705-1709: remove the branch along the path,
 * leaving each tree in subtree/maintree form under a root of age maxtime */
void arghmm_remove_arg_thread_path(LocalTrees *trees, int *removal_path, int maxtime);
/* src/argweaver/thread.cpp:1711-1716: as above, with the removed branch as
 * internal state indices per column */
void arghmm_remove_arg_thread_path2(LocalTrees *trees, int *removal_path, int maxtime,
                                    int *original_thread);
/* src/argweaver/thread.cpp:1719-1735: the time of each column's internal
 * state */
void arghmm_get_thread_times(LocalTrees *trees, int ntimes, int *path, int *path_times);

/*---------------------------------------------------------------------------
 * on-disk formats (host)
 */
/* src/argweaver/local_tree.cpp:1904-1931: .smc (plain or gzip); NULL on a
 * malformed file */
LocalTrees *read_local_trees(const char *filename, const double *times, int ntimes);
/* src/argweaver/local_tree.cpp:1897-1901 (writer :1286-1341); names may be
 * NULL (no NAMES line) */
void write_local_trees(char *filename, LocalTrees *trees, char **names, double *times,
                       int ntimes);
/* src/argweaver/sequences.cpp:783-793: .sites alignment (plain or gzip);
 * subregion bounds of -1 keep the file's REGION; NULL on a malformed file */
void *arghmm_read_sites(const char *filename, int subregion_start, int subregion_end);
/* src/argweaver/sequences.cpp:796-799 */
void arghmm_delete_sites(void *sites);

/*---------------------------------------------------------------------------
 * extensions (no counterpart in the reference)
 */
/* A sites handle: shape4 = {nseqs, nsites, start_coord, end_coord}; the
 * name of sequence i; positions[nsites] and the columns as nsites rows of
 * nseqs bases (either may be NULL); the alignment as nseqs rows of
 * end_coord - start_coord bases with default_char off the variant sites
 * (make_sequences_from_sites, sequences.cpp:315-343). */
int awamd_sites_shape(void *sites, int *shape4);
const char *awamd_sites_name(void *sites, int i);
int awamd_sites_data(void *sites, int *positions, char *cols);
int awamd_sites_sequences(void *sites, char default_char, char *seqs);
/* Compression (find_compress_cols + compress_sites, sequences.cpp:523-611):
 * maps every `compress` bases to one column, keeping every variant site, in
 * place; returns the mapping (NULL when not possible), which also
 * (un)compresses the block lengths of local trees (local_tree.cpp:790-848)
 * and restores the sites. */
void *awamd_sites_compress(void *sites, int compress);
int awamd_sites_uncompress(void *sites, void *mapping);
int awamd_compress_local_trees(LocalTrees *trees, void *mapping);
int awamd_uncompress_local_trees(LocalTrees *trees, void *mapping);
void awamd_delete_sites_mapping(void *mapping);

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
/* Empty the batch so that it takes new instances (its device buffers are
 * kept and reused: a loop that re-fills one batch allocates only when it
 * grows).  Returns 0. */
int awamd_batch_clear(void *batch);
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
 * stats[14]: the nine of awamd_batch_run, then forward tiles, re-run tiles,
 * the ms of the main forward launch, and, when forward tiles were re-run,
 * the ms from the forward's end to the end of the re-run instances'
 * traceback and of the other instances' traceback (0 otherwise). */
int awamd_batch_run_ex(void *batch, int precision, int tile_cols, int traceback, int fixup,
                       int fw_tile, int fw_warm, double fw_tol, double *stats);
/* Traceback tiles of later runs: warm-up columns of a tile's speculative
 * chain (default 256) and options per tile, 1 .. 4 (default 3: the
 * speculative chain and chains from the neighbouring states of its branch;
 * a tile no option fits is re-run).  The sampled thread is the same for
 * every setting.  0 on success, -1 for a value out of range. */
int awamd_batch_set_traceback(void *batch, int warm, int options);
/* Traceback tile length of the instances whose FP32 forward tiles are
 * re-run (their traceback runs after the forward: shorter tiles spread it
 * over more of the device); 0: the run's tile length.  The thread does not
 * depend on it.  -1 for a negative length. */
int awamd_batch_set_tail_tiles(void *batch, int cols);
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
/* columns [c0, c1) of instance i's table, as awamd_batch_alpha lays them out
 * (the states of each column's block in state order); 1 for a bad range */
int awamd_batch_alpha_range(void *batch, int i, int c0, int c1, double *alpha);
int awamd_batch_norms(void *batch, int i, double *norms);
double awamd_batch_lnl(void *batch, int i);
long long awamd_batch_W(void *batch, int i);

/* Seeded DSMC workload: k haplotypes over seqlen (compressed) columns with
 * per-column rho and mu.  Fills seqs[k*seqlen] row-major and returns the
 * ARG of all k leaves (NULL on failure). */
LocalTrees *awamd_simulate(int k, int seqlen, const double *times, int ntimes,
                           const double *popsizes, double rho, double mu,
                           unsigned long long seed, char *seqs);

/* Continuous-batching server (no counterpart in the reference).  Driver
 * threads (two by default) own the device work, each with a batch of its
 * own: a free driver takes the queued requests, in arrival order, while
 * their estimated device footprint fits its share of hbm_budget_gb (at most
 * max_batch; a request larger than the share runs alone) and runs them as
 * one batch, so requests that arrive while a batch is on the GPU are
 * admitted by the other driver and run beside it.  submit copies the sequence rows and
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
/* the same with the number of driver threads, 1 .. 4 (NULL outside) */
void *awamd_server_new_ex(double hbm_budget_gb, int precision, int tile_cols, int max_batch,
                          int drivers);
void awamd_server_free(void *server);
int awamd_server_submit(void *server, LocalTrees *trees, const double *times, int ntimes,
                        const double *popsizes, double rho, double mu, char **seqs,
                        int new_chrom, const double *uniforms);
int awamd_server_wait(void *server, int id, double timeout_s, int *path, double *lnl,
                      double *times3);
/* awamd_server_wait with info5 = {submitted, started, finished, requests in
 * its batch, driver index} */
int awamd_server_wait_ex(void *server, int id, double timeout_s, int *path, double *lnl,
                         double *info5);
double awamd_server_now(void *server);

#ifdef __cplusplus
}
#endif
#endif
