/*
 * spaqr_gpu.h -- C ABI of the MI355X (gfx950) spaQR hot path.
 *
 * Two layers, both plain C (pointers and sizes only):
 *
 *  1. The ParTree facade: the host API of the reference's parallel tree
 *     (include/ptree.h:123-159, include/tree.h:137-143, src/is.cpp:447-527), so a
 *     caller that drove ParTree + dist_gmres drives this instead.  Vectors passed to
 *     solve / spmv / gmres are DEVICE pointers (FP64) in the tree's permuted ordering,
 *     the ordering the reference's solve(x) / spmv(x) / dist_gmres work in.
 *
 *  2. Grouped batched kernels: LAPACK-semantics routines the reference calls per block
 *     (inventory of call sites: geqrt src/ptree.cpp:187, tpqrt :206, tpmqrt :248,
 *     larfb src/util.cpp:468-493, trsm src/util.cpp:233-241, geqp3 + choose_rank
 *     src/util.cpp:501-512,830-851, larft :403-411).  Each call takes an array of task
 *     descriptors in HOST memory whose matrix pointers are DEVICE pointers (column-major,
 *     explicit leading dimensions); all tasks run in one launch, one workgroup per task.
 *     These calls copy the descriptors, launch on `stream` and synchronize it.
 *
 * Precision: `prec` is 64 (double) or 32 (float) for every matrix pointer of a call.
 * T factors are stored in 32-column blocks: T is 32 x k (ld 32), one 32 x 32 upper
 * triangular block per panel of 32 reflectors (LAPACK dgeqrt / dtpqrt storage, NB = 32).
 *
 * Return values: 0 on success, < 0 on a bad argument (LAPACK info convention: -i names
 * the offending argument), -100 on a HIP runtime failure; spaqr_last_error() describes it.
 */
#ifndef SPAQR_GPU_H
#define SPAQR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ ParTree facade */
typedef struct spaqr_tree spaqr_tree;

typedef struct {
    int iters;        /* as dist_gmres returns: 0-based k of the converged step (summed over restarts) */
    int cycles;       /* restart cycles run (1 without restart) */
    double prec_res;  /* final preconditioned residual estimate |beta_{k+1}| / ||M^-1 r0|| */
    double true_res;  /* ||b - A x|| / ||b|| computed in FP64 on the device */
} spaqr_gmres_result;

/* ParTree(int nlevels, int skip)                                     include/ptree.h:135 */
spaqr_tree* spaqr_tree_create(int nlevels, int skip);
void spaqr_tree_destroy(spaqr_tree* t);
/* Tree::set_tol / set_skip / set_scale / set_square / set_order       src/tree.cpp:14-19 */
int spaqr_tree_set_tol(spaqr_tree* t, double tol);
int spaqr_tree_set_skip(spaqr_tree* t, int skip);
int spaqr_tree_set_scale(spaqr_tree* t, int scale);
int spaqr_tree_set_square(spaqr_tree* t, int square);
int spaqr_tree_set_order(spaqr_tree* t, float order);
/* ParTree::set_verbose                                                  src/ptree.cpp:22 */
int spaqr_tree_set_verbose(spaqr_tree* t, int verbose);
/* Tree::set_Xcoo: coordinates X (d x n, column-major)                    src/tree.cpp:20 */
int spaqr_tree_set_Xcoo(spaqr_tree* t, const double* X, int d, int n);
/* Orderings without coordinates (the reference's METIS path, src/ptree.cpp:503-518 ->
   MetisPartition, src/partition.cpp:354-424).  With no coordinates, partition() orders by
   leaf parts: one per column in [0, 2^(levels-1)), numbered as a recursive bisection numbers
   them (part k of depth d splits into 2k and 2k + 1), e.g. the output of
   METIS_PartGraphRecursive on A^T A with 2^(levels-1) parts.  set_parts gives them (n = 0
   clears); without them the library bisects the graph of A^T A itself (level structures from
   a pseudo-peripheral vertex, component by component; not METIS's partition, so not its
   ordering).  Separators and interfaces then follow the METIS path's rules (src/partition.cpp:
   380-424, getInterfacesUsingATA :637-682).  Coordinates, when set, take precedence (as in the
   reference).  get_parts copies up to cap leaf parts of the last partition (natural order, every
   path) and returns their number. */
int spaqr_tree_set_parts(spaqr_tree* t, const int* parts, int n);
int spaqr_tree_get_parts(const spaqr_tree* t, int* out, int cap);
/* Factor precision: 64 (reference FP64 path) or 32 (FP32 fronts; GMRES stays FP64).
   No counterpart in the reference.  Set before spaqr_tree_assemble. */
int spaqr_tree_set_precision(spaqr_tree* t, int bits);
/* ParTree::partition(SpMat& A): A is nrows x ncols CSC (host), sorted row indices, nrows >= ncols
   (nrows > ncols: a least-squares system, below).                          src/ptree.cpp:495 */
int spaqr_tree_partition(spaqr_tree* t, int nrows, int ncols, const int* colptr, const int* rowind,
                         const double* val);
/* A batch of independent square systems factored together (no reference counterpart: the
   reference factors one system per run).  A (n x n CSC, host) is their block-diagonal matrix,
   system s owning unknowns [off_s, off_s + sizes[s]) in order; the coordinates (set_Xcoo) cover
   all n unknowns.  Each system gets its own nested dissection (as spaqr_tree_partition) with
   levels[s] levels; the tree must have been created with the largest of them.  assemble,
   factorize, solve and spmv then act on all systems at once, every tree level's phases running
   as one set of grouped launches over the systems' blocks of uneven sizes.  The permutation
   keeps each system's unknowns in its own range. */
int spaqr_tree_partition_batch(spaqr_tree* t, int nsys, const int* sizes, const int* levels, int n,
                               const int* colptr, const int* rowind, const double* val);
int spaqr_tree_num_systems(const spaqr_tree* t);
/* Planner tunables (no reference counterpart; the reference fixes nb = min(32, k) at
   src/ptree.cpp:203-206 and its laqps block at :407).  value 0 restores the built-in choice.
     0 big-front threshold: min(m, n) from which a front's panels run one launch per panel
       (built in 96; 16 .. 2^30)
     1 trailing-update chain column tile (64; 16, 32, 64)
     2 Q^T C column chunk (128; 16, 32, 64, 128 -- FP32 blocks over 512 rows use at most 64)
     3 C R^{-1} row chunk (128 FP32 / 256 FP64; 32, 64, 128, 256)
     4 sparsify right-transform row chunk (128 FP32 / 256 FP64; 32, 64, 128, 256)
     5 sparsify left-transform column chunk (128; 32, 64, 128)
     6 rank-path column-norm / tail-test column chunk (128; 32, 64, 128, 256)
     7 sketch width from which the sketch QR runs unpivoted (96; 16 .. 2^30)
     8 multi-panel coupling chains on the cooperative kernel (1 on, 2 off)
     9 blocks of at most 64 x 64 on the one-wave kernels (1 on, 2 off)
    10 FP32 fronts: rank-path sketches C Omega in bf16 with FP32 accumulation (1, built in) or
       in FP32 (2)
    11 output-tile columns of the bf16 sketch (128; 64, 128, 256)
    12 rank mode 1's sketch width s = ceil(grow/100 x predicted rank) + pad: grow (125; 100 .. 200)
    13 ... and pad + 1 (built in 8, i.e. 9; 1 .. 65).  Any width is safe: an interface whose
       sketch fails the tail test is retried wider, then falls back to geqp3
    14 column chunk of the trailing update in a big front's panel steps (128; 32, 64, 128, 256)
   The committed autotuner (python -m pspaqr_public_amd.autotune) picks them per workload. */
int spaqr_tree_set_tuning(spaqr_tree* t, int param, int value);
int spaqr_tree_get_tuning(const spaqr_tree* t, int param);
/* Numerical status (the per-block info[] of the factorization).  After each level's
   elimination a check reads the diagonal of every interior's R: a non-finite entry (flag 1) or a
   pivot at or below rel * max |R_ii| (flag 2; rel 0, the default, flags exact zeros only) is
   read back at the level's next host synchronisation and spaqr_tree_factorize returns 1 (a
   numerical failure, in the LAPACK info convention) naming the level and cluster.  The
   reference asserts on a NaN residual (src/is.cpp:460) and exits on an empty interior
   (src/ptree.cpp:171-175).  get_info copies up to cap flags, one per cluster (rows of
   spaqr_tree_get_clusters), also flag 4 for an interface whose sketch failed the tail test and
   took geqp3's rank; returns the number of clusters.  spaqr_tree_partition rejects non-finite
   entries and zero columns before any GPU work. */
int spaqr_tree_set_pivot_check(spaqr_tree* t, double rel);
int spaqr_tree_get_info(spaqr_tree* t, int* out, int cap);
/* Orthogonalisation of spaqr_gmres (src/is.cpp:447-527 uses modified Gram-Schmidt with the
   Givens rotations on the host).  0: that scheme, one dot and one axpy launch per basis
   vector.  1 (default): classical Gram-Schmidt applied twice in fused passes (eight launches
   per iteration at any basis size) with the Givens rotations, the least-squares solve and the
   update of x on the device; one residual estimate is read back per iteration.  Both give the
   same Hessenberg matrix in exact arithmetic and the same stopping test. */
int spaqr_tree_set_krylov_mode(spaqr_tree* t, int mode);
/* hipGraph mode (no reference counterpart; it replaces the per-level task flows of
   src/ptree.cpp:885-2123 being re-issued on every factorization).  With on != 0, a
   factorization records every stretch of launches between two host synchronisations (the
   per-level rank read-backs) as a graph -- after one ordinary factorization that sizes the
   device arenas -- and launches it; later factorizations of the same assembled matrix with the
   same settings replay the graphs without host planning, read the ranks back at the same
   points and plan afresh if any differs.  assemble() and any setting change drop the plan.
   Needs rank mode 0.  info (3 long longs): graph mode on, recorded segments, replays so far. */
int spaqr_tree_set_graphs(spaqr_tree* t, int on);
int spaqr_tree_graph_info(const spaqr_tree* t, long long* info);
/* per system: first position, size (2 ints) */
int spaqr_tree_get_systems(const spaqr_tree* t, int* out);
/* system of each cluster (rows of spaqr_tree_get_clusters) */
int spaqr_tree_get_cluster_systems(const spaqr_tree* t, int* out);
/* GMRES on every system of a batch (spaqr_gmres, or spaqr_fgmres with flexible != 0, per
   system: its own Krylov space, stopping test and counts) with shared SpMV and preconditioner
   launches; out gets one result per system. */
int spaqr_gmres_batch(spaqr_tree* t, const double* b_dev, double* x_dev, int iters, double tol, int restart,
                      double true_tol, int flexible, spaqr_gmres_result* out);
/* ParTree::assemble(SpMat& A): builds the blocks and uploads them to HBM. src/ptree.cpp:696 */
int spaqr_tree_assemble(spaqr_tree* t);
/* ParTree::factorize()                                                 src/ptree.cpp:873 */
int spaqr_tree_factorize(spaqr_tree* t);
/* ParTree::solve(VectorXd& x): x <- M^{-1} x, in place (device)       src/ptree.cpp:2788 */
int spaqr_tree_solve(spaqr_tree* t, double* x_dev);
/* The preconditioner apply in phases, for callers that replay it as a list of operations (the
   fwd()/bwd() objects of include/operations.h:13-19 and include/toperations.h:14-20):
   solve_load copies x (device, FP64, tree ordering) into the solve workspace; solve_phase runs
   phase `phase` of level `level` forward (dir 1) or backward (dir 0) -- 0 merge
   (src/ptree.cpp:2244-2252 / 3105-3123), 1 interior QR (Q^T and the coupling chains
   :2846-2851, 2160-2186 / triangular solve with the off-diagonal reduction :2188-2192),
   2 sparsification (:2957-2963 / 3146-3174); solve_store copies the workspace back to x.
   Forward over levels 0.. and phases 0, 1, 2, then backward over levels ..0 and phases 2, 1, 0
   is spaqr_tree_solve, kernel for kernel. */
int spaqr_tree_solve_load(spaqr_tree* t, const double* x_dev);
int spaqr_tree_solve_phase(spaqr_tree* t, int level, int phase, int dir);
int spaqr_tree_solve_store(spaqr_tree* t, double* x_dev);
/* Host vectors between the natural and the tree ordering: to_natural != 0: out[cperm[k]] = in[k]
   (the cperm.asPermutation() * x of src/ptree.cpp:2783); else out[k] = in[cperm[k]]
   (its transpose, :2272).                                                                   */
int spaqr_tree_permute(const spaqr_tree* t, int to_natural, const double* in, double* out);
/* ParTree::spmv(VectorXd x): y = A x (device)                         src/ptree.cpp:3212 */
int spaqr_tree_spmv(spaqr_tree* t, const double* x_dev, double* y_dev);
/* dist_gmres(A, b, M, iter, tol, verb=true): x solves A x = b (device vectors, permuted
   ordering).  restart <= 0 is the reference's single cycle; restart > 0 restarts from the
   current iterate until ||b - A x|| / ||b|| <= true_tol or iters steps.  hist (host) gets
   the per-step preconditioned residual estimates.                         src/is.cpp:447 */
int spaqr_gmres(spaqr_tree* t, const double* b_dev, double* x_dev, int iters, double tol, int restart,
                double true_tol, spaqr_gmres_result* out, double* hist, int hist_cap, int* hist_len);
/* Flexible (right-preconditioned) GMRES with the same arguments: z_k = M^{-1} q_k is stored
   and the recurrence estimates the true residual ||b - A x|| / ||b||, which is what tol and
   the hist entries refer to here.  For FP32 factors, whose rounding makes M^{-1} slightly
   non-linear.  No counterpart in the reference (its dist_gmres is left-preconditioned,
   src/is.cpp:447). */
int spaqr_fgmres(spaqr_tree* t, const double* b_dev, double* x_dev, int iters, double tol, int restart,
                 double true_tol, spaqr_gmres_result* out, double* hist, int hist_cap, int* hist_len);
/* Tree::rows() / cols() and ParTree::nrows_local / ncols_local (one rank)  src/tree.cpp:23 */
int spaqr_tree_n(const spaqr_tree* t);
int spaqr_tree_m(const spaqr_tree* t);
/* The per-matrix symbolic phase -- ParTree::partition and ParTree::assemble (src/ptree.cpp:
   495-864): the geometric nested dissection, initPermutation, the block extraction, the SpMV
   storage and the distance-two connections -- on the device (on != 0) for square systems with
   coordinates; other systems keep the host phase.  The result is the host phase's: the same
   permutation, clusters, blocks and SpMV storage.  Set before partition.  symbolic_ms: the
   last partition's and assemble's wall time in ms (2 doubles). */
int spaqr_tree_set_device_symbolic(spaqr_tree* t, int on);
int spaqr_tree_symbolic_ms(const spaqr_tree* t, double* out);

/* ---- Rectangular systems (least squares, m > n).  spaqr_tree_partition accepts nrows > ncols
   without coordinates (leaf parts or the graph bisection; the reference's geometric path assigns
   rows by HSL's bipartite matching, src/partition.cpp:103-134, absent here): each row is filed
   under a cluster by form_rmap (src/tree.cpp:102-165) and the factorization reassigns the extra
   rows of every eliminated interior (src/ptree.cpp:1149-1356) and compresses them
   (sparsify_extra :441-493).  GMRES and spaqr_tree_solve are square-only. */
/* row permutation: natural index of permuted row i (include/tree.h:53)                        */
int spaqr_tree_get_rperm(const spaqr_tree* t, int* out);
/* host vectors of m entries between the natural and the permuted row ordering (as permute) */
int spaqr_tree_permute_rows(const spaqr_tree* t, int to_natural, const double* in, double* out);
/* per cluster, 5 ints: first permuted row, rows at creation, extra rows received, rows kept
   beyond the columns (rows - cols; after sparsify_extra its rank), an interior's extra rows that
   went to a neighbour (include/cluster.h:153, 224-233)                                        */
int spaqr_tree_get_rows(const spaqr_tree* t, int* out);
/* ParTree::solve(VectorXd b, VectorXd& x): b (m, permuted rows) -> x (n, permuted columns),
   device vectors.  For a square system it is spaqr_tree_solve out of place.  src/ptree.cpp:2267 */
int spaqr_tree_solve_ls(spaqr_tree* t, const double* b_dev, double* x_dev);
/* solve_nrml (src/tree.cpp:1524-1570, the sequential Tree's; ParTree has none): x = M^{-1}
   M^{-T} y with M^{-1} the backward half of the solve, the transposed operations of
   include/toperations.h:14-195 on the way up -- an approximate (A^T A)^{-1} y.  n-vectors,
   permuted columns, device.                                                                  */
int spaqr_tree_solve_nrml(spaqr_tree* t, const double* y_dev, double* x_dev);
/* y = A^T r (r m entries, y n; permuted orderings, device): mat.adjoint() in src/is.cpp:294-321 */
int spaqr_tree_spmv_t(spaqr_tree* t, const double* r_dev, double* y_dev);
/* cgls(A, b, x, M, iters, tol) with M = the tree's solve_nrml and x0 = 0: CG on the normal
   equations until ||A^T (b - A x)|| < tol ||A^T b||.  *iters_out is the reference's count,
   *rel = ||A^T r|| / ||A^T b||.  b (m) and x (n) device, permuted orderings.  src/is.cpp:274-345 */
int spaqr_cgls(spaqr_tree* t, const double* b_dev, double* x_dev, int iters, double tol, int* iters_out,
               double* rel);
/* column permutation (cperm, natural index of permuted position i)        include/tree.h:53 */
int spaqr_tree_get_cperm(const spaqr_tree* t, int* out);
/* cluster table, 13 ints per cluster in creation order: merge level, self (lvl, sep), l (lvl,
   sep), r (lvl, sep), section, part, cstart, cols at creation, parent order, rank (-1 if the
   cluster was not sparsified)                                            include/cluster.h:142 */
int spaqr_tree_num_clusters(const spaqr_tree* t);
int spaqr_tree_get_clusters(const spaqr_tree* t, int* out);
/* per-level statistics, 8 doubles per level: interiors, interfaces, sparsified, max front,
   colour layers, sum of k before, sum of ranks after, algorithmic flops              */
int spaqr_tree_level_stats(const spaqr_tree* t, double* out, int max_levels);
/* upper triangle of the factored diagonal block of an interior (k x k col-major); returns k */
int spaqr_tree_diag_R(spaqr_tree* t, int order, double* out, int maxn);
double spaqr_tree_factor_ms(const spaqr_tree* t);
/* the HIP stream (hipStream_t) every call of this tree is ordered on */
void* spaqr_tree_stream(spaqr_tree* t);
const char* spaqr_last_error(const spaqr_tree* t);
/* wait for every call issued on this tree's stream */
int spaqr_tree_sync(spaqr_tree* t);
/* Kernel-class timing with HIP events on the tree's stream (no reference counterpart).
   Classes, in order: 0 geqrt (interior diagonal QR), 1 larfb (Q^T on interior rows),
   2 tpqrt chains, 3 tpmqrt chains (trailing updates), 4 scaling QR, 5 scaling apply,
   6 RRQR gather, 7 geqp3, 8 sparsify apply + merge, 9 fills, 10 preconditioner solve,
   11 SpMV, 12 GMRES orthogonalisation, 13 sketch / features / predictor / tail test of the
   rank path.  Per class 5 doubles: milliseconds, launches,
   tasks, algorithmic flops, algorithmic bytes.  Timings are folded in at each sync. */
int spaqr_tree_set_profiling(spaqr_tree* t, int on);
int spaqr_tree_profile(spaqr_tree* t, double* out, int max_classes);
int spaqr_tree_profile_reset(spaqr_tree* t);

/* ------------------------------------------------------------------ rank path
   The sparsification rank of each interface (src/ptree.cpp:373-439).  Mode 0 (default) is the
   reference's geqp3 + choose_rank on the gathered C.  Mode 1 replaces it by a pivoted QR of a
   Gaussian sketch C*Omega whose width comes from a learned rank predictor (MLP 32 -> 64 -> 64
   -> 1); the chosen rank is the smallest whose tail passes the test geqp3 stops on
   (every column of (Q^T C)[r:, :] has norm <= tol * max column norm of C), and an interface
   whose sketch fails falls back to geqp3 on C.  Mode 2 is mode 1 with the sketch as wide as
   min(k, w) (no predictor).  The closest reference analogue is the unused adaptive range
   finder random_rrqr (src/util.cpp:643-715); none of these calls has a reference counterpart. */
int spaqr_tree_set_rank_mode(spaqr_tree* t, int mode);
/* predictor weights (host floats, spaqr_mlp_params() of them):
   W1 (64 x 32, row-major) b1 (64) W2 (64 x 64) b2 (64) w3 (64) b3 */
int spaqr_tree_set_rank_model(spaqr_tree* t, const float* w, int nparams);
/* record (32 features, rank) of every sparsified interface wide enough to sketch */
int spaqr_tree_set_rank_log(spaqr_tree* t, int on);
/* copies up to cap entries: x (32 floats each), meta (5 ints each: rank, k, w, level, mode
   used); returns the number of entries recorded */
int spaqr_tree_rank_log(const spaqr_tree* t, float* x, int* meta, int cap);
int spaqr_tree_rank_log_clear(spaqr_tree* t);
/* 7 doubles: interfaces, sketched, fell back to geqp3, direct geqp3, sum of sketch widths
   (retries included), sum of min(k, w) over the sketched interfaces, sketches retried with a
   doubled width (since the last reset) */
int spaqr_tree_rank_stats(const spaqr_tree* t, double* out);
int spaqr_tree_rank_stats_reset(spaqr_tree* t);

/* The predictor on its own (host buffers; the GPU kernels run underneath and synchronize).
   Features X are n x 32 row-major; targets y are rank fractions rank / min(k, w) in [0, 1].
   predict: out = sigmoid(MLP(X)).  grad: full-batch gradient of the mean logistic loss
   mean_i softplus(o_i) - y_i o_i.  train: `epochs` full-batch Adam steps (lr) from W, in
   place; loss_hist (optional, epochs floats) gets the loss before each step. */
int spaqr_mlp_params(void);
int spaqr_mlp_features(void);
int spaqr_mlp_predict(const float* X, int n, const float* W, float* out);
int spaqr_mlp_grad(const float* X, const float* y, int n, const float* W, float* grad, float* loss);
int spaqr_mlp_train(const float* X, const float* y, int n, float* W, int epochs, float lr, float* loss_hist);

/* ------------------------------------------------------------------ device memory helpers
   For callers without a GPU runtime of their own (cgo, JNI, ctypes).  No reference
   counterpart: the reference keeps everything in host memory. */
void* spaqr_device_alloc(size_t bytes);
int spaqr_device_free(void* p);
int spaqr_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int spaqr_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int spaqr_device_synchronize(void);

/* ------------------------------------------------------------------ batched kernels */
/* dgeqrt(m, n, nb=32, A, lda, T, 32): A <- V, R; T is 32 x min(m,n).    LAPACKE_dgeqrt */
typedef struct { void* A; void* T; int m, n, lda, pad; } spaqr_geqrt_desc;
int spaqr_geqrt_batched(int prec, int count, const spaqr_geqrt_desc* d, void* stream);

/* C (m x n) <- Q^T C with Q from a geqrt (V m x k, T 32 x k); then, if R != NULL,
   C <- C R^{-1} (R upper k2 x k2, n == k2).  dlarfb('L','T','F','C') / cblas_dtrsm(Right,
   Upper, NoTrans, NonUnit).  V may be NULL (trsm only). */
typedef struct { const void* V; const void* T; void* C; const void* R; int m, n, k, ldv, ldc, k2, ldr, pad; } spaqr_larfb_desc;
int spaqr_larfb_batched(int prec, int count, const spaqr_larfb_desc* d, void* stream);

/* dtpqrt(M=rows, N=k, L=0, NB=32) chained: for each link j of a task, in order,
   [R; B_j] <- QR, V_j in B_j, T_j (32 x k).  One task per interior front: the chain of
   tsqrt_edge calls over edgesOut (src/ptree.cpp:197-214, 1165-1167).  Chains are dispatched
   as the level planner does: k <= 64 with links of <= 64 rows one per wavefront, several
   panels (k > 32) with more than one link on the cooperative kernel, the rest one workgroup
   each.                                                                                   */
typedef struct { void* B; void* T; int rows, ldb; } spaqr_tpqrt_link;
typedef struct { void* R; int k, ldr, link0, nlinks; } spaqr_tpqrt_desc;
int spaqr_tpqrt_batched(int prec, int count, const spaqr_tpqrt_desc* d, int nlinks, const spaqr_tpqrt_link* links,
                        void* stream);

/* dtpmqrt('L','T', M=rows, N=ncols, K=k, L=0, NB=32) chained over the links of a task:
   [Atop (k x ncols, lda); B_j(:, col0:col0+ncols)] <- Q_j^T [...].                      */
typedef struct { const void* V; const void* T; void* B; int rows, ldv, ldb, pad; } spaqr_tpmqrt_link;
typedef struct { void* Atop; int k, lda, ncols, link0, nlinks, col0; } spaqr_tpmqrt_desc;
int spaqr_tpmqrt_batched(int prec, int count, const spaqr_tpmqrt_desc* d, int nlinks, const spaqr_tpmqrt_link* links,
                         void* stream);

/* dgeqp3 + choose_rank(|R_ii|, tol) on C (k x w, ld k), stopped at the rank: rank_out (device
   int), V = C(:, 0:rank) in Vout (ld k, k x min(k,w) capacity), tau, T (32 x rank), plus
   scratch vn1, vn2 (w each) and jpvt (w ints, 0-based pivots).  C is overwritten.
   tie_rel = 0 pivots on LAPACK's argmax (first maximum); tie_rel > 0 picks the lowest index
   whose partial norm is >= (1 - tie_rel) * max (the tree path uses 1e-10). */
typedef struct { void* C; void* Vout; void* tau; void* T; void* vn1; void* vn2; int* jpvt; int* rank_out; int k, w, pad0, pad1; } spaqr_geqp3_desc;
int spaqr_geqp3_batched(int prec, int count, const spaqr_geqp3_desc* d, double tol, double tie_rel, void* stream);

/* spaqr_geqp3_batched without any device allocation: same arguments, kernels and results.
   No counterpart in the reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its scratch
   (cooperative exchange slots, barrier counters, error flag, descriptor tables) lives in a
   fixed array of spaqr_geqp3_noalloc_scratch_bytes() bytes (16 MiB) that each GPU holds from
   library load to process exit.  A call needing more returns -6 before launching anything.
   Calls that overlap on one GPU share that array and race: issue them one at a time per GPU
   (every call synchronizes `stream` before returning). */
int spaqr_geqp3_batched_noalloc(int prec, int count, const spaqr_geqp3_desc* d, double tol, double tie_rel,
                                void* stream);
int spaqr_geqp3_noalloc_scratch_bytes(void);

/* Sparsification apply + scatter (src/ptree.cpp:1867-1897 then :330): A (mc x nc) <- A Q_n
   (Vn != NULL, first rn reflectors of Vn kn x rn), A[0:kc, 0:rn] <- Q_c^T A (Vc != NULL,
   first rc reflectors), dst(i, j) = A(i, j) for i < rc, j < rn; ident != 0 writes I_rc.
   pad = 0 runs every phase; 1 only A <- A Q_n; 2 only the left transform and the scatter. */
typedef struct {
    void* A; const void* Vn; const void* Tn; const void* Vc; const void* Tc; void* dst;
    int mc, nc, lda, kn, rn, kc, rc, ldvn, ldvc, ldd, ident, pad;
} spaqr_spars_desc;
int spaqr_spars_apply_batched(int prec, int count, const spaqr_spars_desc* d, void* stream);

/* ------------------------------------------------------------------ rank-path kernels
   The randomized sparsification's kernels on caller buffers (rank modes 1 and 2 above).  No
   reference counterpart; the closest analogue is random_rrqr (src/util.cpp:643-715).  All
   matrices are column-major with ld = rows; every call synchronizes `stream`. */

/* Sketch rounds per interface in rank modes 1 and 2: retries (0..2) doublings of the sketch
   width after a failed tail test before the geqp3 fallback (default 2). */
int spaqr_tree_set_rank_retries(spaqr_tree* t, int retries);
/* cluster order of each rank-log entry (the order of spaqr_tree_get_clusters' rows) */
int spaqr_tree_rank_log_clusters(const spaqr_tree* t, int* out, int cap);

/* Gaussian test matrix Omega (rows x cols, ld): entry (i, j) depends on (i, j, seed) only
   (splitmix64 of seed ^ (j << 32 | i), then Box-Muller), so a larger matrix agrees with a
   smaller one on their overlap.  The planner uses seed 0x5eed5ca1e.  prec 16 writes bf16: the
   FP32 matrix rounded to nearest even. */
int spaqr_gauss_fill(int prec, void* omega, int rows, int cols, int ld, uint64_t seed, void* stream);

/* Column norms: cn[j] = ||C(:, j)||^2 (w values) and *cmax2 = max(*cmax2, max_j cn[j]); tasks
   may share one cmax2, which must hold 0 (or a lower bound) on entry. */
typedef struct { const void* C; void* cn; void* cmax2; int k, w; } spaqr_colnorm_desc;
int spaqr_colnorm_batched(int prec, int count, const spaqr_colnorm_desc* d, void* stream);

/* Sketch Y (k x s, ld k) = C (k x w, ld k) * Omega(0:w, 0:s) (Omega ld ldo >= every w). */
typedef struct { const void* C; void* Y; int k, w, s, pad; } spaqr_sketch_desc;
int spaqr_sketch_batched(int prec, int count, const spaqr_sketch_desc* d, const void* omega, int ldo, void* stream);
/* The BF16 sketch of FP32 C (no reference counterpart; the reference sketches in FP64,
   random_rrqr, src/util.cpp:643-715): C and Omega rounded to bf16 (nearest even), products
   accumulated in FP32 on the bf16 matrix cores, so |Y - C Omega| <= (2 u + u^2 + c w 2^-24)
   |C| |Omega| elementwise with u = 2^-8.  Omega is bf16 (spaqr_gauss_fill prec 16) with ldo a
   multiple of 8 and >= every w rounded up to 32 (rows w .. of each 32-row step are read and
   discarded).  tile_cols (64, 128, 256): output-tile columns per workgroup; C is read once per
   tile_cols columns of Y.  -4 Omega null or not 16-byte aligned, -5 bad ldo, -6 bad tile_cols. */
int spaqr_sketch_bf16_batched(int count, const spaqr_sketch_desc* d, const void* omega_bf16, int ldo, int tile_cols,
                              void* stream);

/* The predictor's 32 input features of one interface (x, floats): from Y16, the geqrt-factored
   16-column sketch (its R diagonal Y16[j + j k]), the column norms cn and cmax2 of C, and the
   tree context (neighbour counts, level, tol).  x[0..15]: 1 + log10(|R_jj| / |R_00|) / 8
   clamped to [0, 1]; x[16..26]: log2(k+1)/11, log2(w+1)/14, min(k,w)/max(k,w), level/16,
   -log10(tol)/8, nsout/32, nsin/32, the share of columns with cn >= tol^2 cmax2 over w and
   over min(k,w) (capped at 1), cmax2 / sum(cn), min(1, sum(cn) / cmax2 / min(k,w));
   x[27..31] = 0. */
typedef struct { const void* Y16; const void* cn; const void* cmax2; float* x; int k, w, nsout, nsin, level, pad; }
    spaqr_rank_features_desc;
int spaqr_rank_features_batched(int prec, int count, const spaqr_rank_features_desc* d, double tol, void* stream);

/* Tail selection on B = Q^T C (k x w): *rank = max(*rank, the smallest r <= *ry such that
   every column j has ||B(r:k, j)||^2 <= tol^2 * (*cmax2)), or INT_MAX when some column's tail
   at *ry still exceeds it.  *rank must hold 0 on entry; tasks (column chunks of one interface)
   may share it.  Sums in FP64. */
typedef struct { const void* B; const void* cmax2; const int* ry; int* rank; int k, w; } spaqr_tail_desc;
int spaqr_tail_select_batched(int prec, int count, const spaqr_tail_desc* d, double tol, void* stream);

/* The same selection from B's leading rows only (the product form the FP32 planner uses for
   interfaces of more than 1024 rows; the tail test of the rank path, which stands in for the
   residual check of random_rrqr src/util.cpp:643-715): B holds rows 0 .. r-1 of Q^T C (r x w,
   ld r, r >= *ry), cn the squared column norms of C (w entries), and the tail of column j at row
   i is cn[j] - ||B(0:i, j)||^2.  Otherwise as spaqr_tail_select_batched. */
typedef struct { const void* B; const void* cn; const void* cmax2; const int* ry; int* rank; int r, w; }
    spaqr_tail_lead_desc;
int spaqr_tail_lead_batched(int prec, int count, const spaqr_tail_lead_desc* d, double tol, void* stream);

/* B (r x w, ld r) = M(0:r, 0:k) C (k x w, ld k), M with ld ldm >= r; FP32 on the matrix cores
   (the product form's Q(:, 0:r)^T C, with M = Q^T formed on the identity). */
typedef struct { const float* M; const float* C; float* B; int r, k, w, ldm; } spaqr_tail_gemm_desc;
int spaqr_tail_gemm_batched(int count, const spaqr_tail_gemm_desc* d, void* stream);

/* Matrix Market files (include/mmio.hpp:174-377 of the reference: sp_mmread, dense_mmread /
   vector_mmread, sp_mmwrite, dense_mmwrite; the driver reads A and the coordinates with them,
   pspaQR.cpp:85,161-167).  Host only, no GPU.
   read_sparse: a coordinate file of type real, integer or pattern (entries 1), property general,
   symmetric / hermitian (off-diagonal entries mirrored) or skew-symmetric (mirrored, negated),
   into CSC with rows sorted in each column and duplicates summed.  Call it first with colptr,
   rowind and val NULL for the sizes (nnz counts the mirrored entries), then with arrays of
   ncols + 1, nnz and nnz entries.  read_dense: an array file (general) into column-major data
   (a vector is an n x 1 array); data NULL for the sizes.  write_sparse: a real coordinate file
   (symmetric != 0 writes the lower triangle with property symmetric); write_dense: a real array
   file.  Values are written with 17 significant digits.  Return 0, -1 the file cannot be opened,
   -2 unsupported or malformed header, -3 malformed line or index out of range, -4 entry count
   differs from the header. */
int spaqr_mm_read_sparse(const char* path, int* nrows, int* ncols, long long* nnz, int* colptr, int* rowind,
                         double* val);
int spaqr_mm_read_dense(const char* path, int* nrows, int* ncols, double* data);
int spaqr_mm_write_sparse(const char* path, int nrows, int ncols, const int* colptr, const int* rowind,
                          const double* val, int symmetric);
int spaqr_mm_write_dense(const char* path, int nrows, int ncols, const double* data);

/* Joins the library's host worker threads.  Call it after the last tree is destroyed and
   before the process exits (the Python binding does so at interpreter exit); a later call
   into the library starts them again.  No reference counterpart. */
int spaqr_shutdown(void);

#ifdef __cplusplus
}
#endif
#endif /* SPAQR_GPU_H */
