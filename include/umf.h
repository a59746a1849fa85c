/* umf.h -- C ABI of the MI355X multifrontal factor / update / solve library (libumf_hip.so).
 *
 * One plan = one nested-dissection factorization of a 2D/3D grid operator held in device
 * memory (HBM) of the current HIP device.  Host arrays passed in are caller-owned and copied;
 * device memory is plan-owned.  Every call is synchronous unless its name ends in _device,
 * whose pointers are device pointers and whose work is ordered on the plan's stream.
 * Calls on one plan must be serialized by the caller; distinct plans are independent.
 *
 * Return status: UMF_OK (0); negative = error (see umf_last_error()); positive = numerical
 * failure: a non-positive pivot, with umf_last_error() naming the level, box and pivot column
 * (the reference aborts on LAPACK info != 0, include/LinAlg/dense.hpp:187).
 *
 * Each entry point names the reference interface it stands in for.
 */
#ifndef UMF_H
#define UMF_H 1

#ifdef __cplusplus
extern "C" {
#endif

#define UMF_OK 0
#define UMF_EINVAL (-1)     /* bad argument */
#define UMF_EGEOM (-2)      /* planner geometry error (reference throws std::runtime_error) */
#define UMF_EDEVICE (-3)    /* HIP runtime error */
#define UMF_ENOMEM (-4)     /* device allocation failed */
#define UMF_EPATTERN (-5)   /* CSR pattern differs from the planned one */
#define UMF_EUPDATEMAP (-6) /* a modified DOF maps outside the leaf box grid */
#define UMF_ESTATE (-7)     /* call out of order (e.g. solve before factor) */
#define UMF_ENOCONV (-8)    /* FP32 form: refinement did not reach the requested backward error */
#define UMF_EBUSY (-9)      /* serving: the slot pool has no free range for the request now */

typedef struct umf_plan umf_plan;

/* Front-block selectors for umf_get_front_block (names of the reference's Box fields,
 * include/DE/mf.hpp:45-55).  A22 is rebuilt as L*L^T from the stored Cholesky factor; XR, XL
 * and S are formed from L and W = L^{-1} A21.  A11post is the update matrix handed to the
 * parent; its diagonal separator blocks are never read by the reference (the parent replaces
 * them from the separators, src/DE/mf2d.cpp:427-443) and hold S there instead, so they come
 * back as NaN. */
#define UMF_BLOCK_A22 0
#define UMF_BLOCK_XL 1
#define UMF_BLOCK_XR 2
#define UMF_BLOCK_S 3
#define UMF_BLOCK_A11POST 4
#define UMF_BLOCK_L 5 /* lower Cholesky factor of A22 (no reference counterpart) */

/* Replaces the constructors MF2D(A, Nx, Ny, width, maxSideLength) (include/DE/mf.hpp:67) and
 * MF3D(A, Nx, Ny, Nz, width, maxSideLength) (include/DE/mf.hpp:173): plans the elimination
 * tree from the grid, uploads the CSR operator (int32 indices, n x n, DOF id i + j*Nx [+ k*Nx*Ny])
 * and allocates the device front pool.  dim = 2 or 3 (nz ignored for 2D).  The pattern must be
 * symmetric with no duplicate entries; symmetric values are factored by Cholesky (they must be
 * positive definite), others by LU fronts (umf_set_lu). */
int umf_plan_create(int dim, int nx, int ny, int nz, int width, int max_side_length, int n,
                    const int* row_ptr, const int* col_ind, const double* val, umf_plan** plan);

/* Batch of independent problems (no reference counterpart: the reference factors one problem
 * per object; this is one MF2D/MF3D per problem merged): problem q is described by
 * desc[7q..7q+6] = {dim, nx, ny, nz, width, maxSideLength, n} and its CSR arrays, as for
 * umf_plan_create.  The batch is one plan of the block-diagonal operator diag(A_0, A_1, ...):
 * DOF d of problem q is merged DOF dof0(q) + d, so umf_factor / umf_update / umf_solve and their
 * device variants work on the whole batch (a right-hand-side column stacks every problem's
 * right-hand side; an update's modified DOFs may touch any problems, each gets its own dirty
 * sets).  Every level of every tree runs in the same batched launches: the trees are aligned at
 * the root.  flags bit 0: plan only (no device memory, val may be NULL).  Per-tree structure
 * and block queries take single-problem plans.  Every problem's values must be symmetric
 * (Cholesky fronts). */
int umf_batch_create(int nprob, const int* desc, const int* const* row_ptr, const int* const* col_ind,
                     const double* const* val, int flags, umf_plan** plan);
/* Problem `prob` of a batch: first merged DOF, size, tree levels, level shift (merged level =
 * own level + shift).  Returns the number of problems (1 for a single-problem plan). */
int umf_batch_info(const umf_plan* plan, int prob, int* dof0, int* n, int* levels, int* shift);

/* Symbolic analysis only (no device memory, no values): the elimination tree, DOF lists and
 * assembly programs of umf_plan_create, for structure queries on hosts without a GPU.
 * umf_factor / umf_update / umf_solve on such a plan return UMF_ESTATE. */
int umf_plan_analyze(int dim, int nx, int ny, int nz, int width, int max_side_length, int n,
                     const int* row_ptr, const int* col_ind, umf_plan** plan);

/* Computes, without running it, the dirty sets an update with these modified DOFs would
 * factor (MF2D::Update's marking, src/DE/mf2d.cpp:161-174, and the level recurrence of
 * src/DE/mf2d.cpp:579-620); afterwards umf_dirty reports them.  Works on analyzed plans. */
int umf_analyze_update(umf_plan* plan, int n_mod, const int* mod_dofs);

/* Frees the plan and all its device memory (the reference's destructor). */
int umf_plan_destroy(umf_plan* plan);

/* Replaces MF2D::Factor / MF3D::Factor (src/DE/mf2d.cpp:449-500, src/DE/mf3d.cpp:705-766):
 * factors every front, leaves to root, on the device. */
int umf_factor(umf_plan* plan);

/* Replaces MF2D::Update(A, modifiedDOFs) / MF3D::Update (src/DE/mf2d.cpp:150-181,
 * src/DE/mf3d.cpp:208-248): installs the new CSR (same pattern as planned) and re-factors the
 * reference's dirty fronts only: the leaf boxes holding the modified DOFs, then per level the
 * parents of both neighbours of every touched separator, and the root. */
int umf_update(umf_plan* plan, const int* row_ptr, const int* col_ind, const double* val, int n_mod,
               const int* mod_dofs);

/* Device-resident variant of umf_update: d_val holds the new CSR values (same pattern as
 * planned, nnz doubles, device memory); the caller guarantees symmetry.  No reference
 * counterpart (the reference copies a host Sparse). */
int umf_update_device(umf_plan* plan, const double* d_val, int n_mod, const int* mod_dofs);

/* Replaces MF2D::Solve / MF3D::Solve (src/DE/mf2d.cpp:976-1054, src/DE/mf3d.cpp:1237-1313) for
 * nrhs right-hand sides at once: b and x are n x nrhs column-major host arrays with leading
 * dimension ld >= n (x may equal b). */
int umf_solve(umf_plan* plan, const double* b, double* x, int nrhs, int ld);

/* Replaces MF2D::Apply / MF3D::Apply (src/DE/mf2d.cpp:904-974, src/DE/mf3d.cpp:1164-1235):
 * applies the factored operator to nrhs columns (host arrays, as umf_solve). */
int umf_apply(umf_plan* plan, const double* x, double* y, int nrhs, int ld);

/* Device-pointer variants of umf_solve / umf_apply (no reference counterpart: the reference
 * takes one host vector per call).  d_b / d_x are n x nrhs column-major in device memory. */
int umf_solve_device(umf_plan* plan, const double* d_b, double* d_x, int nrhs, int ld);
int umf_apply_device(umf_plan* plan, const double* d_x, double* d_y, int nrhs, int ld);

/* umf_solve_device that allocates nothing (no counterpart in the reference): for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Its work block, per-front right-hand-side scratch and id list live in
 * one file-scope device array of 64 MiB (each GPU the library runs on holds its own until the
 * process ends); it calls no hipMalloc / hipFree.  Limits: FP64 factors only (UMF_ESTATE
 * otherwise), and plans whose solve fits the array -- about n <= 230k unknowns in 2D -- larger
 * ones return UMF_EINVAL naming the size needed.  Every call on one GPU shares the array: calls
 * that overlap on one GPU (other host threads, other plans' streams) race and must be serialized
 * by the caller.  Results are bit-identical to umf_solve_device. */
int umf_solve_device_noalloc(umf_plan* plan, const double* d_b, double* d_x, int nrhs, int ld);

/* CSR SpMV y = A x with the plan's current operator, device pointers (the reference's
 * Sparse::operator*(Vec), src/LinAlg/sparse.cpp:63-77), for residuals. */
int umf_spmv_device(umf_plan* plan, const double* d_x, double* d_y, int nrhs, int ld);

/* Coefficient-change request on the harness operator -div(a grad u) of
 * test_include/DE/buildLaplacian.hpp:10-132, :244-391 (the nodal, padded coefficient field a;
 * face value (a_p + a_q)/2; stencil width = the plan's width), done on the device: scale a by
 * `factor` on the box of padded nodes [lo, lo + side) per axis, recompute in d_val the values of
 * every row the change reaches (d_val: the plan's CSR pattern, holding the generator's values),
 * and return the modified DOFs -- the columns of the entries whose value changed, the
 * reference's rule test_src/DE/testMF2D.cpp:143-150 -- ascending in mod_dofs (host, capacity
 * cap; *n_mod = their count, also when cap is too small).  d_coeff: device array of a, last
 * index fastest, (N_axis + 2w) per axis; h: the generator's grid spacing.  The reference
 * rebuilds the whole operator on the host (its test mains); this keeps coefficient-change
 * requests on the device.  Synchronous. */
int umf_box_scale_device(umf_plan* plan, double* d_coeff, double h, const int* lo, int side, double factor,
                         double* d_val, int* mod_dofs, int cap, int* n_mod);

/* Plan files (no reference counterpart; the reference keeps a factorization only in memory):
 * umf_plan_save writes a plan's geometry, operator (its current values) and, when factored, its
 * factors in the precision they are held in; umf_plan_load rebuilds the plan from the file and
 * restores the factors without refactoring, so Solve / Apply / Update can follow at once.  For
 * reproducible inputs and warm starts of a serving process.  Single-problem plans only; the file
 * is specific to this library's build (its front-pool layout is checked on load). */
int umf_plan_save(umf_plan* plan, const char* path);
int umf_plan_load(const char* path, umf_plan** out);

/* Last error message of the calling thread. */
const char* umf_last_error(void);

/* ---- structure queries (parity with the reference's private state, include/DE/mf.hpp:103-139) */
/* levels_ and the leaf box widths boxWidth{X,Y,Z}Bottom_ (axis 0..2). */
int umf_levels(const umf_plan* plan);
int umf_leaf_width(const umf_plan* plan, int axis);
int umf_num_boxes(const umf_plan* plan, int level);
/* Separators between boxes of `level` (1 <= level < levels), the reference's edgeData_[level-1]
 * / faceData_[level-1]. */
int umf_num_seps(const umf_plan* plan, int level);
/* Box DOF lists: which = 0 intDOFs, 1 bdDOFs.  Writes to out when non-null; returns the count. */
int umf_box_dofs(const umf_plan* plan, int level, int box, int which, int* out);
int umf_sep_dofs(const umf_plan* plan, int level, int sep, int* out);
/* Dirty sets of the last factor/update: which = 0 boxes factored at `level`, 1 separators that
 * received Schur contributions at `level`, 2 separators rebuilt from their children at `level`.
 * Sorted ascending.  Returns the count. */
int umf_dirty(const umf_plan* plan, int level, int which, int* out);

/* ---- factor-entry queries (parity); see UMF_BLOCK_*.  Writes column-major rows x cols. */
int umf_get_front_block(umf_plan* plan, int level, int box, int which, double* out, int* rows, int* cols);
/* The separator's A11 (k x k), symmetric, both triangles filled. */
int umf_get_sep_block(umf_plan* plan, int level, int sep, double* out, int* k);

/* ---- measurement */
typedef struct umf_stats {
  double t_factor_ms;   /* device time of the last umf_factor */
  double t_update_ms;   /* device time of the last umf_update */
  double t_solve_ms;    /* device time of the last umf_solve / umf_solve_device */
  double flops_factor;  /* sum over fronts of n^3/3 + n^2 m + n m^2 */
  double flops_update;  /* the same sum over the fronts of the last update */
  double flops_solve;   /* nrhs * sum (2 n^2 + 4 n m) of the last solve */
  double bytes_front_pool; /* device bytes held by fronts and separators */
  int n_fronts;
  int n_fronts_update;
  int ir_iters;         /* FP32 form: refinement iterations of the last solve (max over columns) */
  double ir_berr;       /* FP32 form: final normwise backward error in the infinity norm (max over columns) */
} umf_stats;
int umf_get_stats(const umf_plan* plan, umf_stats* out);

/* Precision form (no reference counterpart; the reference is FP64 throughout).
 *   UMF_PREC_FP64: fronts stored and factored in FP64 (the drop-in form, the default).
 *   UMF_PREC_FP32: fronts stored and factored in FP32 on v_mfma_f32_16x16x4 (half the front
 *     memory); umf_solve then runs preconditioned conjugate gradients in FP64 (FP64 operator
 *     and residuals, the FP32 multifrontal solve as preconditioner) until every column reaches
 *     the normwise backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= tol, at most max_iter
 *     iterations (UMF_ENOCONV when a column does not); umf_apply applies the FP32 factors.
 *   UMF_PREC_BF16X3: the FP32 form with the factorization's Schur and trailing-update GEMMs
 *     on v_mfma_f32_16x16x32_bf16: each FP32 operand is split into two bfloat16 pieces,
 *     x = x0 + x1, and the three products x0 y0 + x0 y1 + x1 y0 are accumulated in FP32 (about
 *     16 significant bits per product).  Fronts, pivot blocks, panels, solve and refinement are
 *     those of UMF_PREC_FP32.  The split runs on the levels whose GEMM work lies mostly in fronts
 *     of 1024+ pivots; the other levels keep FP32 MFMA.
 *   UMF_PREC_FP32T64: the top three separator levels (0..2; every level of a plan with at most
 *     three) stored and factored in FP64, pivot blocks, TRSM and Schur updates alike; the levels
 *     below in FP32 as UMF_PREC_FP32.  The lowest FP64 level assembles from the FP32 blocks of
 *     the level below, converting them; the solve sweeps the FP32 levels in FP32 and the top
 *     levels in FP64.  umf_solve refines as UMF_PREC_FP32, umf_apply applies the mixed factors.
 *     Not served (umf_server_*); umf_plan_save stores such a plan unfactored.
 * Takes effect at the next umf_factor (the plan must be factored again). */
#define UMF_PREC_FP64 0
#define UMF_PREC_FP32 1
#define UMF_PREC_BF16X3 2
#define UMF_PREC_FP32T64 3
int umf_set_precision(umf_plan* plan, int precision, double tol, int max_iter);

/* Per-kernel timing (no reference counterpart; the reference's HASCPP11TIMING timers are
 * the nearest, src/DE/mf2d.cpp:463-499): when enabled, every kernel launch of the plan is
 * bracketed by HIP events on the plan stream and its algorithmic flops / bytes are counted.
 * umf_profile_read(plan, i, ...) fills entry i and returns the number of kernels seen. */
int umf_profile_enable(umf_plan* plan, int on);
/* Restrict the timing to launches of one kernel (base name, e.g. "k_chol_panel"), so that a
 * timed region carries events only around the kernel being measured; NULL or "" = all. */
int umf_profile_select(umf_plan* plan, const char* kernel);
int umf_profile_reset(umf_plan* plan);
int umf_profile_read(const umf_plan* plan, int i, char* name, int name_len, double* ms_total, int* launches,
                     double* flops, double* bytes);

/* Front factorization (the reference factors every front by LU, include/LinAlg/dense.hpp:180-227):
 * operators with symmetric values are factored by Cholesky (the default, which requires SPD
 * input); an operator whose values are not symmetric -- pattern-symmetric, as the reference
 * requires (include/DE/mf.hpp:9-12), e.g. the harness operator with stencil width 2-3 and a
 * variable coefficient -- gets LU fronts at umf_plan_create: partial pivoting within each front's
 * interior rows, FP64, the same blocks (A22, XL, XR, S, A11post; UMF_BLOCK_L = the packed L\U of
 * the row-permuted A22), Solve, Apply and Update.  umf_set_lu(plan, 1) makes a plan use LU fronts
 * also for symmetric values (indefinite operators), 0 returns to Cholesky (symmetric values
 * only); either takes effect at the next umf_factor.  LU fronts are FP64 only and are not served
 * (umf_server_*) or batched.  umf_get_lu reports the current form. */
int umf_set_lu(umf_plan* plan, int on);
int umf_get_lu(const umf_plan* plan);

/* Publication between the workgroups of the blocked solve sweeps (no reference counterpart):
 * fenced = 0 (default) publishes each block as 64-bit words that carry the launch's epoch beside
 * the data (relaxed agent-scope atomics, polled by the readers: no flag, no cache maintenance);
 * fenced = 1 writes the block, then sets a per-block flag between agent-scope release / acquire
 * fences (slower: L2 write-back / invalidate per block).  Both are ordered by the HIP memory
 * model; results are identical. */
int umf_set_sweep_sync(umf_plan* plan, int fenced);

/* hipStream_t of the plan (for callers ordering their own work), returned as void*. */
void* umf_stream(umf_plan* plan);

/* Factorization as a HIP graph (no reference counterpart; the reference factors with host
 * calls, src/DE/mf2d.cpp:449-500): with on != 0, umf_factor runs the whole level loop as one
 * captured graph, captured at the first umf_factor after enabling (and again after a precision
 * change) and replayed by the later ones; results are bit-identical to the direct launches.
 * umf_update and the solves keep direct launches (their schedules depend on the call).  Not
 * used while per-kernel timing is enabled. */
int umf_set_graph(umf_plan* plan, int on);

/* Tuning of the factorization's GEMM stages (no reference counterpart; the reference's BLAS
 * picks its own blocking): gw2 = width of the outer column groups of the two-level trailing
 * update (256..8192, a multiple of 256; default 1024), tile = 64 or 128 for the trailing and
 * Schur GEMM tiles, or 0 (default) for 128 when a stage's largest front has 16 such tiles and
 * its fronts 2048 in all, else 64.  The tile size does not change results (each entry accumulates the same
 * products in the same order); gw2 changes the order of the trailing updates. */
int umf_set_tune(umf_plan* plan, int gw2, int tile);
int umf_get_tune(const umf_plan* plan, int* gw2, int* tile);
/* The autotuner: factors the plan's current operator in its precision with every candidate
 * (gw2 in {1024, 2048} x tile in {0, 64, 128}), `reps` times each, keeps the fastest by the
 * factorization's event time (*best_ms, may be NULL) and refactors with it. */
int umf_autotune(umf_plan* plan, int reps, double* best_ms);

/* ---- serving (no reference counterpart; the reference's use case is Update after a local
 * coefficient change followed by Solve, test_src/DE/testMF3D.cpp:137-215, on one object).
 * A server answers requests against an immutable factored base plan: request r = the base
 * operator with a local change (its values, its modified DOFs by the reference's rule) and
 * right-hand sides; its answer is A_r^{-1} b -- what Update(A_r) + Solve on the base would give,
 * bit for bit -- while the base stays unchanged and requests never see each other's changes.
 * A request runs in a slot (its own stream, descriptors, operator values and scratch); its dirty
 * fronts and separators (the reference's Update recurrence) are factored into a range of the
 * server's slot pool, every other front is read from the base.  Its footprint is the bytes of
 * those fronts (with their inverted diagonal blocks) and separators.  Slots run concurrently:
 * distinct slots may be driven by distinct host threads; one slot by one thread at a time.
 * While a server lives, umf_factor / umf_update / umf_set_precision / umf_autotune /
 * umf_plan_destroy on its base return UMF_ESTATE. */
typedef struct umf_server umf_server;
/* slots: 1..32 execution contexts; pool_bytes: device memory for the requests' dirty fronts. */
int umf_server_create(umf_plan* base, int slots, double pool_bytes, umf_server** server);
int umf_server_destroy(umf_server* server);
/* Slot-pool bytes a request with these modified DOFs needs. */
int umf_server_footprint(umf_server* server, int n_mod, const int* mod_dofs, double* bytes);
/* Admission: reserves the request's footprint in the pool (first fit) and returns a ticket;
 * UMF_EBUSY when no free range holds it now (retry after a release), UMF_EINVAL when the whole
 * pool could not. */
int umf_server_admit(umf_server* server, int n_mod, const int* mod_dofs, int* ticket);
int umf_server_release(umf_server* server, int ticket);
int umf_server_pool(umf_server* server, double* total_bytes, double* used_bytes);
/* The request's operator in slot `slot`: full CSR values (device pointer, the plan's pattern),
 * or a coefficient box change on the base operator, as umf_box_scale_device (the base field
 * d_coeff is read, not changed; the modified DOFs come back in mod_dofs). */
int umf_server_set_values(umf_server* server, int slot, const double* d_val);
int umf_server_box_scale(umf_server* server, int slot, const double* d_coeff, double h, const int* lo, int side,
                         double factor, int* mod_dofs, int cap, int* n_mod);
/* Runs an admitted request in `slot` with the operator set there: factors its dirty fronts into
 * its pool range, then solves nrhs columns (device pointers, as umf_solve_device; the refined
 * forms refine against the request's operator).  Synchronous on the slot's stream. */
int umf_server_run(umf_server* server, int slot, int ticket, const double* d_b, double* d_x, int nrhs, int ld);
/* Timing and refinement statistics of the slot's last request (t_update_ms = its factor pass). */
int umf_server_stats(const umf_server* server, int slot, umf_stats* out);

#ifdef __cplusplus
}
#endif
#endif
