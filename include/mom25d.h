/* C-ABI of libmom25d.so: the gfx950 kernels of mom25d's per-frequency solve.
 *
 * Conventions
 *   - every pointer argument except `args` / host structs is a device pointer
 *     (the caller, e.g. PyTorch, owns every buffer; the library allocates
 *     nothing);
 *   - complex numbers are interleaved (re, im) doubles (complex128) or floats
 *     (complex64); matrices are row-major with a leading dimension counted in
 *     complex elements;
 *   - `stream` is a hipStream_t; all calls are asynchronous on it;
 *   - the return value is 0 on success and < 0 on error, with the text in
 *     mom_last_error() (thread-local).  No C++ exception crosses the ABI.
 *
 * Each entry point names the reference interface (mom25d, pure Python) whose
 * work it takes over.
 */
#ifndef MOM25D_H
#define MOM25D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* mom_last_error(void);
int mom_abi_version(void);

/* Sommerfeld integrals of the Green's-function remainders, batched over the
 * table rows of every frequency of a sweep:
 *   out[row, c] = sum_k J0(rho[row] kc_k) Cc[k, c] + sum_k J0(rho[row] kr_k) Cr[k, c]
 * over the nodes of problem row_prob[row] (kc: complex lifted path, kr: real
 * tail; offsets / lengths per problem).  ncol = kernels * levels^2 with 1, 2
 * or 3 kernels and 1 to 3 levels.
 * Replaces GroundedSlabGreens._sommerfeld / residual
 * (Engine/mom25d/greens.py:157-173) and LayeredGreens._sommerfeld / _residual
 * (Engine/mom25d/layered.py:360-377). */
int mom_sommerfeld(int n_rows, const double* rho, const int32_t* row_prob, const int64_t* kc_off,
                   const int32_t* kc_len, const double* kc, const double* Cc, const int64_t* kr_off,
                   const int32_t* kr_len, const double* kr, const double* Cr, int ncol, double* out, void* stream);

/* Element-wise Bessel J0 (complex z, or real part only when real_axis != 0):
 * the scipy.special.jv(0, .) / j0(.) calls of Engine/mom25d/greens.py:164. */
int mom_j0(int n, const double* z, double* out, int real_axis, void* stream);

/* FP64 impedance-matrix fill of one tier (0 near, 1 mid, 2 far) for n_items
 * work items; `args` points to a host FillArgs block (layout: csrc/fill.hip,
 * size mom_fill_args_size()).  Adds into Z (zeroed by the caller; 16-byte
 * aligned): through atomics, or with plain loads and stores for the entries
 * the ownership arrays row_own / col_own mark as one work item's alone
 * (null arrays: atomics everywhere).
 * Replaces _fast.chunk_vals / group_vals / scatter (Engine/mom25d/_fast.py:82-158,
 * 190-211) and the block loop of assemble (Engine/mom25d/solver.py:474-509). */
int mom_fill_tier(int tier, int n_items, const void* args, void* stream);
/* The near tier's frequency-independent closed-form parts (S0 = I0 / 4 pi,
 * S1 = ((r - o) I0 + I1) / 4 pi per test point and source triangle) of the
 * n_items near work items in args->items, written to `out` at the offsets
 * args->item_soff; pair_dz[lt * nl + ls] is the distance of the two levels.
 * Once per mesh; the near tier of mom_fill_tier reads them from
 * args->near_static.  Replaces the near-pair statics of AssemblyGeometry
 * (Engine/mom25d/solver.py:373-376, integrals.py:24-67). */
int mom_near_statics(int n_items, const void* args, const double* pair_dz, double* out, void* stream);
int mom_fill_args_size(void);
int mom_pair_tab_size(void);

/* ---- plan object of the fill (one per mesh) ----------------------------
 * The arrays of the reference's frequency-independent set-up,
 * AssemblyGeometry (Engine/mom25d/solver.py:307-426), flattened: host
 * pointers, row-major, read during mom_plan_create only.
 *   V (nt,3,2), qpts (nt,7,2), qw (nt,7), mpts (nt,3,2), mw (nt,3),
 *   fpts (nt,2) centroids, fw (nt) areas, cA / cV (nt,3), tri_basis (nt,3)
 *   (geom.b, -1: no basis), zl (n_levels);
 *   test blocks (geom.blocks[k][:2]): blk_off (n_blocks+1) into blk_tris,
 *   blk_layer (n_blocks) = lt;
 *   source groups (geom.blocks[k][3], in order): grp_block, grp_layer (ls),
 *   grp_tier (0: near, the group has statics; 1: mid, 3-point rule;
 *   2: far, centroid), grp_off (n_groups+1) into grp_src (= sel);
 *   pad_pos (nt): row of each triangle in IVpad, -1, or NULL without vias. */
typedef struct MomAssemblyGeometry {
  int32_t nt, nb, n_levels, n_blocks, n_groups, pad_;
  const double* V;
  const double* qpts;
  const double* qw;
  const double* mpts;
  const double* mw;
  const double* fpts;
  const double* fw;
  const double* cA;
  const double* cV;
  const int64_t* tri_basis;
  const double* zl;
  const int64_t* blk_off;
  const int64_t* blk_tris;
  const int32_t* blk_layer;
  const int32_t* grp_block;
  const int32_t* grp_layer;
  const int32_t* grp_tier;
  const int64_t* grp_off;
  const int64_t* grp_src;
  const int64_t* pad_pos;
} MomAssemblyGeometry;

/* One fill: jwmu = j omega mu0, ijwe = 1 / (j omega eps0) (the reference's
 * jwmu / inv_jwe, solver.py:462-463); the spline tables of the frequency as
 * device arrays of this library's layout (pairs, tab_x, tab_c, tab_b), or
 * pairs = NULL for the plan's own (mom_plan_set_pair); Z the complex128 RWG
 * block (nb x nb at row stride ldz), IVpad complex128 (n_pad, nt) or NULL
 * without pads; stage_max (0: 64) and tier_mask (0: all three tiers). */
typedef struct MomFillCall {
  double jwmu_re, jwmu_im, ijwe_re, ijwe_im;
  const void* pairs;
  const double* tab_x;
  const double* tab_c;
  const void* tab_b;
  void* Z;
  int64_t ldz;
  void* IVpad;
  int32_t stage_max, tier_mask;
} MomFillCall;

typedef struct MomPlanInfo {
  int32_t n_items[3];
  int32_t n_pad;
  int64_t acc_bytes[3];
  int64_t tier_pairs[3];
  int64_t near_static_len;
  int64_t n_src;
} MomPlanInfo;

/* sizeof of the three blocks above, for bindings to check their mirrors. */
int mom_plan_struct_sizes(int32_t* out3);
/* Builds the plan of one mesh (work items of <= 64 sources per block and
 * group, scatter maps, distance bounds); with on_device, uploads it on the
 * current device and computes the near tier's statics on `stream`
 * (synchronised before returning).  Replaces the per-sweep set-up that
 * assemble's loop of _fast.chunk_vals re-derives per block
 * (Engine/mom25d/solver.py:307-426, 472-483). */
int mom_plan_create(const MomAssemblyGeometry* g, int on_device, void* stream, void** plan);
int mom_plan_destroy(void* plan);
int mom_plan_info(const void* plan, MomPlanInfo* info);
/* Read access to the plan's host arrays by name (src, items0..2, item_r0..2,
 * blk_tris, blk_row_off, blk_rows, blk_rloc, item_cols, src_cloc,
 * item_soff), valid until mom_plan_destroy. */
int mom_plan_host_array(const void* plan, const char* name, const void** data, int64_t* count);
/* The spline of level pair (lt, ls) for the plan's own tables, in the
 * reference's layout (_fast.spline_data, Engine/mom25d/_fast.py:35-45):
 * nodes x (n), PPoly coefficients c (4, n-1, 2) complex128 of the smooth
 * kernels [S_A, S_V], the pair's dz and the coefficients cA, cV of its
 * 1/(4 pi R) terms (pair.cA, pair.cV); zero marks an uncoupled pair. */
int mom_plan_set_pair(void* plan, int lt, int ls, int n, const double* x, const double* c, double dz,
                      double cA_re, double cA_im, double cV_re, double cV_im, int zero);
/* Adds the RWG block of Z and IVpad (device pointers, zeroed by the caller)
 * on `stream`: what assemble's loop of _fast.chunk_vals + scatter adds into
 * Z[:nb, :nb] and IVpad (Engine/mom25d/solver.py:472-483,
 * Engine/mom25d/_fast.py:145-211). */
int mom_plan_fill(void* plan, const MomFillCall* call, void* stream);
/* The same into host arrays (numpy: Z at row stride ldz, IVpad (n_pad, nt));
 * stores the values (as chunk_vals + scatter into the zeroed Z of
 * assemble do) and returns when they are there.  Device staging is
 * plan-owned and reused. */
int mom_plan_fill_host(void* plan, const MomFillCall* call);

/* Z[b_i, b_j] += zs * g[t][i][j]: surface-impedance (conductor loss) term,
 * Engine/mom25d/solver.py:511-513. */
int mom_gram_add(int nt, const double* g, const int32_t* tb, double zs_re, double zs_im, double* Z, int64_t ldz,
                 void* stream);

/* Via rows / columns and the via-via block of Z from IVpad (pad triangles x
 * all triangles) and the host part E (nv x nv), Engine/mom25d/solver.py:527-566. */
int mom_via_blocks(int nb, int nv, int nt, const double* IVpad, const int32_t* d_off, const int32_t* d_pad,
                   const double* d_val, const int32_t* pad_tri, const int32_t* tp, const int32_t* vp,
                   const int32_t* tm, const int32_t* vm, const double* cV, const double* E, double ijwe_re,
                   double ijwe_im, double* Z, int64_t ldz, void* stream);

/* Z <- (Z + Z^T)/2 in place; a complex64 copy into Z32 when non-null, scaled
 * to d_i Z_ij d_j when d is non-null (Engine/mom25d/solver.py:522-524). */
int mom_symmetrize(int n, double* Z, int64_t ldz, float* Z32, int64_t ld32, const double* d, void* stream);

/* d_i = 1 / sqrt(|Z_ii|): symmetric diagonal scaling for the complex64 LU
 * (no counterpart in the reference, which factors in complex128 only). */
int mom_diag_scale(int n, const double* Z, int64_t ldz, double* d, void* stream);

/* Line voltage at points of level li from the triangle charges q (nt x nrhs),
 * Engine/mom25d/solver.py:1101-1125 (Solver.line_voltage).  tab_b: the bucket
 * entries of the tables, 16 bytes each ({double split; int32 i0; int32 pad}). */
int mom_line_voltage(int npts, const double* pts, int li, int nl, int nt, int nrhs, const int32_t* tri_layer,
                     const double* V, const double* qp7, const double* qw7, const void* pairs, const double* tab_x,
                     const double* tab_cw, const void* tab_b, const double* q, double inv_eps0, double* out,
                     void* stream);

/* LU with partial pivoting of an n x n complex matrix in place (prec 64:
 * complex128, 32: complex64), panel width jb; ipiv[k] = row swapped with
 * row k (0-based); info[0] = first zero pivot + 1 or 0; info[1] = 1 when a
 * panel's workgroups stopped waiting for each other (a bounded wait ran out;
 * the factorisation is then invalid), else 0 (info holds 2 ints).  `work` holds
 * mom_getrf_work_bytes(prec, n, jb) bytes.  max_groups bounds the
 * workgroups of the panel kernel, whose workgroups must all be resident at
 * once (0: all compute units); a caller that factors L matrices concurrently
 * on L streams passes CUs / L so that every panel can be resident together.
 * Replaces scipy.linalg.lu_factor (LAPACK zgetrf) of _dense_solve,
 * Engine/mom25d/solver.py:11-23. */
int64_t mom_getrf_work_bytes(int prec, int n, int jb);
int mom_getrf(int prec, int n, void* A, int64_t lda, int32_t* ipiv, int32_t* info, void* work, int64_t work_bytes,
              int jb, int max_groups, void* stream);
/* The same factorisation without row interchanges (ipiv[k] = k), for the
 * preconditioner of the mixed-precision GMRES (no counterpart in the
 * reference, which factors with LAPACK's partial pivoting only).  A panel is
 * then its diagonal block factored by one workgroup plus an independent solve
 * per row below it, with no cross-workgroup exchange per column, and the
 * factorisation is bounded by its update GEMM.  On the diagonally scaled MoM
 * matrices the element growth stays near 2.  An exactly zero pivot is
 * replaced by 1 and reported in info[0] (the factors are then finite but not
 * those of A).  Panel width min(jb, 128) in complex64, min(jb, 64) in
 * complex128 inside update blocks of width jb.  No max_groups: every launch
 * runs to completion by itself. */
int mom_getrf_nopiv(int prec, int n, void* A, int64_t lda, int32_t* ipiv, int32_t* info, void* work,
                    int64_t work_bytes, int jb, void* stream);
/* The same factorisation of a complex-symmetric matrix (A = A^T, as the
 * diagonally scaled MoM matrix D Z D is: Galerkin testing with the basis
 * itself), A = L D L^T without pivoting.  Only the lower triangle of A is
 * read; on return the strict lower triangle holds L (unit diagonal) and the
 * upper triangle U = D L^T, so that the factors are laid out exactly as
 * mom_getrf_nopiv's and every solve entry point takes them unchanged.  Each
 * update block's row of U is the transposed, pivot-scaled column of L
 * already computed (no triangular solve), and the trailing updates skip the
 * tiles strictly above the diagonal (which the later blocks re-derive from
 * L): about half the update flops of mom_getrf_nopiv.  ipiv[k] = k; zero
 * pivots as mom_getrf_nopiv.  No counterpart in the reference. */
int mom_sytrf_nopiv(int prec, int n, void* A, int64_t lda, int32_t* ipiv, int32_t* info, void* work,
                    int64_t work_bytes, int jb, void* stream);

/* Workgroup tile (rows x real columns) of the trailing-update GEMM of
 * mom_getrf and mom_cgemm_sub, per precision: 64x64, 64x128, 128x64 or
 * 128x128 (the autotuner's choice; default 64x128).  Applies to the calls
 * issued after it, process-wide. */
int mom_set_gemm_tile(int prec, int bm, int bn);
int mom_get_gemm_tile(int prec, int* bm, int* bn);

/* Polls before a wait of the LU panel or of the triangular-solve chain gives
 * up and flags the call as failed (info[1] / the solve workspace's fifth
 * counter); default 2^21 (about a second).  Process-wide. */
int mom_set_spin_limit(unsigned polls);

/* Block width nb of mom_getrf's trailing updates, per precision (0: the panel
 * width; otherwise a multiple of 8 up to 256, rounded down to a multiple of
 * the panel width): a block of nb columns is factored as panels of jb
 * columns, then the rest of the matrix takes one update of depth nb.
 * Applies to the calls issued after it, process-wide (the autotuner's
 * choice). */
int mom_set_lu_block(int prec, int nb);
int mom_get_lu_block(int prec);

/* Row permutation of the factorisation: perm[i] = original row of row i. */
int mom_build_perm(int n, const int32_t* ipiv, int32_t* perm, void* stream);

/* X = (LU)^-1 P B for nrhs <= 8 right-hand sides (n x nrhs, row-major).
 * `work` holds mom_getrs_work_bytes(prec, n) bytes of device memory (block
 * tickets, cleared by the call itself, and the intermediate L^-1 P B); its
 * fifth uint32 is the failure word: a solve whose bounded wait ran out sets
 * it to 1 (X is then invalid) and no solve clears it, so the caller zeroes it
 * (with a new workspace, and before a run of solves it wants to check) and
 * reads it once after the run.  Replaces scipy.linalg.lu_solve (LAPACK zgetrs),
 * Engine/mom25d/solver.py:23. */
int64_t mom_getrs_work_bytes(int prec, int n);
int mom_getrs(int prec, int n, int nrhs, const void* LU, int64_t lda, const int32_t* perm, const void* B, void* X,
              void* work, void* stream);

/* mom_getrs without a workspace argument and without any allocation: the
 * tickets and the intermediate L^-1 P B live in the library's static device
 * array (about 4 MB, held on each device the library runs on until the
 * process ends), for n <= 16384 and nrhs <= 8; larger sizes are rejected
 * with an error.  *status (a caller word, device memory) is set to 1 when a
 * bounded wait ran out and is never cleared by the call.  Calls that overlap
 * on one device race on the static array.  No counterpart in the reference:
 * for callers that cannot allocate (graph capture, serving loops). */
int mom_getrs_noalloc(int prec, int n, int nrhs, const void* LU, int64_t lda, const int32_t* perm, const void* B,
                      void* X, uint32_t* status, void* stream);

/* The same solve with the 64 x 64 diagonal blocks of L and U applied as
 * explicit inverses (a parallel product per block instead of the serial
 * substitution; about twice as fast, backward error ~ cond(block) eps instead
 * of eps).  mom_trtri_blocks computes the inverses (in FP64, rounded to prec)
 * into `work` (mom_getrs_inv_work_bytes(prec, n) bytes) once per
 * factorisation; every mom_getrs_inv with that `work` then reuses them.  Used
 * as the complex64 preconditioner of the FP64 GMRES refinement, where the
 * outer iteration absorbs the difference.  Same role as scipy.linalg.lu_solve,
 * Engine/mom25d/solver.py:23, in the mixed-precision path that the reference
 * does not have. */
int64_t mom_getrs_inv_work_bytes(int prec, int n);
/* Row blocks of the complex64 mom_getrs_inv(_z) chain: 64 (default) or 128
 * (the 128-row inverses composed by mom_trtri_blocks from the 64-row ones;
 * half as many blocks in the chain).  Process-wide. */
int mom_set_solve_rows(int rows);
int mom_get_solve_rows(void);
/* Chain form of the 64-row solves with inverted blocks (default on):
 * mom_trtri_blocks also forms M_k = D_k^-1 C_k, C_k the coupling of block k to
 * the block solved just before it, so that a block computes
 * y_k = D_k^-1 (b_k - all other couplings) while it waits and then only
 * x_k = y_k - M_k x_prev.  Off: x_k = D_k^-1 (b_k - all couplings) after the
 * wait.  Process-wide. */
int mom_set_solve_chain(int on);
int mom_get_solve_chain(void);
int mom_trtri_blocks(int prec, int n, const void* LU, int64_t lda, void* work, void* stream);
int mom_getrs_inv(int prec, int n, int nrhs, const void* LU, int64_t lda, const int32_t* perm, const void* B,
                  void* X, void* work, void* stream);
/* mom_getrs_inv with complex128 B and X (the LU in prec): B is rounded to
 * prec on the way in and the solution widened on the way out, inside the
 * solve's own kernels (the GMRES preconditioner step). */
int mom_getrs_inv_z(int prec, int n, int nrhs, const void* LU, int64_t lda, const int32_t* perm, const void* B,
                    void* X, void* work, void* stream);

/* W = D Z D X (complex128, D = diag(dsc), nrhs <= 4): the product of the
 * diagonally scaled system in the GMRES refinement, one HBM pass over Z.  No
 * counterpart in the reference, which solves in FP64 directly
 * (Engine/mom25d/solver.py:11-31). */
int mom_scaled_matvec(int n, int nrhs, const void* Z, int64_t ldz, const double* dsc, const void* X, void* W,
                      void* stream);

/* The same product for a symmetric Z (Z = Z^T, as the fill and the frequency
 * interpolation leave it), reading only the 64 x 64 tiles on and above the
 * diagonal: half the HBM traffic of mom_scaled_matvec.  work: device
 * workspace of mom_scaled_symv_work_bytes(n, nrhs) bytes (the tiles' partial
 * products, summed in a fixed order: the result does not depend on
 * scheduling).  No counterpart in the reference, which solves in FP64
 * directly (Engine/mom25d/solver.py:11-31). */
int64_t mom_scaled_symv_work_bytes(int n, int nrhs);
int mom_scaled_symv(int n, int nrhs, const void* Z, int64_t ldz, const double* dsc, const void* X, void* W,
                    void* work, int64_t work_bytes, void* stream);

/* One Arnoldi step of the mixed path's restarted GMRES in one call: W_j =
 * M^-1 V_j (mom_getrs_inv_z with the factors LU, perm and their solve
 * workspace), w = D Z D W_j into V_{j+1} (mom_scaled_symv when symv_work is
 * non-null, else mom_scaled_matvec), classical Gram-Schmidt twice against
 * V_0 .. V_j with the summed coefficients in H[0 .. j], H[j+1] = |w| and
 * V_{j+1} = w / |w| per column, then H[0 .. j+1] ((j + 2) x nrhs) copied to
 * the pinned host buffer Hh on the stream.  V: (m + 1) x n x nrhs and W:
 * m x n x nrhs complex128; H: device workspace of mom_gmres_work_elems(m,
 * nrhs) complex128 (the column, the second Gram-Schmidt pass, and the
 * partial dot products and norms of 8 row ranges, summed in a fixed order);
 * 0 <= j < m, nrhs <= 4.  No counterpart in the reference, which solves in
 * FP64 directly (Engine/mom25d/solver.py:11-31). */
int64_t mom_gmres_work_elems(int m, int nrhs);
int mom_gmres_arnoldi(int prec, int n, int nrhs, int j, int m, const void* LU, int64_t lda, const int32_t* perm,
                      void* solve_work, const void* Z, int64_t ldz, const double* dsc, void* symv_work,
                      int64_t symv_bytes, void* V, void* W, void* H, void* Hh, void* stream);

/* Super-block triangular solves of unpivoted factors (the mixed path's
 * preconditioner M = L U of D Z D).  With the factors cut into super-blocks
 * of sb = 512 rows, D_k the inverse of the k-th diagonal super-block of L
 * (of U) and T_kj = -D_k L_kj (j < k), -D_k U_kj (j > k), the solves read
 * x_k = D_k b_k + sum over the super-blocks solved before k of T_kj x_j:
 * each row is an independent dot product, and only the term of the
 * super-block solved just before sits on the chain's critical path (n / sb
 * hops instead of mom_getrs's n / 64).
 * mom_sbsolve_work_bytes(prec, n, sb): workspace bytes (-1 unless sb = 512;
 * about 2 n^2 elements of prec for D and T plus the vectors).
 * mom_sbsolve_setup: D and T from the factors LU (lda) into work, once per
 * factorisation (inverting the diagonal super-blocks in FP64).
 * mom_sbsolve_z: X = M^-1 P B for complex128 B, X (n x nrhs, nrhs <= 4,
 * B != X), perm as mom_getrs's; the workspace's status words hold the
 * workgroup tickets and a sticky failure word (word 4, set by a bounded wait
 * that gave up or a ticket outside the grid, cleared only by the caller).
 * mom_gmres_arnoldi_sb: mom_gmres_arnoldi with this solve in place of
 * mom_getrs_inv_z.  No counterpart in the reference, which solves in FP64
 * directly (Engine/mom25d/solver.py:11-31). */
int64_t mom_sbsolve_work_bytes(int prec, int n, int sb);
int mom_sbsolve_setup(int prec, int n, const void* LU, int64_t lda, int sb, void* work, int64_t work_bytes,
                      void* stream);
int mom_sbsolve_z(int prec, int n, int nrhs, int sb, const int32_t* perm, const void* B, void* X, void* work,
                  void* stream);
int mom_gmres_arnoldi_sb(int prec, int n, int nrhs, int j, int m, int sb, const int32_t* perm, void* sb_work,
                         const void* Z, int64_t ldz, const double* dsc, void* symv_work, int64_t symv_bytes, void* V,
                         void* W, void* H, void* Hh, void* stream);

/* Frequency interpolation of Z (Engine/mom25d/freqinterp.py:86-139).
 * mom_interp_anchor: W = j w Z exp(+j ke R) (N x N, W stored in prec_w = 64 or 32
 * bits per real part), R the distance of the basis-function centres pos
 * (N x 3, FP64), 0 on the via-via block (indices >= nb).
 * mom_interp_combine: Z = (sum_k L[k] W[k]) exp(-j ke R) / (j w) from K <= 8
 * anchors; W and L are host arrays of device pointers and weights. */
int mom_interp_anchor(int prec_w, int N, int nb, const void* Z, int64_t ldz, const double* pos, double omega,
                      double ke, void* W, int64_t ldw, void* stream);
int mom_interp_combine(int prec_w, int N, int nb, int K, const void* const* W, const double* L, int64_t ldw,
                       const double* pos, double omega, double ke, void* Z, int64_t ldz, void* stream);

/* C (m x n) -= A (m x k) B (k x n), complex, on the matrix cores (the LU's
 * trailing update, exposed for testing and tuning). */
int mom_cgemm_sub(int prec, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                  int64_t ldc, void* stream);

/* Event timing of selected launches (roofline report; no reference
 * counterpart): kind 0 / 4 the complex64 / complex128 LU's trailing-update
 * GEMMs (work in flops), 1 the strip pass of mom_scaled_symv (bytes: the
 * symmetric half of Z and the columns), 2 / 3 each triangular chain of the
 * complex64 / complex128 super-block solves (bytes: one triangle of the
 * factors and the columns).
 * begin(kind, capacity) records up to `capacity` launches of that kind,
 * each between two HIP events on its own stream; end(kind, out) waits and
 * writes out[0] = launches, out[1] = total ms, out[2] = total work. */
int mom_timing_begin(int kind, int capacity);
int mom_timing_end(int kind, double* out);

/* FP64 iterative refinement pieces: R = B - Z X (nrhs <= 4; rowsum, when
 * non-null, receives the row sums of |Re z| + |Im z| of diag(d) Z diag(d),
 * d = 1 when null), complex128 -> complex64 conversion, and X += D with the
 * column maxima |D|, |X| (|re|+|im|, as IEEE bit patterns in uint64). */
int mom_residual(int n, int nrhs, const void* Z, int64_t ldz, const void* X, const void* B, void* R, double* rowsum,
                 const double* d, void* stream);
int mom_cvt_z_to_c(int64_t n_reals, const double* src, float* dst, void* stream);
int mom_refine_update(int n, int nrhs, const void* D, void* X, void* dmax, void* xmax, void* stream);

/* GMRES orthogonalisation against a basis V (j1 x n x nrhs, complex128):
 * h = V^H w per column, and w -= V h (no counterpart in the reference, whose
 * dense solve is a direct LAPACK factorisation). */
int mom_basis_dot(int n, int nrhs, int j1, const void* V, const void* w, void* h, void* stream);
int mom_basis_sub(int n, int nrhs, int j1, const void* V, const void* h, void* w, void* stream);

/* HIP graphs of library calls (no counterpart in the reference, which has no
 * GPU path): calls issued on `stream` between mom_graph_begin and
 * mom_graph_end are captured (the stream must not be the default stream,
 * nothing else may be enqueued on it meanwhile, and the calls must not
 * allocate); mom_graph_launch replays them with the same arguments. */
int mom_graph_begin(void* stream);
int mom_graph_end(void* stream, void** graph_exec);
int mom_graph_launch(void* graph_exec, void* stream);
int mom_graph_destroy(void* graph_exec);

/* Surrogate of the smooth Green's-function remainders R_A, R_V: an MLP
 * 6 -> 128 -> 128 -> 128 -> 4 with SiLU (FP32 parameter vector of
 * mom_mlp_param_count() values: W1[128][6] b1 W2[128][128] b2 W3 b3 W4[4][128]
 * b4; BF16 copies of W2, W3 from mom_mlp_pack).  It stands in for the
 * Sommerfeld integrals of Engine/mom25d/layered.py:333-377 (`_remainder`,
 * the Bessel sums) when the stack's tables are built; the integrals stay the
 * judge of its error.
 * mom_mlp_forward: out[n][4] from X[n][6]; precision 16 = BF16 matrix cores
 * (the training arithmetic), 32 = FP32 throughout.
 * mom_mlp_train_step: squared-error gradient of one batch, one partial
 * gradient of mom_mlp_partial_stride() floats per 64-sample tile
 * (nslots = ceil(n / 64)); mom_mlp_adam sums the partials (and the loss,
 * into *loss) and, when step_lr != 0, applies Adam to P, m1, m2 and the
 * BF16 copies; grad, when non-null, receives the summed gradient. */
int mom_mlp_param_count(void);
int mom_mlp_partial_stride(void);
int mom_mlp_pack(const float* P, uint16_t* Wbf, void* stream);
int mom_mlp_forward(int precision, int n, const float* X, const float* P, const uint16_t* Wbf, float* out,
                    void* stream);
int mom_mlp_train_step(int n, const float* X, const float* Y, const float* wout, float grad_scale, const float* P,
                       const uint16_t* Wbf, float* part, int nslots, void* stream);
int mom_mlp_adam(int nslots, const float* part, float* P, float* m1, float* m2, uint16_t* Wbf, float* grad,
                 float* loss, float beta1, float beta2, float eps, float step_lr, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MOM25D_H */
