/*
 * hssc.h -- C ABI of the MI355X-native HSS compressed direct solver.
 *
 * The entry points mirror the MATLAB functions of HSSC_DirectSolver (Corona, Martinsson,
 * Zorin) that make up the compress -> invert -> apply path.  Each declaration names the
 * reference function it replaces (paths relative to the reference repository root).
 *
 * Conventions (the MEX / ctypes layers convert from MATLAB's):
 *   - arrays are column-major with explicit sizes; point sets are N x 2 (x column, then y);
 *   - index vectors are int64 and 0-based (MATLAB is 1-based);
 *   - every call returns an int status (HSSC_OK on success); the message of the last
 *     failure on the calling thread is returned by hssc_last_error();
 *   - handles own device memory on the device that was current when they were created;
 *     calls on one handle must be serialised by the caller, distinct handles are
 *     independent.
 */
#ifndef HSSC_H
#define HSSC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------ */
#define HSSC_OK 0
#define HSSC_ERR_INVALID_ARG 1
#define HSSC_ERR_OUT_OF_MEMORY 2
#define HSSC_ERR_SINGULAR 3
#define HSSC_ERR_RANK_OVERFLOW 4
#define HSSC_ERR_HIP 5
#define HSSC_ERR_NOT_CONVERGED 6
#define HSSC_ERR_UNSUPPORTED 7

/* ---- enumerations ------------------------------------------------------------------ */
#define HSSC_POT_SL_L_2D 0 /* 2D Laplace single layer, Common/Kernel_Eval.m:60-61     */
#define HSSC_POT_SL_H_2D 1 /* 2D Helmholtz single layer, Common/Kernel_Eval.m:44-49   */

#define HSSC_INTER_DENSE 0   /* params.INTERform = 'dense'   */
#define HSSC_INTER_LOWRANK 1 /* params.INTERform = 'lowrank' */

#define HSSC_PREC_FP64 0 /* factorization (F^-1, E) stored and applied in FP64            */
#define HSSC_PREC_FP32 1 /* factorization in FP32; solves refine against the FP64 operator */

/* The params struct of Common/HSS_tree_parameters.m:116-139 (field names kept). */
typedef struct hssc_params {
  int32_t flag_pot;   /* HSSC_POT_*                                   (.flag_pot)   */
  double kh;          /* wavenumber                                   (.kh)         */
  int32_t sym;        /* symmetric kernel                             (.sym)        */
  int32_t Ng;         /* points per panel side                        (.Ng)         */
  int32_t np;         /* panels per domain side                                     */
  double acc;         /* ID / interpolation accuracy                  (.acc)        */
  int32_t layers;     /* boundary layers of the skeleton              (.layers)     */
  int32_t n_cut;      /* HSS switch threshold (levHSS)                (.n_cut)      */
  int32_t transinv;   /* translation invariant kernel                 (.transinv)   */
  int32_t depth;      /* tree depth, leaves are panels                (.depth)      */
  double h;           /* grid spacing 2/(Ng*np)                       (.h)          */
  int32_t max_particles; /* Ng*Ng                                     (.max_particles) */
  int32_t interform;  /* HSSC_INTER_*                                 (.INTERform)  */
  int32_t k_cut;      /* 2500                                         (.k_cut)      */
  double dr_weights_re; /* Duan-Rokhlin corr0, real part              (.dr_weights) */
  double dr_weights_im; /* Duan-Rokhlin corr0, imaginary part                       */
  int32_t factor_precision; /* HSSC_PREC_* (no reference counterpart: GPU policy)   */
  int32_t proxy_layers; /* > 0: proxy ring depth at lev > 0 instead of lay_prox
                           (HSS2D_build_tree_fsym.m:339-347); 0 keeps the reference rule */
  int32_t preset_all;  /* kh > 0: 0 = skeletons of the two bottom levels chosen by an ID of
                          K(X_proxy, X_src) (HSS2D_build_tree_fsym.m:376, :686-729), as the
                          reference; 1 = the 'bdry' rule at every level (uniform shapes) */
  /* Options without a reference counterpart (the benchmark configurations' literal forms;
     hssc_params_init sets the reference's choices): */
  int32_t contrast;     /* kh > 0, NTI, sym: HSSC_CONTRAST_PLATEAU = b = c = sqrt of the tanh
                           plateau (Bump_function option 4, Common/Kernel_Eval.m:133,143);
                           HSSC_CONTRAST_GAUSSIAN = b = c = sqrt(q),
                           q(x) = contrast_amp exp(-|x|^2 / (2 contrast_width^2)) */
  double contrast_amp;
  double contrast_width;
  double kernel_scale;  /* SL_L_2D: factor s of the kernel s log r (1 = the reference's
                           h^2 log r, Kernel_Eval.m:61; 1/(2 pi) = (1/2 pi) log r) */
  double length_scale;  /* SL_L_2D: physical length L of the reference's unit: spacing L h and
                           distances L r (1 = [-1,1]^2 as the reference; 1/2 = [0,1]^2) */
  int32_t tree_precision; /* HSSC_PREC_* of the interpolation operators T (no reference
                           counterpart: GPU policy); FP32 T serves only as a preconditioner,
                           with residuals taken against the exact operator (hssc_grid_op) */
  int32_t proxy;        /* params.proxy (Common/HSS_tree_parameters.m:138, default 0): in
                           Kernel_Eval's LOCAL_get_bc (Common/Kernel_Eval.m:111-159), > 0 drops
                           the row factor b, < 0 the column factor c (HSS1D tests use 1); the
                           HSS2D (fsym) entry points need 0 */
  int32_t proxy_rule;   /* params.proxy_rule (:123): HSSC_PROXY_BDRY (default) or
                           HSSC_PROXY_RAND (HSS2D_build_tree_fsym.m:899-921) */
  int32_t skel_extra_pts; /* params.skel_extra_pts (:122, default 0): 1 adds skeleton points
                           chosen by a randomized ID of the interpolation residual
                           (LOCAL_ID_rand_update, HSS2D_build_tree_fsym.m:1069-1124) */
  int32_t reserved0;
  double lambda;        /* params.lambda (:124, default 2): decay of the 'rand' proxy density */
  int32_t reserved[4];
} hssc_params;

#define HSSC_PROXY_BDRY 0
#define HSSC_PROXY_RAND 1

#define HSSC_CONTRAST_PLATEAU 0
#define HSSC_CONTRAST_GAUSSIAN 1

typedef struct hssc_tree hssc_tree;         /* INFO_TREE + params.X_source            */
typedef struct hssc_inverse hssc_inverse;   /* A_HSS of HSS2D_invert_* (FI, E, JE)    */
typedef struct hssc_operator hssc_operator; /* A_HSS of HSS2D_compress_* (Bself, B12) */

/* Thread-local text of the last failure ("" if none). */
const char* hssc_last_error(void);

/* Library version string. */
const char* hssc_version(void);

/* Replaces Common/HSS_tree_parameters.m:1 (params = HSS_tree_parameters(pot,kh,sym,Ng,np,
 * acc,lay,n_cut,TI)), defaults of :116-139.  For kh > 0 it also sets the order-4
 * Duan-Rokhlin weight (:167-175) from hssc_dr_corr0(h, kh). */
int hssc_params_init(hssc_params* p, int32_t pot, double kh, int32_t sym, int32_t Ng,
                     int32_t np, double acc, int32_t lay, int32_t n_cut, int32_t TI);

/* Order-4 Duan-Rokhlin centre weight corr0(h, k) of the 2D Helmholtz kernel H0(k r):
 * replaces Common/DR_stencil.m:64-107 (LOCAL_compute_duan_rokhlin_weights, DVEC(1)), which
 * runs the Fortran side-car Common/FORTRAN_get_DR_weights.f:85-490 (corr_hank).  Computed
 * from the lattice-sum closed form corr0 = h^2 + i (2h^2/pi)[log(hk/2) + gamma + Z'(0) +
 * sum_m (-1)^m (hk/2)^2m / (m!)^2 Z'(-2m)], Z the Epstein zeta function of Z^2.  Accepts
 * 0 < hk <= 2 (the side-car's tables stop at 1.01); HSSC_ERR_INVALID_ARG otherwise. */
int hssc_dr_corr0(double h, double k, double* corr0_re, double* corr0_im);

/* Sets params.dr_weights = corr0 (Common/DR_stencil.m:10-25 output, order 4). */
int hssc_params_set_dr_weights(hssc_params* p, double corr0_re, double corr0_im);

/* Number of grid points N = (Ng*np)^2. */
int64_t hssc_params_npoints(const hssc_params* p);

/* params.X_source of Common/HSS_tree_parameters.m:144-160, written N x 2 column-major. */
int hssc_source_points(const hssc_params* p, double* X);

/* Replaces Common/Kernel_Eval.m:1 (A = Kernel_Eval(X1,X2,params)) for the 2D plane:
 * A (m x n, column-major, leading dimension lda >= m) = K[X1, X2].  X1 is m x 2, X2 is
 * n x 2, both column-major host arrays.  Real kernels write doubles; complex kernels
 * write interleaved (re, im) pairs.  Computed on the GPU. */
int hssc_kernel_eval(const hssc_params* p, const double* X1, int64_t m, const double* X2,
                     int64_t n, double* A, int64_t lda);

/* Replaces HSS2D/HSS2D_build_tree_fsym.m:1 (INFO_TREE = HSS2D_build_tree_fsym(params)):
 * uniform tree, 'bdry' skeletons and proxies, and for kh > 0 (unless preset_all) skeletons
 * of the two bottom levels chosen by a batched GPU ID of K(X_proxy, X_src)
 * (LOCAL_Non_Preset_Skeleton_Info, :686-729).  Every box of a level keeps the level's
 * largest rank: a box of adaptive rank k_b < k takes the next k - k_b columns of its ID
 * order, and the preset levels above pad their 'bdry' sets to the level's largest count
 * with the points nearest the box boundary.  Then the interpolation operators
 * T = K(Xp,X_sk) \ K(Xp,X_rs), computed on the GPU by Householder least squares (at the
 * ID levels the same T as the ID's R11^-1 R12, :726) and stored dense (the 'lowrank'
 * INTERform of the reference stores the same operator to tolerance as T.U * T.V.').
 * Complex (SL_H_2D) trees hold complex T. */
int hssc_hss2d_build_tree(const hssc_params* p, hssc_tree** tree);

/* Geometry and index sets only (no interpolation operators).  Used by the same builder;
 * exposed so index-set parity can be checked without a device.  Needs the GPU only for the
 * ID-selected skeletons of kh > 0 trees (preset_all = 0). */
int hssc_hss2d_build_index_sets(const hssc_params* p, hssc_tree** tree);

/* Tree queries (the fields of INFO_TREE, HSS2D/HSS2D_build_tree_fsym.m:655-680). */
int hssc_tree_depth(const hssc_tree* t, int32_t* depth);
/* Per level: n = |I_src|, k = |I_sk|, M = proxy size, nbox = boxes, nkey = distinct
 * (J,T) sets (1 if transinv, nbox otherwise), q = stored rank of T (n-k if dense). */
int hssc_tree_level_info(const hssc_tree* t, int32_t lev, int64_t* n, int64_t* k,
                         int64_t* M, int64_t* nbox, int64_t* nkey, int64_t* q);
/* Heap box number b (1-based, as in the reference) -> I_src / I_sk (0-based global). */
int hssc_tree_get_isrc(const hssc_tree* t, int64_t box, int64_t* out);
/* Per key box of a level (nkey entries): its own skeleton count before the level's
 * padding, i.e. the adaptive ID rank (ID levels) or the 'bdry' count (preset levels);
 * *id_level = 1 if the level's skeletons come from the ID. */
int hssc_tree_box_ranks(const hssc_tree* t, int32_t lev, int32_t* ranks, int32_t* id_level);
int hssc_tree_get_isk(const hssc_tree* t, int64_t box, int64_t* out);
/* J of box b (NTI) or of b's level (TI), 0-based local indices, skeleton first. */
int hssc_tree_get_J(const hssc_tree* t, int64_t box, int64_t* out);
/* Proxy points of box b (or its level), M x 2 column-major. */
int hssc_tree_get_proxy(const hssc_tree* t, int64_t box, double* out);
/* Interpolation operator of box b (or its level) as a dense k x (n-k) column-major
 * matrix (T = T.U * T.V.' in the reference's low-rank form); interleaved complex doubles
 * for SL_H_2D. */
int hssc_tree_get_T(const hssc_tree* t, int64_t box, double* out);
/* Releases the caller's handle.  Inverses and operators made from the tree read its device
 * data and keep it alive: the tree's memory is freed when the handle and every such inverse
 * and operator have been destroyed, in any order. */
void hssc_tree_destroy(hssc_tree* t);

/* Replaces HSS2D/levHSS.m:1 (lHSS = levHSS(TREE,n_cut,trinv)). */
int hssc_levhss(const hssc_tree* t, int32_t n_cut, int32_t* lhss);

/* Replaces HSS2D/HSS2D_invert_fsym.m:1 with params.Fopts = 'dense' (dense F^-1 and E per
 * box, or per level when transinv).  precision = HSSC_PREC_FP64 or HSSC_PREC_FP32. */
int hssc_hss2d_invert(hssc_tree* t, int32_t precision, hssc_inverse** inv);

/* Replaces MATLAB's inv on the dense blocks of the inversion (HSS2D_invert_fsym.m:108,
 * HSS2D_build_Finv_fsym.m:101, HSS2D_build_E_fsym.m:51,64): each of `batch` n x n blocks of
 * device memory A (column-major, leading dimension lda, block stride sA elements) is replaced
 * by its inverse, by the library's Gauss-Jordan sweeps (pivoting inside each diagonal block
 * of 64 columns, or 128 for real entries).  complex_entries != 0: interleaved complex;
 * precision = HSSC_PREC_FP64 or HSSC_PREC_FP32 selects the entry type; sym != 0 (complex
 * only) for exactly symmetric blocks (A^T = A): upper-tile sweeps and a symmetric result.
 * Returns HSSC_ERR_SINGULAR when a pivot is zero, not finite or at most n eps max|A_block|
 * (the blocks are overwritten all the same). */
int hssc_dense_invert_device(int32_t complex_entries, int32_t precision, int32_t sym, int64_t n,
                             void* A, int64_t lda, int64_t sA, int64_t batch);
/* F^-1 (n x n, rows/cols in I_src order) and E (k x k) of box b (or its level), as doubles
 * or interleaved complex doubles (FP32 factorizations are widened). */
int hssc_inverse_get_FI(const hssc_inverse* inv, int64_t box, double* out);
int hssc_inverse_get_E(const hssc_inverse* inv, int64_t box, double* out);
void hssc_inverse_destroy(hssc_inverse* inv);

/* Replaces HSS2D/HSS2D_compress_fsym.m:1 with dense Bself / B12 blocks. */
int hssc_hss2d_compress(hssc_tree* t, hssc_operator** op);
void hssc_operator_destroy(hssc_operator* op);

/* Replaces HSS2D/HSS2D_applyinv_TI.m:1 and HSS2D/HSS2D_applyinv_fsym.m:1
 * (Y = HSS2D_applyinv_*(INFO_TREE,A_HSS,X,params)).  X, Y: N x m column-major host
 * arrays in the X_source ordering; interleaved complex doubles (2 per entry) for SL_H_2D. */
int hssc_hss2d_applyinv(const hssc_inverse* inv, const double* X, double* Y, int64_t m);

/* Replaces HSS2D/HSS2D_apply_TI.m:1 and HSS2D/HSS2D_apply_fsym.m:1. */
int hssc_hss2d_apply(const hssc_operator* op, const double* X, double* Y, int64_t m);

/* Device-pointer variants (X, Y already in device memory, double or interleaved complex
 * double as above; stream = hipStream_t or 0). */
int hssc_hss2d_applyinv_device(const hssc_inverse* inv, const void* dX, void* dY, int64_t m,
                               void* stream);
int hssc_hss2d_apply_device(const hssc_operator* op, const void* dX, void* dY, int64_t m,
                            void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * result as hssc_hss2d_applyinv_device, with the solve workspace (the per-level vectors of
 * the two sweeps) laid out in a fixed device buffer of 128 MiB that every GPU the library runs
 * on holds until the process ends, instead of library allocations; it calls no allocator.
 * HSSC_ERR_INVALID_ARG if the workspace of the tree and m does not fit (the message gives both
 * sizes; the 512^2 Helmholtz tree with m = 1 needs 80 MB in complex double, 40 MB in
 * complex float), if the current device is not
 * the inverse's, or if the tree's build is still pending.  Every call on a GPU uses that GPU's
 * buffer, which keeps the last call's contents: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller. */
int hssc_hss2d_applyinv_device_noalloc(const hssc_inverse* inv, const void* dX, void* dY,
                                       int64_t m, void* stream);

/* Replaces Common/Iterative_Refinement.m:1 with A = the FP64 compressed operator and
 * Ainv = the stored inverse: Y = Ainv(B); while ||B - A(Y)|| > tol*||B|| and
 * iter < maxit: Y += Ainv(B - A(Y)).  Returns the iterations done and the final relative
 * residual.  Device pointers. */
int hssc_hss2d_solve_device(const hssc_inverse* inv, const hssc_operator* op, const void* dB,
                            void* dY, int64_t m, double tol, int32_t maxit, int32_t* iters,
                            double* relres, void* stream);

/* ---- exact grid operator -----------------------------------------------------------
 * Replaces the dense products with A = Kernel_Eval(X_source, X_source, params) that the
 * reference's tests form to check a solve (HSS2D/Test/Test_HSS2D_InverseCompression.m:
 * 300-311, Test_HSS2D_Matrix_Apply.m:227-235): the exact operator of the uniform grid, applied
 * in O(N log N) by FFT convolution on the GPU (the kernel depends on point offsets only; the
 * weights b, c and the diagonal are applied pointwise).  Grids up to 512 x 512 points.  X, Y:
 * N x m column-major in the X_source ordering (double, or interleaved complex double for
 * SL_H_2D).  Calls on one handle must be serialised (it owns its work arrays). */
typedef struct hssc_grid_op hssc_grid_op;
int hssc_grid_op_create(const hssc_params* p, hssc_grid_op** op);
int hssc_grid_op_apply(const hssc_grid_op* op, const double* X, double* Y, int64_t m);
int hssc_grid_op_apply_device(const hssc_grid_op* op, const void* dX, void* dY, int64_t m,
                              void* stream);
void hssc_grid_op_destroy(hssc_grid_op* op);

/* Common/Iterative_Refinement.m:20-35 with A = the exact grid operator (so the stopping
 * residual is the true one, ||B - A Y|| / ||B||) and Ainv = the stored inverse (FP64 or FP32,
 * built on an FP64 or FP32 tree).  Returns the solves done and the final relative residual;
 * HSSC_ERR_NOT_CONVERGED if tol is not reached within maxit solves.  Device pointers. */
int hssc_hss2d_solve_exact_device(const hssc_inverse* inv, const hssc_grid_op* A,
                                  const void* dB, void* dY, int64_t m, double tol,
                                  int32_t maxit, int32_t* iters, double* relres, void* stream);

/* ---- interpolative decompositions ------------------------------------------------- */
/* Replaces Common/ID.m:1 ([T,I] = ID(A,par)): column-pivoted QR on the GPU; par < 1 is the
 * accuracy (rank = #{|R_ii| > par |R_11|} plus the reference's randomized check),
 * par >= 1 the fixed rank.  A: m x n host array (leading dimension lda).  Outputs: the
 * rank k, the column permutation perm (n, 0-based, skeleton first) and T (k x (n-k),
 * column-major, leading dimension k; room for n*n doubles). */
int hssc_id(const double* A, int64_t m, int64_t n, int64_t lda, double par, int64_t* perm,
            double* T, int64_t* k);

/* Replaces Common/LRID_rand.m:1 with dense matvecs: A ~ U V^T with U = A(:, J(1:k)) and
 * V(J,:) = [I T]^T, from a Gaussian sketch Y = A^T G of k0 columns grown by C columns while
 * the last pivoted-QR diagonal of Y exceeds acc (QR_findrank.m), then the fixed-rank ID of
 * Y^T (ID_rand.m).  U: m x min(m,n), V: n x min(m,n) host buffers; J: n entries. */
int hssc_lrid_rand(const double* A, int64_t m, int64_t n, int64_t lda, double acc, int64_t k0,
                   int64_t C, uint64_t seed, int64_t* k, int64_t* J, double* U, double* V);

/* ---- HSS1D: boundary integral equations on curves ----------------------------------
 * The operator is the second-kind Laplace double layer A = -1/2 I + D on a closed curve
 * given by its N nodes (X, N x 2), outward unit normals (nrm, N x 2), quadrature weights
 * w and curvature kappa (D_jj = -kappa_j w_j / 4pi).  The reference's Kernel_Eval takes
 * positions only; its 'green' compression with this kernel is the HSS1D workload measured
 * here.  Proxy circles carry log charges. */
typedef struct hssc_hss1d hssc_hss1d;

/* Replaces HSS1D/HSS1D_compress_nsym.m:1 (K_HSS = HSS1D_compress_nsym('green',X,nbox_max,
 * acc,params)): binary tree (halve until <= nbox_max points), proxy circles of nproxy
 * points at 1.75 x the circumradius, skeletons by GPU interpolative decompositions. */
int hssc_hss1d_compress(const double* X, const double* nrm, const double* w, const double* kappa,
                        int64_t N, int32_t nbox_max, double acc, int32_t nproxy,
                        hssc_hss1d** out);
/* Replaces HSS1D/HSS1D_compress_nsym.m:1 with type 'green' on the reference's own kernels:
 * every block, proxy rows (K(Zp, X_out)) and proxy columns (K(X_in, Zp)) included, is
 * Kernel_Eval(X1, X2, params) (Common/Kernel_Eval.m; :152-156, :222-249), as
 * HSS1D/Test/Test_HSS1D_compress.m runs it on the skeleton points of a 2D box.  X is N x 2
 * column-major; nproxy = 50 is the reference's (:52).  Real kernels only (SL_L_2D; the
 * params.proxy option selects Kernel_Eval's b / c factors): HSSC_ERR_UNSUPPORTED for SL_H_2D. */
int hssc_hss1d_compress_kernel(const hssc_params* p, const double* X, int64_t N,
                               int32_t nbox_max, double acc, int32_t nproxy, hssc_hss1d** out);
/* Replaces HSS1D/HSS1D_compress_nsym.m:1 with type 'brute' (:158-181): compression of the
 * given dense N x N matrix K (column-major, leading dimension ldk >= N).  A leaf's block row
 * and column are every index outside it; a parent's are the skeletons of the level below
 * outside it (:164-177, :210-214).  K is read during the call only. */
int hssc_hss1d_compress_dense(const double* K, int64_t N, int64_t ldk, int32_t nbox_max,
                              double acc, hssc_hss1d** out);
/* ---- rank prediction (HSS2D/predict_rank.m:5-23, interpolation_operator.m:101-117) --- */
/* A small network (8 -> 64 -> 64 -> 1, GELU, FP32) predicting log2 of a box's ID rank from
 * geometric features; trained and evaluated by the library's own GPU kernels.  It takes
 * the place of the closed-form rank guesses that size the randomized ID's sample count. */
typedef struct hssc_mlp hssc_mlp;
int hssc_mlp_create(uint64_t seed, hssc_mlp** out);
void hssc_mlp_destroy(hssc_mlp* m);
int hssc_mlp_num_params(void);
/* Flat parameters: W1 (8x64, column-major), b1 (64), W2 (64x64), b2 (64), w3 (64), b3. */
int hssc_mlp_get_params(const hssc_mlp* m, float* out);
int hssc_mlp_set_params(hssc_mlp* m, const float* in);
/* Standardisation: mu[0..7], sd[0..7] of the features, mu[8], sd[8] of the target. */
int hssc_mlp_set_scaling(hssc_mlp* m, const float* mu, const float* sd);
/* X: B x 8 column-major host array of raw features; y: B predicted targets. */
int hssc_mlp_predict(hssc_mlp* m, const float* X, int32_t B, float* y);
/* Mean squared error (standardised units) and its gradient w.r.t. the flat parameters. */
int hssc_mlp_grads(hssc_mlp* m, const float* X, const float* t, int32_t B, double* loss,
                   float* grads);
/* Full-batch Adam for `epochs` steps; loss = MSE of the last step (standardised units). */
int hssc_mlp_train(hssc_mlp* m, const float* X, const float* t, int32_t B, int32_t epochs,
                   float lr, double* loss);

/* HSS2D_build_tree_fsym (hssc_hss2d_build_tree) with rank-predicted randomized IDs for the
 * ID-selected skeletons (kh > 0, two bottom levels; HSS2D_build_tree_fsym.m:686-729), in
 * place of the closed-form guesses of HSS2D/predict_rank.m:5-23 and
 * interpolation_operator.m:101-117 that size the reference's randomized IDs.  rank_mode 0: the
 * plain ID (as hssc_hss2d_build_tree); 1: sketch rows from the closed-form guess; 2: from the
 * network `mlp` (features of hssc_tree_box_features).  A level's boxes share one Gaussian
 * sketch of s rows (the largest predicted rank plus `oversample`, rounded up to 16) applied by
 * one batched complex MFMA product; each sketched ID is judged on its full block by the check
 * of Common/ID.m:55-59 and redone by the plain ID when it fails. */
int hssc_hss2d_build_tree_ranked(const hssc_params* p, int32_t rank_mode, hssc_mlp* mlp,
                                 int32_t oversample, hssc_tree** tree);
/* Counters of the skeleton IDs of a tree: blocks, blocks redone by the plain ID, sketch rows
 * used and rows of the blocks factored by the plain ID. */
int hssc_tree_id_stats(const hssc_tree* t, int64_t* boxes, int64_t* fallbacks,
                       int64_t* sketch_rows, int64_t* full_rows);
/* Rank-network features of the key boxes of a level (nkey x 8, column-major): level, log2 n,
 * log2 M, kh times the box width, mean / min / max of the points' kernel weights, -log10 acc. */
int hssc_tree_box_features(const hssc_tree* t, int32_t lev, float* X);

/* HSS1D compression with rank-predicted randomized IDs.  rank_mode 0: the plain ID
 * (hssc_hss1d_compress); 1: sketch rows from the closed-form rank guess of
 * HSS2D/predict_rank.m:5-23; 2: from the network `mlp`.  Each block's sketch has the
 * predicted rank + oversample rows rounded up to 16; a sketch whose ID is saturated or
 * fails the acceptance test on the full block falls back to the plain ID, which remains
 * the judge. */
int hssc_hss1d_compress_ranked(const double* X, const double* nrm, const double* w,
                               const double* kappa, int64_t N, int32_t nbox_max, double acc,
                               int32_t nproxy, int32_t rank_mode, hssc_mlp* mlp,
                               int32_t oversample, hssc_hss1d** out);
/* Batched HSS1D (many curves in one operator): curve p has Ns[p] nodes, stored after the
 * nodes of curves 0..p-1 in X (sum Ns x 2), nrm, w, kappa.  Every curve gets its own tree;
 * the boxes of all curves at a level are compressed, inverted and swept by the same
 * launches.  apply / applyinv act on the concatenated vectors (length sum Ns).  Box
 * numbers: the roots are 1..nprob. */
int hssc_hss1d_compress_batched(int32_t nprob, const int64_t* Ns, const double* X,
                                const double* nrm, const double* w, const double* kappa,
                                int32_t nbox_max, double acc, int32_t nproxy, int32_t rank_mode,
                                hssc_mlp* mlp, int32_t oversample, hssc_hss1d** out);
/* Features of a non-root box b as fed to the rank network: level, log2 points,
 * log2 candidate skeleton size, log2(R_box / R_curve), max |kappa| R_box, mean |kappa|
 * R_box, log2 rows of the outgoing block, -log10 acc. */
int hssc_hss1d_box_features(const hssc_hss1d* h, int32_t box, float* f);
/* Counters of the IDs done by the compression: blocks, fallbacks to the plain ID, sketch
 * rows used and rows of the blocks factored by the plain ID. */
int hssc_hss1d_id_stats(const hssc_hss1d* h, int64_t* boxes, int64_t* fallbacks,
                        int64_t* sketch_rows, int64_t* full_rows);

/* Replaces HSS1D/HSS1D_invert_nsym.m:1 ([HSSinv,FTinv] = HSS1D_invert_nsym(K_HSS,acc)). */
int hssc_hss1d_invert(hssc_hss1d* h);
/* The same inverse factors in HSSC_PREC_FP64 (as hssc_hss1d_invert) or HSSC_PREC_FP32 (kept
 * beside the FP64 operator, no reference counterpart: the factorization of the mixed
 * precision solve). */
int hssc_hss1d_invert_prec(hssc_hss1d* h, int32_t precision);
/* Common/Iterative_Refinement.m:20-35 on an HSS1D operator: X = A^-1 F by the FP32 inverse
 * (the FP64 one if no FP32 inverse exists), refined against the FP64 compressed operator
 * until ||F - A X|| <= tol ||F|| or maxit solves.  Returns the solves done and the relative
 * residual; HSSC_ERR_NOT_CONVERGED if tol is not reached (X holds the last iterate). */
int hssc_hss1d_solve(const hssc_hss1d* h, const double* F, double* X, double tol,
                     int32_t maxit, int32_t* iters, double* relres);
int hssc_hss1d_solve_device(const hssc_hss1d* h, const void* dF, void* dX, double tol,
                            int32_t maxit, int32_t* iters, double* relres);
/* Replaces HSS1D/HSS1D_applyinv_nsym.m:1 and HSS1D/HSS1D_apply_nsym.m:1: m right-hand sides
 * (1 <= m <= 65535), N x m column-major (X = A^-1 F, U = A Q). */
int hssc_hss1d_applyinv(const hssc_hss1d* h, const double* F, double* X, int64_t m);
int hssc_hss1d_apply(const hssc_hss1d* h, const double* Q, double* U, int64_t m);
int hssc_hss1d_applyinv_device(const hssc_hss1d* h, const void* dF, void* dX, int64_t m);
int hssc_hss1d_apply_device(const hssc_hss1d* h, const void* dQ, void* dU, int64_t m);
/* Tree size and per-box skeleton data (the K_HSS cell rows LEVEL, NSKEL, KSKEL, I_SKUP,
 * I_SKDN of Common/box_constants.m:33-52; indices 0-based). */
int hssc_hss1d_info(const hssc_hss1d* h, int32_t* nboxes, int32_t* depth);
int hssc_hss1d_box(const hssc_hss1d* h, int32_t box, int32_t* level, int32_t* nskel,
                   int32_t* kskel, int64_t* isk_up, int64_t* isk_dn);
void hssc_hss1d_destroy(hssc_hss1d* h);

/* ---- device memory cache -----------------------------------------------------------
 * No counterpart in the reference (MATLAB owns its memory): device blocks released by the
 * library (tree workspaces, inverses, operators) are kept per device and reused by later
 * calls of the same sizes, up to HSSC_POOL_CAP_GB (environment, default 96) GiB.  A released
 * block is reused only after a device-wide synchronisation that follows its release.
 * hssc_release_cached_memory synchronises the current device and frees every cached block
 * (for callers that share the GPU with another allocator); hssc_cached_memory reports the
 * bytes held. */
int hssc_release_cached_memory(void);
int hssc_cached_memory(int64_t* bytes);

/* ---- instrumentation -------------------------------------------------------------
 * No counterpart in the reference (which times stages with tic/toc,
 * HSS2D/Test/Test_HSS2D_InverseCompression.m:44-50): when enabled, every launch of a
 * kernel class ("gemm_f64", "gemv_f64", "gj_small_f64", "gen_block_f64", "cqr_panel_f64",
 * ... and the _f32 variants) is bracketed by HIP events on its own stream, and its
 * algorithmic flops and bytes are accumulated.  hssc_stats_get synchronises the pending
 * events and returns the totals since the last reset. */
int hssc_stats_enable(int32_t on);
int hssc_stats_reset(void);
int hssc_stats_get(const char* kernel, int64_t* launches, double* flops, double* bytes,
                   double* ms);

#ifdef __cplusplus
}
#endif
#endif /* HSSC_H */
