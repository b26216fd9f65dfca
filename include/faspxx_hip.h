/* faspxx_hip.h -- C ABI of the MI355X (gfx950) iterative-solve path.
 *
 * Every entry point returns an int equal to a FaspRetCode value
 * (src/RetCode.hxx:24-55 in the reference): 0 on success, negative on error.
 * After an error, fx_last_error() returns a message for the calling thread.
 * No C++ exception crosses this boundary.
 *
 * Objects:
 *   fx_ctx     one GPU, one HIP stream (default priority), reduction scratch.
 *   fx_vec     a device vector of f64 or f32 (VEC, src/VEC.hxx:25-146).
 *   fx_csr     a CSRx matrix resident in HBM (MAT, src/MAT.hxx:39-213):
 *              u32 row pointers / column indices / diagonal pointers, f64
 *              values (plus an optional f32 copy), and the row-block table
 *              the SpMV kernels use.
 *   fx_pc      a preconditioner (Identity or Jacobi, src/Iter.hxx:26-71).
 *   fx_solver  a Krylov or Jacobi solver (SOL and subclasses, src/SOL.hxx).
 * Handles created on one fx_ctx are used on that context's stream only and
 * are not thread-safe, like the reference objects.  The caller owns every
 * handle and keeps a matrix and preconditioner alive while a solver uses
 * them (the reference's SOL also holds non-owning pointers, src/SOL.hxx:28-29).
 */
#ifndef FASPXX_HIP_H
#define FASPXX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fx_ctx_s* fx_ctx;
typedef struct fx_vec_s* fx_vec;
typedef struct fx_csr_s* fx_csr;
typedef struct fx_pc_s* fx_pc;
typedef struct fx_solver_s* fx_solver;
typedef struct fx_ir_s* fx_ir;
typedef struct fx_mcgs_s* fx_mcgs;
typedef struct fx_op_s* fx_op;

/* y = A x for a callback operator: called on the host with the context's
 * stream current; it enqueues its device work on that stream (or works on
 * host copies it synchronises itself) and returns 0 or a negative
 * FaspRetCode. */
typedef int (*fx_op_apply_fn)(void* user, fx_vec x, fx_vec y);

/* Value types of vectors and matrix values. */
enum { FX_F64 = 0, FX_F32 = 1 };

/* Return codes: the reference's FaspRetCode values (src/RetCode.hxx:24-55). */
enum {
  FX_SUCCESS = 0,
  FX_ERROR_INPUT_PAR = -12,
  FX_ERROR_VEC_SIZE = -14,
  FX_ERROR_MAT_SIZE = -15,
  FX_ERROR_NONMATCH_SIZE = -16,
  FX_ERROR_MAT_DATA = -17,
  FX_ERROR_DIVIDE_ZERO = -18,
  FX_ERROR_ALLOC_MEM = -20,
  FX_ERROR_SOLVER_TYPE = -30,
  FX_ERROR_SOLVER_PCD_TYPE = -31,
  FX_ERROR_SOLVER_STAG = -32,
  FX_ERROR_SOLVER_SOLSTAG = -33,
  FX_ERROR_SOLVER_TOLSMALL = -34,
  FX_ERROR_SOLVER_MAXIT = -39,
  FX_ERROR_DEVICE = -90 /* HIP runtime failure; no reference counterpart */
};

/* Solver types: the reference's SOLType values (src/Param.hxx:36-54). */
enum {
  FX_SOLVER_CG = 1,
  FX_SOLVER_BICGSTAB = 2,
  FX_SOLVER_GMRES = 4,
  FX_SOLVER_FGMRES = 5,
  FX_SOLVER_JACOBI = 11
};

/* SOLParams (src/Param.hxx:57-80) plus the GMRES restart bounds
 * (src/GMRES.hxx:40-42) and GMRES's left/right switch (src/GMRES.hxx:70-73).
 * fx_params_default() gives the reference defaults: maxIter 1, minIter 0,
 * savIter 100000, restart 30, relTol 1e-6, absTol 1e-8, maxRestart 30,
 * minRestart 10, right preconditioning. */
typedef struct {
  uint32_t maxIter, minIter, savIter, restart;
  double relTol, absTol;
  int verbose;
  uint32_t maxRestart, minRestart;
  int leftPrecond;
} fx_params;

/* What SOL::GetIterations / GetNorm2 / GetInfNorm report (src/SOL.cxx:199-205),
 * the return code of Solve, and the device time of the solve in ms. */
typedef struct {
  int retcode;
  uint32_t numIter;
  double norm2, normInf;
  double solve_ms;
} fx_result;

const char* fx_last_error(void);
void fx_params_default(fx_params* p);

/* ---- context --------------------------------------------------------- */
int fx_ctx_create(int device, fx_ctx* out);
/* Run on an existing HIP stream (e.g. the caller's framework stream). */
int fx_ctx_create_on_stream(int device, void* hip_stream, fx_ctx* out);
int fx_ctx_destroy(fx_ctx ctx);
int fx_ctx_sync(fx_ctx ctx);
void* fx_ctx_stream(fx_ctx ctx);

/* ---- VEC (src/VEC.hxx:25-146) --------------------------------------- */
/* VEC(size, 0.0), src/VEC.cxx:19-23 */
int fx_vec_create(fx_ctx ctx, uint32_t n, int dtype, fx_vec* out);
/* Non-owning view of caller device memory (n elements of dtype). */
int fx_vec_wrap(fx_ctx ctx, uint32_t n, int dtype, void* dev_ptr, fx_vec* out);
int fx_vec_destroy(fx_vec v);
uint32_t fx_vec_size(fx_vec v);
int fx_vec_dtype(fx_vec v);
void* fx_vec_data(fx_vec v);
/* VEC::SetValues(size, ptr), src/VEC.cxx:113-128 (host -> device) */
int fx_vec_upload(fx_vec v, const void* host);
/* VEC::GetArray + copy, src/VEC.cxx:144-148 (device -> host) */
int fx_vec_download(fx_vec v, void* host);
/* VEC::SetValues(size, value), src/VEC.cxx:99-103 */
int fx_vec_set(fx_vec v, double value);
/* VEC::CopyTo / operator=, src/VEC.cxx:51-57,210-214; converts f64<->f32 */
int fx_vec_copy(fx_vec src, fx_vec dst);

/* y += a x          VEC::AXPY, src/VEC.cxx:245-256 */
int fx_axpy(fx_ctx ctx, double a, fx_vec x, fx_vec y);
/* y = a y + x       VEC::XPAY, src/VEC.cxx:259-271 */
int fx_xpay(fx_ctx ctx, double a, fx_vec x, fx_vec y);
/* y = a y + b x     VEC::AXPBY, src/VEC.cxx:274-326 */
int fx_axpby(fx_ctx ctx, double a, double b, fx_vec x, fx_vec y);
/* w = a x + b y     VEC::WAXPBY, src/VEC.cxx:329-379 */
int fx_waxpby(fx_ctx ctx, double a, fx_vec x, double b, fx_vec y, fx_vec w);
/* y *= a            VEC::Scale, src/VEC.cxx:154-165 */
int fx_scale(fx_ctx ctx, double a, fx_vec y);
/* y .*= v           VEC::PointwiseMult, src/VEC.cxx:168-179 */
int fx_pmul(fx_ctx ctx, fx_vec v, fx_vec y);
/* y = 1 ./ y        VEC::Reciprocal, src/VEC.cxx:182-193 */
int fx_recip(fx_ctx ctx, fx_vec y);
/* y += v            VEC::operator+=, src/VEC.cxx:66-78 */
int fx_add(fx_ctx ctx, fx_vec v, fx_vec y);
/* y += a           VEC::Shift, src/VEC.cxx:217-228 */
int fx_shift(fx_ctx ctx, double a, fx_vec y);
/* y = |y|           VEC::Abs, src/VEC.cxx:231-242 */
int fx_abs(fx_ctx ctx, fx_vec y);
/* y ./= v           VEC::PointwiseDivide, src/VEC.cxx:196-207 */
int fx_pdiv(fx_ctx ctx, fx_vec v, fx_vec y);
/* y -= v            VEC::operator-=, src/VEC.cxx:81-93 */
int fx_sub(fx_ctx ctx, fx_vec v, fx_vec y);
/* max / min entry   VEC::Max / Min, src/VEC.cxx:382-421 (start values -1e60 / 1e60) */
int fx_max(fx_ctx ctx, fx_vec x, double* out);
int fx_min(fx_ctx ctx, fx_vec x, double* out);
/* Reductions returning a host double (synchronising), VEC::Dot / Norm2 /
 * NormInf, src/VEC.cxx:424-485.  Summation order is a fixed two-stage tree
 * (deterministic run to run); it differs from the reference's 4-way order. */
int fx_dot(fx_ctx ctx, fx_vec x, fx_vec y, double* out);
int fx_norm2(fx_ctx ctx, fx_vec x, double* out);
int fx_norminf(fx_ctx ctx, fx_vec x, double* out);
/* Same reductions into a device double (no synchronisation). */
int fx_dot_dev(fx_ctx ctx, fx_vec x, fx_vec y, double* dev_out);
int fx_norm2_dev(fx_ctx ctx, fx_vec x, double* dev_out);
/* Fused: y += a x, then *out = (y, y)^(1/2)  (one pass) */
int fx_axpy_norm2(fx_ctx ctx, double a, fx_vec x, fx_vec y, double* out);
/* Fused: y = a y + x (XPAY), then *out = (y, y)^(1/2)  (one pass) */
int fx_xpay_norm2(fx_ctx ctx, double a, fx_vec x, fx_vec y, double* out);
/* Fused: y += a x, then *out = (y, z)        (one pass) */
int fx_axpy_dot(fx_ctx ctx, double a, fx_vec x, fx_vec y, fx_vec z, double* out);

/* ---- MAT (src/MAT.hxx:39-213) ---------------------------------------- */
/* MAT(nrow, mcol, nnz, values, colInd, rowPtr[, diagPtr]),
 * src/MAT.cxx:22-94.  Host arrays.  values == NULL gives a structure-only
 * matrix; diagPtr == NULL forms it like MAT::FormDiagPtr (src/MAT.cxx:917-928).
 * The structure is validated as CheckCSR does (src/MATUtil.cxx:110-174):
 * rowPtr[0] == 0, rowPtr non-decreasing, rowPtr[nrow] == nnz, every column
 * index < mcol.  nrow, mcol or nnz equal to 0 gives an empty matrix. */
int fx_csr_create(fx_ctx ctx, uint32_t nrow, uint32_t mcol, uint32_t nnz,
                  const uint32_t* rowPtr, const uint32_t* colInd, const double* values,
                  const uint32_t* diagPtr, fx_csr* out);
int fx_csr_destroy(fx_csr A);
int fx_csr_info(fx_csr A, uint32_t* nrow, uint32_t* mcol, uint32_t* nnz);
/* Bytes per column index that the SpMV kernels read: 0 for row patterns
 * (a matrix created from host arrays whose rows repeat at most 255
 * column-offset patterns of at most 32 entries, with at most 512 offsets in
 * all: one pattern byte per row instead), 2 when every row block's columns
 * lie within 32K of the block's first row (16-bit deltas), else 4.
 * Environment at creation: FASPXX_FX_NOPAT=1 disables row patterns,
 * FASPXX_FX_COL32=1 keeps 4.  All forms give the same bits.  No reference
 * counterpart (the reference stores u32 columns only, src/MAT.hxx:43-47). */
int fx_csr_index_bytes(fx_csr A, int* bytes);
/* How the SpMV kernels read the matrix: FX_LAYOUT_U32 (row blocks, u32
 * columns), FX_LAYOUT_D16 (row blocks, 16-bit column deltas),
 * FX_LAYOUT_PATTERN (row patterns), FX_LAYOUT_SORTED (u32 row blocks over a
 * copy of the matrix whose rows are in order of decreasing length, each row
 * written to its own slot: the default for a matrix made from host arrays
 * whose columns suit neither row patterns nor 16-bit deltas;
 * FASPXX_FX_NOSORT=1 at creation turns it off), FX_LAYOUT_SLICED (for the
 * f64 kernels, windows of 256 rows whose rows of at most 32 entries are
 * sorted by length into 64-row slices stored column-major, longer rows a
 * wavefront or the workgroup each; the f32 kernels keep u32 row blocks: used
 * with FASPXX_FX_NOSORT=1, or for every matrix made from host arrays with
 * FASPXX_FX_SELL=1; FASPXX_FX_NOSELL=1 turns it off).  A row summed by one
 * thread, left to right (in the sliced layout every row of at most 32
 * entries; in row blocks the rows of blocks averaging at most 32 entries per
 * row), gives the same bits in every layout; rows summed by lane groups
 * agree to rounding.  No reference counterpart. */
#define FX_LAYOUT_U32 0
#define FX_LAYOUT_D16 1
#define FX_LAYOUT_PATTERN 2
#define FX_LAYOUT_SLICED 3
#define FX_LAYOUT_SORTED 4
int fx_csr_layout(fx_csr A, int* layout);
/* Copy the device arrays back (any pointer may be NULL). */
int fx_csr_download(fx_csr A, uint32_t* rowPtr, uint32_t* colInd, double* values,
                    uint32_t* diagPtr);
/* Replace the f64 values from device memory (nnz doubles) on the context
 * stream; refreshes the f32 copy if one exists. */
int fx_csr_set_values_dev(fx_csr A, const double* dev_values);
/* Device pointer to the f64 values (nnz doubles). */
double* fx_csr_values_dev(fx_csr A);
/* Build (or rebuild) the f32 value copy used by f32 SpMV / mixed precision. */
int fx_csr_make_f32(fx_csr A);
/* w = A v  MAT::Apply, src/MAT.cxx:373-452.  f64 vectors use the f64
 * values; f32 vectors use the f32 copy (fx_csr_make_f32). */
int fx_spmv(fx_ctx ctx, fx_csr A, fx_vec v, fx_vec w);
/* fx_spmv with HIP events attached to the kernel's own dispatch; waits for
 * it and returns its duration in *ms.  A measurement aid with no reference
 * counterpart (the reference times whole solves, example/TestCG.cxx:86-88). */
int fx_spmv_timed(fx_ctx ctx, fx_csr A, fx_vec v, fx_vec w, float* ms);
/* r = b - A x, or r = b if ||x||_inf < 1e-20   MAT::Residual, src/MAT.cxx:455-476 */
int fx_residual(fx_ctx ctx, fx_csr A, fx_vec b, fx_vec x, fx_vec r);
/* v = diag(A)  MAT::GetDiag, src/MAT.cxx:256-267 */
int fx_csr_get_diag(fx_ctx ctx, fx_csr A, fx_vec v);
/* *At = A^T    MAT::TransInPlace, src/MAT.cxx:479-538 (new handle; columns
 * ascending in each row; diagPtr formed as FormDiagPtr does). */
int fx_csr_transpose(fx_ctx ctx, fx_csr A, fx_csr* At);

/* ---- setup-time sparse algebra: each result is a new handle -----------
 * Lower / upper triangular part by position: row i keeps its entries up to
 * and including its diagonal entry, or from it to the row's end; every row
 * needs a diagonal entry (FX_ERROR_MAT_DATA otherwise).
 *   MAT::GetLowerTri / GetUpperTri, src/MAT.cxx:290-345 */
int fx_csr_lower_tri(fx_ctx ctx, fx_csr A, fx_csr* L);
int fx_csr_upper_tri(fx_ctx ctx, fx_csr A, fx_csr* U);
/* *C = a A + b B: per row the union of the patterns, columns ascending; an
 * entry in both is fma(b, B_ij, a A_ij).  Both need values and one shape.
 * With one operand empty the result is the other, scaled (the reference
 * then leaves its target unchanged; the drop-in MAT::Add keeps that).
 *   MAT::Add, src/MAT.cxx:652-719 */
int fx_csr_add(fx_ctx ctx, double a, fx_csr A, double b, fx_csr B, fx_csr* C);
/* *C = L R: row i holds the distinct columns reached through row i of L in
 * order of first appearance (not sorted), values the fma chain of the
 * products in that order.  Structure-only operands give a structure-only
 * product; more than 2^31 candidate products is FX_ERROR_ALLOC_MEM.
 *   MAT::Mult, src/MAT.cxx:722-800 */
int fx_csr_mult(fx_ctx ctx, fx_csr L, fx_csr R, fx_csr* C);
/* values[diagPtr[j]] += a in place for every row j, as the reference's loop
 * over diagPtr: a row without a diagonal entry has diagPtr 0 and adds a to
 * values[0] (once per such row).   MAT::Shift, src/MAT.cxx:359-364 */
int fx_csr_shift_diag(fx_ctx ctx, fx_csr A, double a);

/* Refresh the values of At (made by fx_csr_transpose from A) from A's
 * current values, without recomputing the structure.  No reference
 * counterpart: the adjoint solve's A^T after A's values change. */
int fx_csr_transpose_update(fx_ctx ctx, fx_csr A, fx_csr At);
/* out_k = -lam[row(k)] * x[col(k)] over A's stored entries (nnz doubles,
 * device memory): dL/dA on the sparsity pattern, from the adjoint solution
 * lam of A^T lam = dL/dx.  No reference counterpart. */
int fx_csr_pattern_grad(fx_ctx ctx, fx_csr A, fx_vec lam, fx_vec x, double* dev_out);

/* ---- preconditioners (src/Iter.hxx:26-71) ---------------------------- */
/* Identity::Solve, src/Iter.cxx:16-20 */
int fx_pc_create_identity(fx_ctx ctx, fx_pc* out);
/* Jacobi with SetWeight(weight), minIter = maxIter = iters, Setup(A):
 * dinv = weight * (1 / diag(A)), src/Iter.cxx:29-53. */
int fx_pc_create_jacobi(fx_ctx ctx, fx_csr A, double weight, uint32_t iters, fx_pc* out);
int fx_pc_destroy(fx_pc pc);
/* x <- P(b) with x as the initial guess: Jacobi::Solve, src/Iter.cxx:56-112 */
int fx_pc_apply(fx_ctx ctx, fx_pc pc, fx_vec b, fx_vec x);
/* No counterpart in the reference: fx_pc_apply for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same result bit for bit,
 * with no allocation: identity and Jacobi preconditioners, f64 vectors, at
 * most 4,194,304 rows (FX_ERROR_INPUT_PAR beyond).  Its work vector and
 * reduction scratch live in one library-owned device array per GPU (32 MiB,
 * held until the process ends) that every call on that GPU shares: calls
 * that overlap on one GPU race. */
int fx_pc_apply_noalloc(fx_ctx ctx, fx_pc pc, fx_vec b, fx_vec x);

/* ---- Multicolor Gauss-Seidel smoother --------------------------------------
 * No reference counterpart: FASP++ has no Gauss-Seidel, and a sequential one
 * has no parallel form, so this separately validated smoother is offered
 * instead of a drop-in.  Colors: rows in natural order take the smallest
 * color unused by already-colored neighbours in the pattern of A + A^T
 * (stencils come out red-black).  One color is one parallel pass:
 * r_i = b_i - sum_k a_ik x_k (row order), x_i = fma(r_i, weight / a_ii, x_i).
 * A sweep runs colors 0..C-1, then C-1..0 when symmetric.  f64 only. */
int fx_mcgs_create(fx_ctx ctx, fx_csr A, double weight, fx_mcgs* out);
/* Number of colors, and (if color != NULL) the color of each of the nrow rows. */
int fx_mcgs_info(fx_mcgs s, uint32_t* ncolors, uint32_t* color);
/* `sweeps` sweeps updating x in place (b and x distinct). */
int fx_mcgs_sweep(fx_mcgs s, fx_vec b, fx_vec x, uint32_t sweeps, int symmetric);
int fx_mcgs_destroy(fx_mcgs s);
/* Preconditioner z = P r: `sweeps` symmetric multicolor Gauss-Seidel sweeps
 * on A z = r from the z given (the solvers pass z = 0), usable by every
 * solver in place of Jacobi (solvers with it run the per-operation path). */
int fx_pc_create_mcgs(fx_ctx ctx, fx_csr A, double weight, uint32_t sweeps, fx_pc* out);

/* ---- solvers --------------------------------------------------------- */
/* new CG / BiCGStab / GMRES / FGMRES / Jacobi (src/Krylov.cxx:21-40) */
int fx_solver_create(fx_ctx ctx, int type, fx_solver* out);
/* SetupPCD(pc) + setters + Setup(A): src/SOL.cxx:249, src/CG.cxx:16-40 etc.
 * For FX_SOLVER_JACOBI, pc must be a Jacobi preconditioner; its weight is
 * the solver's weight and its iteration count is ignored.  As Setup does,
 * it resets the GMRES / FGMRES restart length (src/GMRES.cxx:44,75-79). */
int fx_solver_setup(fx_solver s, fx_csr A, fx_pc pc, const fx_params* prm);
/* A, pc and params as the reference's Solve finds them without a new Setup
 * (SOL keeps pointers to A and pcd, src/SOL.hxx:28-29, and setters change
 * params): bound again, while GMRES / FGMRES keep the restart length the
 * last Solve ended with (the member `restart`, adapted per cycle,
 * src/GMRES.cxx:255-265) and FGMRES its Z (src/FGMRES.cxx:158-159). */
int fx_solver_update(fx_solver s, fx_csr A, fx_pc pc, const fx_params* prm);
/* Clean(): the GMRES / FGMRES restart length back to the set-up value and
 * FGMRES's kept Z zeroed (src/GMRES.cxx:90-110, src/FGMRES.cxx:83-105). */
int fx_solver_clean(fx_solver s);
/* A linear operator given by its action: a LOP subclass with its own Apply
 * (src/LOP.hxx:22-66; the matrix-free pattern of example/TestMatFree.cxx).
 * The Krylov solvers (CG, BiCGStab, GMRES, FGMRES) accept it through
 * fx_solver_setup_op with any preconditioner and run their per-operation
 * form, calling apply wherever the reference calls A->Apply.  The Jacobi
 * solver needs a matrix (its diagonal), as in the reference. */
int fx_op_create(fx_ctx ctx, uint32_t nrow, uint32_t mcol, fx_op_apply_fn apply, void* user,
                 fx_op* out);
/* Declares that op's apply only enqueues device work on the context's stream
 * (no host synchronisation, allocation or host-side reads of device data)
 * and enqueues the same work on every call, e.g. a device stencil kernel
 * for the MatFree::Apply of example/TestMatFree.cxx:86-176.  The Krylov
 * solvers then run their fused form on it (the identity or a one-sweep
 * Jacobi preconditioner): apply is called while replay graphs are captured
 * (on the first Solve after fx_solver_setup_op / fx_solver_update_op, and
 * on a later Solve for each shorter last replay it first needs), and its
 * device work is replayed from then on.  No counterpart in the reference,
 * whose LOP::Apply is a host call. */
int fx_op_set_device_only(fx_op op, int device_only);
int fx_op_destroy(fx_op op);
int fx_solver_setup_op(fx_solver s, fx_op A, fx_pc pc, const fx_params* prm);
/* fx_solver_update for an operator (see fx_solver_update). */
int fx_solver_update_op(fx_solver s, fx_op A, fx_pc pc, const fx_params* prm);
/* Solve(b, x): x is the initial guess on entry (CG::Solve, src/CG.cxx:53-249;
 * BiCGStab::Solve, src/BiCGStab.cxx:68-270; GMRES::Solve, src/GMRES.cxx:16-22;
 * FGMRES::Solve, src/FGMRES.cxx:108-271; Jacobi::Solve, src/Iter.cxx:56-112).
 * The return value equals result->retcode. */
int fx_solver_solve(fx_solver s, fx_vec b, fx_vec x, fx_result* result);
/* Residual norms the reference prints per iteration (SOL::PrintInfo,
 * src/SOL.cxx:46-64), index = iteration number. */
int fx_solver_history(fx_solver s, double* out, uint32_t cap, uint32_t* count);
/* Which form ran the last Solve: FX_PATH_FUSED (graph-replayed fused
 * kernels, scalars on the device) or FX_PATH_OPS (one device call per
 * reference operation, scalars read back).  No reference counterpart (the
 * reference has one CPU path). */
#define FX_PATH_OPS 0
#define FX_PATH_FUSED 1
int fx_solver_last_path(fx_solver s, int* path);
int fx_solver_destroy(fx_solver s);

/* Free function Krylov(A, b, x, pcd, params), src/Krylov.cxx:16-53: name
 * (case-insensitive "cg", "bicgstab", "gmres", "fgmres"; anything else -> CG)
 * -> solver, Setup, SetupPCD, Solve. */
int fx_krylov(fx_ctx ctx, fx_csr A, fx_vec b, fx_vec x, fx_pc pc, const char* algName,
              const fx_params* prm, fx_result* result);

/* ---- mixed-precision iterative refinement (no reference counterpart) --
 * Outer loop in f64: r = b - A x (f64 values), stop when ||r|| / ||r0|| <
 * outer->relTol or ||r|| < outer->absTol or after outer->maxIter steps;
 * inner solve A d = r in f32 values and vectors by `inner_type` (CG, or
 * GMRES with f64 Hessenberg / Givens) with inner->relTol, from d = 0,
 * Jacobi(weight) preconditioned; x += d in f64.  result->numIter is the
 * number of outer steps; *inner_iters (optional) the summed inner count.
 * CG inner, every a_ii > 0: the inner solve is CG on S A S with
 * S = diag(a_ii)^-1/2 (Jacobi-PCG in exact arithmetic; the weight does not
 * change CG's iterates), in the single-reduction form of Chronopoulos and
 * Gear, stopped when ||S r_k|| < max(inner->relTol ||S r_0||,
 * inner->absTol) or after inner->maxIter iterations; then d = S y.  Other
 * matrices, or FASPXX_FX_IR_REFCG=1 at creation, use the f32 Jacobi-PCG
 * with the reference CG's checks. */
int fx_solve_ir(fx_ctx ctx, int inner_type, fx_csr A, double weight, const fx_params* outer,
                const fx_params* inner, fx_vec b, fx_vec x, fx_result* result,
                uint32_t* inner_iters);
/* The same as a reusable object: create allocates the workspace (the f32
 * values of S A S and S, or an f32 value copy and the Jacobi diagonal;
 * inner-solver buffers); solve allocates nothing except the replay graphs
 * it makes on first use: the full replays on the first call, and a shorter
 * last replay per predicted length on later calls.  Values of A changed
 * after creation are not seen by the inner solves. */
int fx_ir_create(fx_ctx ctx, int inner_type, fx_csr A, double weight, const fx_params* outer,
                 const fx_params* inner, fx_ir* out);
int fx_ir_solve(fx_ir ir, fx_vec b, fx_vec x, fx_result* result, uint32_t* inner_iters);
int fx_ir_destroy(fx_ir ir);

#ifdef __cplusplus
}
#endif

#endif /* FASPXX_HIP_H */
