/*
 * libmfdfd — MI355X (gfx950) implementation of the MaxwellFDFD frequency-domain
 * hot path behind a plain C ABI: Yee-grid curl and material-operator assembly
 * with PML-stretched ∆l, material assignment with subpixel smoothing, the
 * matrix-free curl-curl operator, and the iterative solve of
 *     (Cₘ μ⁻¹ Cₑ − ω² ε) e = −iω jₑ − Cₘ μ⁻¹ jₘ.
 *
 * Conventions follow the reference package (EMinsight/MaxwellFDM.jl, module
 * MaxwellFDFD, src/model/model.jl:32-81):
 *   - complex values are two interleaved doubles (Julia ComplexF64);
 *   - CSC outputs are Julia SparseMatrixCSC{ComplexF64,Int64}: 1-based colptr
 *     (length n+1) and rowval (length nnz), rows ascending within a column;
 *   - parameter arrays εarr/μarr are (Nx,Ny,Nz,3,3) column-major, field arrays
 *     (Nx,Ny,Nz,3) column-major; DOF vectors are component-first
 *     (order_cmpfirst) or component-last;
 *   - every function returns MFD_OK or an error code; mfd_last_error() gives
 *     a message for the calling thread.  Nothing keeps caller pointers after
 *     returning and nothing calls back into the caller.
 *   - host pointers unless the name ends in _dev (device pointers on the
 *     context's GPU, in the context's HIP stream order after a full sync).
 */
#ifndef MFDFD_H
#define MFDFD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    MFD_OK = 0,
    MFD_ERR_ARG = 1,          /* Julia ArgumentError analogue (source.jl:214,230) */
    MFD_ERR_UNSUPPORTED = 2,  /* model.jl:241-242 (ft not EE/HH), :236 (Pμ not diagonal) */
    MFD_ERR_HIP = 3,
    MFD_ERR_NOCONV = 4,
    MFD_ERR_ALLOC = 5
};
enum { MFD_EE = 0, MFD_HH = 1 };      /* FieldType */
enum { MFD_PRIM = 0, MFD_DUAL = 1 };  /* GridType */
enum { MFD_FP64 = 0, MFD_MIXED = 1 }; /* solve precision */
/* Krylov method: SQMR and COCG (with reliable updates) need the complex-symmetric
 * operator of a grid without complex Bloch phases; BiCGSTAB serves the rest. */
enum { MFD_SQMR = 0, MFD_COCG = 1, MFD_BICGSTAB = 2 };

typedef struct { double re, im; } mfd_c64;

/* A Julia SparseMatrixCSC{ComplexF64,Int64} as it lies in memory: m × n with
 * nnz stored entries, 1-based colptr (n + 1 values, colptr[0] = 1,
 * colptr[n] = nnz + 1) and rowval (ascending within a column), nzval. */
typedef struct {
    int64_t m, n, nnz;
    const int64_t *colptr;
    const int64_t *rowval;
    const mfd_c64 *nzval;
} mfd_csc;

/* Grid + Model boundary settings (MaxwellBase Grid(lprim, isbloch) and the
 * Model fields boundft/Npml/kbloch/ωpml/order_cmpfirst, model.jl:37-72). */
typedef struct {
    int64_t N[3];
    const double *lprim[3];   /* N[w]+1 primal plane locations, both bounds */
    int32_t isbloch[3];
    int32_t boundft[3];       /* MFD_EE / MFD_HH per axis */
    int64_t npml[2][3];       /* Npml[negative/positive end][axis] */
    mfd_c64 omega_pml;        /* ωpml (set_ωpml!, model.jl:86) */
    double kbloch[3];         /* set_kbloch!, model.jl:89 */
    double pml_m;             /* SC-PML polynomial degree (4) */
    double pml_lnR;           /* SC-PML ln(target reflection) (-16) */
    int32_t order_cmpfirst;   /* model.jl:72 */
} mfd_grid;

/* Shapes of add_obj! (model.jl:107-120): kind 0 = ball (c, r), 1 = cuboid (c, s),
 * 2 = cylinder along z (c, r, height s[2]; a 2D disk extruded). */
typedef struct {
    int32_t kind;
    int32_t pad;
    double c[3];
    double r;
    double s[3];
} mfd_shape;

typedef struct mfd_ctx mfd_ctx;
typedef struct mfd_op mfd_op;

typedef struct {
    double tol;               /* relative residual ||b-Ae||/||b|| in FP64 (1e-8) */
    int64_t maxit;            /* total Krylov iterations allowed per RHS */
    int32_t precision;        /* MFD_FP64 or MFD_MIXED (FP32 inner, FP64 refinement) */
    int32_t method;           /* MFD_SQMR, MFD_COCG, MFD_BICGSTAB, or -1 (SQMR; BiCGSTAB with a complex Bloch phase) */
    double inner_tol;         /* mixed: tightest inner reduction per FP32 cycle (1e-5) */
    int32_t check_every;      /* iterations between host convergence checks (32) */
    int32_t use_graph;        /* capture a chunk of iterations in a hipGraph */
} mfd_solve_opts;

typedef struct {
    int64_t iterations;       /* Krylov iterations (max over RHS) */
    int64_t outer_cycles;     /* refinement cycles (SQMR) / reliable updates (COCG) */
    double relres;            /* max over RHS of the final FP64 relative residual */
    double seconds;           /* device time of the solve */
    int32_t converged;
    int64_t iterations_fp32;  /* of `iterations`, those run in complex-FP32 inner cycles */
} mfd_solve_stats;

const char *mfd_version(void);
const char *mfd_last_error(void);

/* One GPU, one HIP stream.  Not thread-safe: the caller serialises. */
int mfd_ctx_create(int device, mfd_ctx **out);
int mfd_ctx_destroy(mfd_ctx *ctx);

/* Stretched ∆l tables, replaces MaxwellBase create_stretched_∆l called at
 * model.jl:126.  out: [gt = PRIM, DUAL][axis 0..2][N[axis]] complex. */
int mfd_stretched_dl(mfd_ctx *ctx, const mfd_grid *g, mfd_c64 *out);

/* distweights (source.jl:153-154, 208-306): 1-based indices as in Julia.
 * Returns MFD_ERR_ARG where the reference throws ArgumentError. */
int mfd_distweights(double c, int gt, const double bounds[2], const double *l, const double *dl,
                    int64_t N, int isbloch, int64_t ind[2], double wt[2]);

/* add_srce!/add_srcm! with a PointSrc (model.jl:184-200, pointsrc.jl:64-103).
 * p is normalised here as PointSrc's constructor does (pointsrc.jl:58).
 * jarr: (Nx,Ny,Nz,3) complex, accumulated into. */
int mfd_add_pointsrc(const mfd_grid *g, int ft, const double c[3], const double p[3], mfd_c64 Idr,
                     mfd_c64 *jarr);
/* add_src!(…, ::PlaneSrc) (planesrc.jl:41-79, via add_srce!/add_srcm!,
 * model.jl:184-200): plane normal to axis nn (0..2) at intercept c, current
 * J∆n along p (normalised here), added into jarr (Nx,Ny,Nz,3) column-major. */
int mfd_add_planesrc(const mfd_grid *g, int ft, int nn, double c, const double p[3], mfd_c64 Jdn, mfd_c64 *jarr);

/* create_curl (model.jl:171-172): which 0 = Cₑ (E -> H), 1 = Cₘ (H -> E).
 * Two phases: query nnz, then fill caller-allocated buffers. */
int mfd_curl_nnz(mfd_ctx *ctx, const mfd_grid *g, int which, int64_t *nnz);
int mfd_curl_csc(mfd_ctx *ctx, const mfd_grid *g, int which, int64_t *colptr, int64_t *rowval,
                 mfd_c64 *nzval);

/* create_Mcs (model.jl:287-306): which 0 = Mcₑ, 1 = Mcₘ, the operators that
 * interpolate the solution fields to the voxel corners (MaxwellBase
 * create_mean with isfwd = boundft .!= EE / .!= HH): F_w goes to the w slot
 * of its corners by the ∆l-weighted average along w that Pε/Pμ's off-diagonal
 * blocks also use.  Same two phases as the curls. */
int mfd_Mcs_nnz(mfd_ctx *ctx, const mfd_grid *g, int which, int64_t *nnz);
int mfd_Mcs_csc(mfd_ctx *ctx, const mfd_grid *g, int which, int64_t *colptr, int64_t *rowval,
                mfd_c64 *nzval);

/* calc_matparams! (full.jl:16-70): assignment + Kottke subpixel smoothing.
 * eps_t/mu_t: nobj 3x3 column-major tensors; objects later in the list win. */
int mfd_matparams(mfd_ctx *ctx, const mfd_grid *g, int64_t nobj, const mfd_shape *shapes,
                  const mfd_c64 *eps_t, const mfd_c64 *mu_t, mfd_c64 *eps_arr, mfd_c64 *mu_arr);

/* create_paramop (model.jl:152-155): which 0 = Pε, 1 = Pμ. */
int mfd_paramop_nnz(mfd_ctx *ctx, const mfd_grid *g, int which, const mfd_c64 *parr, int64_t *nnz);
int mfd_paramop_csc(mfd_ctx *ctx, const mfd_grid *g, int which, const mfd_c64 *parr,
                    int64_t *colptr, int64_t *rowval, mfd_c64 *nzval);

/* create_A (model.jl:225-246): A = Cₘ (Pμ \ Cₑ) − ω² Pε for ft = EE,
 * A = Cₑ (Pε \ Cₘ) − ω² Pμ for ft = HH (the divided operator must be diagonal). */
int mfd_A_nnz(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr,
              const mfd_c64 *mu_arr, int64_t *nnz);
int mfd_A_csc(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr,
              const mfd_c64 *mu_arr, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval);

/* create_A(ft, ω, Pε, Pμ, Cₑ, Cₘ) (model.jl:222-246) from the caller's
 * assembled matrices, formed on the device: A = Cₘ (Pμ \ Cₑ) − ω² Pε (EE) or
 * Cₑ (Pε \ Cₘ) − ω² Pμ (HH), with Julia SparseArrays' structure rules: the
 * divisor must be diagonal (MFD_ERR_UNSUPPORTED otherwise, as model.jl:236
 * notes), the product keeps its structural entries, and the subtraction keeps
 * only nonzero results (skipped for ω = 0).  The same two phases as mfd_A_csc;
 * the results equal mfd_A_csc's for the matrices mfd_paramop_csc and
 * mfd_curl_csc return. */
int mfd_A_from_csc_nnz(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
                       const mfd_csc *Cm, int64_t *nnz);
int mfd_A_from_csc(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
                   const mfd_csc *Cm, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval);
/* create_b(ft, ω, Pε, Pμ, Cₑ, Cₘ, jₑ, jₘ) (model.jl:248-274) from the
 * matrices: b = −Cₘ (Pμ \ jₘ) − iω jₑ (EE) or Cₑ (Pε \ jₑ) − iω jₘ (HH). */
int mfd_b_from_csc(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
                   const mfd_csc *Cm, const mfd_c64 *je, const mfd_c64 *jm, mfd_c64 *b);
/* h_from_e(e, ω, Pμ, Cₑ, jₘ) = (i/ω) (Pμ \ (Cₑ e + jₘ)) and
 * e_from_h(h, ω, Pε, Cₘ, jₑ) = (−i/ω) (Pε \ (Cₘ h − jₑ)) (model.jl:276-284)
 * from the matrices; jₘ / jₑ may be null (zero). */
int mfd_h_from_e_csc(mfd_ctx *ctx, const mfd_c64 *e, mfd_c64 omega, const mfd_csc *Pm, const mfd_csc *Ce,
                     const mfd_c64 *jm, mfd_c64 *h);
int mfd_e_from_h_csc(mfd_ctx *ctx, const mfd_c64 *h, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Cm,
                     const mfd_c64 *je, mfd_c64 *e);

/* create_b (model.jl:251-274): b = −Cₘ (Pμ \ jₘ) − iω jₑ for ft = EE (uses
 * mu_arr), b = Cₑ (Pε \ jₑ) − iω jₘ for ft = HH (uses eps_arr); any boundft. */
int mfd_create_b(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr,
                 const mfd_c64 *mu_arr, const mfd_c64 *je, const mfd_c64 *jm, mfd_c64 *b);

/* Matrix-free operator for the A of create_A (model.jl:225-246), device
 * resident.  eps_arr/mu_arr as produced by create_paramops (host).
 * ft = EE: A = Cₘ (Pμ \ Cₑ) − ω² Pε; ft = HH: A = Cₑ (Pε \ Cₘ) − ω² Pμ, built as
 * the identical EE operator of the same grid with every boundft flipped and
 * ε, μ swapped.  Any boundft; the tiled stencils serve every operator but
 * one with both μ ≠ 1 and off-diagonal ε, which uses the direct kernel
 * (boundft = HH axes run on the grid reflected along them).  Off-diagonal
 * ε (EE) needs boundft = EE on every axis, off-diagonal μ (HH) boundft = HH. */
int mfd_op_create(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr,
                  const mfd_c64 *mu_arr, mfd_op **out);
/* Operator from the caller's assembled A (the A of create_A / create_linsys,
 * model.jl:209-246; n = 3 × cells): a row-major device copy applied by a
 * sparse product, Jacobi-scaled, for mfd_op_apply(_dev) and the solvers.  If A
 * equals its transpose it is complex-symmetric (SQMR by default, COCG, the
 * adjoint solve); otherwise BiCGSTAB serves it (A from a PML grid is symmetric
 * only up to the cell volumes, which A alone does not carry).  The matrix-free
 * mfd_op_create is the fast form of the same operator.  Tile and field
 * conversion calls do not apply to it (MFD_ERR_UNSUPPORTED). */
int mfd_op_create_csc(mfd_ctx *ctx, const mfd_csc *A, mfd_op **out);
int mfd_op_destroy(mfd_op *op);
/* y = A x for nrhs vectors stored one after another (host pointers). */
int mfd_op_apply(mfd_op *op, const mfd_c64 *x, mfd_c64 *y, int64_t nrhs);
/* device pointers; prec MFD_FP64 (complex double) or MFD_MIXED (complex float) */
int mfd_op_apply_dev(mfd_op *op, const void *x, void *y, int64_t nrhs, int prec);
/* Inner operator of the solver: the grad-div augmented A' = A − ε G W Div ε
 * (W = 1/ε̄² at the primal nodes) with right-hand side r + ω⁻² ε G W Div r.
 * Div Cₘ = 0 makes every solution of A e = b a solution of the augmented
 * system, so the discretisation and the FP64 residual checked are unchanged;
 * the longitudinal near-null space of Cₘ Cₑ is moved away from zero.  Off by
 * default (opt-in; eligible with diagonal ε, μ = 1, ω ≠ 0).  No reference
 * counterpart: the reference has no solver (model.jl stops at create_A). */
int mfd_op_set_graddiv(mfd_op *op, int enable);
/* Unknowns of the 2D and 1D models (replaces the 3D-embedded solve of
 * ModelTE te.jl:4-13, E = (Ex, Ey): c0 = 0, nc = 2; ModelTM tm.jl:4-13,
 * E = Ez: c0 = 2, nc = 1; ModelTEM tem.jl:4-13, E = Ex: c0 = 0, nc = 1):
 * restrict the operator of their single-cell 3D model to the E components
 * c0 .. c0+nc−1, i.e. to the model's own create_A, A[w,w].  Afterwards every
 * apply and solve takes vectors of nc·M values in the component-last order
 * (p + M·k for component c0 + k), and the solver moves nc/3 of the bytes.
 * Needs an operator created with order_cmpfirst = 0, μ = 1 and boundft = EE
 * on every axis (MFD_ERR_UNSUPPORTED otherwise), grad-div off; h_from_e,
 * e_from_h, the adjoint and TY = 0 then return MFD_ERR_UNSUPPORTED.
 * c0 = 0, nc = 3 restores the full vectors. */
int mfd_op_set_components(mfd_op *op, int c0, int nc);
/* Tile-shape autotuner for the stencil (rows per block TY, z-planes per block
 * KZ, tile order: bit 0 set = each XCD gets a contiguous run of tiles, clear =
 * dispatch order; bit 1 set = 32-wide tiles of 2·TY rows, a candidate when
 * ε is diagonal and the x size is not a multiple of 64): times each
 * candidate inside complete solver iterations for nrhs
 * right-hand sides (iters of them, on the solver workspace) and keeps the one
 * with the fastest iteration per precision; ms[] is that iteration time.
 * Outputs may be null.  No reference counterpart. */
int mfd_op_autotune(mfd_op *op, int64_t nrhs, int iters, int tile64[3], int tile32[3], double ms[2]);
/* Fix the tile shape by hand: TY in {1,2,4,8} (0 = cache-reliant direct
 * kernel; MFD_ERR_UNSUPPORTED with a component window), KZ >= 1, xcd_order
 * in {0,1,2,3} (bit 0: XCD tile order; bit 1: 32-wide tiles of 2·TY rows,
 * used by the diagonal-ε, μ = 1 stencil of an EE-boundft operator and
 * ignored by the others).  Until this or mfd_op_autotune is called, an
 * operator uses TY = 4, KZ = 16, XCD order, with KZ and then TY halved (down
 * to 4 and 1) while a launch would have fewer than 256 tiles, TY = 8 for
 * full ε on a one-cell z axis where that still gives 256 tiles, and 32-wide
 * tiles (xcd_order 3) when that stencil serves and they leave fewer lanes of
 * the last x tile idle than 64-wide ones. */
int mfd_op_set_tile(mfd_op *op, int prec, int ty, int kz, int xcd_order);
/* Average device time (ms) of one apply of the solver's stencil kernel
 * (MFD_FP64: A; MFD_MIXED: the complex-FP32 inner operator) */
int mfd_op_bench_apply(mfd_op *op, int64_t nrhs, int prec, int iters, double *ms_per_apply);
/* Average device time (ms) of each kernel of one SQMR iteration, in launch
 * order: ms[0] stencil with the qᵀPq epilogue, ms[1] r update with its
 * reductions, ms[2] d/z/q update.  Uses (and scribbles over) the solver
 * workspace.  No reference counterpart. */
int mfd_op_bench_iteration(mfd_op *op, int64_t nrhs, int prec, int iters, double ms[3]);

/* Solve A e = b (the user's `A \ b`); e holds the initial guess on entry.
 * The result meets tol on the FP64 residual b − A e (computed with the FP64
 * operator after each cycle); SQMR FP64 cycles end as soon as a probe of that
 * residual is at or below 0.9·tol.  A right-hand side b = 0 gets e = 0 (the
 * exact solution) whatever the initial guess.  A right-hand side whose
 * residual is NaN stops iterating and the solve returns MFD_ERR_NOCONV with
 * stats.relres NaN.  A SQMR or BiCGSTAB solve that ends without converging
 * (MFD_ERR_NOCONV) leaves in each column of e the iterate with that column's
 * smallest such residual (converged columns keep their converged iterate);
 * stats.relres reports the largest of those residuals.  BiCGSTAB cycles also
 * probe that residual every 2,048 steps: a probe that improves a column's
 * best keeps its iterate as the best, the cycle ends once every column meets
 * 0.9·tol on a probe or the probed residual rises 10× above the cycle's best
 * probe, and a column that ends a cycle above twice its best starts the next
 * cycle from its best iterate. */
void mfd_solve_opts_default(mfd_solve_opts *o);
int mfd_solve(mfd_op *op, const mfd_c64 *b, mfd_c64 *e, int64_t nrhs, const mfd_solve_opts *o,
              mfd_solve_stats *st);
int mfd_solve_dev(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *o,
                  mfd_solve_stats *st);
/* Adjoint solve Aᵀ λ = g using Aᵀ = S_E A S_E⁻¹ (complex symmetry). */
/* mfd_solve_dev without allocating: no reference counterpart; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, method choice,
 * convergence rules and results as mfd_solve_dev with use_graph = 0 (graphs
 * are never used here), but the solver workspace lives in a fixed 64 MiB
 * array per GPU that the library holds until the process ends (set up by
 * mfd_ctx_create): about 150 bytes per unknown and right-hand side plus
 * 1.5 MB per right-hand side, so a 32³ grid with 4 right-hand sides fits.
 * A larger request fails with MFD_ERR_ARG before any GPU work.  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU race.  The
 * operator must have been created (mfd_op_create allocates) beforehand. */
int mfd_solve_dev_noalloc(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *o,
                          mfd_solve_stats *st);
/* Solve `count` independent systems A_p e_p = b_p, one RHS each (every
 * problem keeps its own grid, ε and ω; all operators share one context).
 * Each chunk of iterations of all unconverged problems is one hipGraph launch
 * whose per-problem branches run concurrently.  b[p], e[p]: host arrays of
 * 3·M_p values (e holds the initial guess); st: `count` entries or null.
 * b = 0, NaN residuals and unconverged problems are handled per problem as
 * mfd_solve handles its columns.
 * SQMR only.  No reference counterpart (the reference solves one `A \ b`). */
int mfd_solve_batch(int count, mfd_op *const *ops, const mfd_c64 *const *b, mfd_c64 *const *e,
                    const mfd_solve_opts *o, mfd_solve_stats *st);
/* The same with device pointers on the context's GPU. */
int mfd_solve_batch_dev(int count, mfd_op *const *ops, const void *const *b, void *const *e,
                        const mfd_solve_opts *o, mfd_solve_stats *st);
/* h_from_e (model.jl:276-279): h = (i/ω) Pμ⁻¹ (Cₑ e + jₘ) with the operator's
 * grid, μ and ω; e, jm, h host DOF vectors of 3M values (jm may be null).
 * Needs an EE-formulation operator (as e_from_h does). */
int mfd_h_from_e(mfd_op *op, const mfd_c64 *e, const mfd_c64 *jm, mfd_c64 *h);
/* e_from_h (model.jl:281-284): e = (−i/ω) Pε⁻¹ (Cₘ h − jₑ); diagonal ε only
 * (MFD_ERR_UNSUPPORTED for an operator with off-diagonal ε); je may be null. */
int mfd_e_from_h(mfd_op *op, const mfd_c64 *h, const mfd_c64 *je, mfd_c64 *e);
int mfd_solve_adjoint(mfd_op *op, const mfd_c64 *g, mfd_c64 *lam, const mfd_solve_opts *o,
                      mfd_solve_stats *st);
/* ∂f/∂ε per E DOF for real f with Aᵀλ = (∂f/∂e)ᵀ: grad = 2ω² Re(λ ∘ e)
 * (derivative through the diagonal of −ω²Pε; for an HH operator, ∂f/∂μ per
 * H DOF through −ω²Pμ). */
int mfd_adjoint_grad(mfd_op *op, const mfd_c64 *e, const mfd_c64 *lam, double *grad);

#ifdef __cplusplus
}
#endif
#endif
