/*
 * CPU restatement of the MaxwellFDFD frequency-domain assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is imported by tests/, by the smoke check and by the
 * cpu_baseline leg of bench.py, and only as the checker.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Functions whose arithmetic lives in the absent MaxwellBase dependency are
 * restated from the named textbook method; their pinning is listed in the
 * header comment of oracle.c.
 *
 * Conventions (reference src/model/model.jl:32-81, Appendix A of the
 * reference's own conventions):
 *   - axis w in {0,1,2}; grid type 0 = PRIM, 1 = DUAL; field type 0 = EE, 1 = HH
 *   - DOF index (0-based) = c + 3*p  (order_cmpfirst) or p + M*c otherwise,
 *     with p = i + Nx*(j + Ny*k), M = Nx*Ny*Nz
 *   - parameter arrays are (Nx,Ny,Nz,3,3) column-major: entry (v,w) of the
 *     tensor at cell p sits at p + M*(v + 3*w)
 *   - CSC outputs are 1-based (Julia SparseMatrixCSC{ComplexF64,Int64})
 */
#ifndef MFDFD_ORACLE_H
#define MFDFD_ORACLE_H
#include <stdint.h>
#include <complex.h>

typedef double _Complex oc64;

typedef struct {
    int64_t N[3];
    const double *lprim[3];   /* N[w]+1 primal plane locations, both bounds included */
    int32_t isbloch[3];
    int32_t boundft[3];       /* 0 = EE, 1 = HH */
    int64_t npml[2][3];       /* [negative/positive end][axis] */
    double omega_pml[2];      /* complex ωpml (re, im) */
    double kbloch[3];
    double pml_m;             /* polynomial grading degree (4) */
    double pml_lnR;           /* ln of target reflection (-16) */
    int32_t order_cmpfirst;
} og_desc;

typedef struct {
    int32_t kind;             /* 0 = ball, 1 = axis-aligned cuboid, 2 = cylinder along z (c, r, height s[2]) */
    int32_t pad;
    double c[3];              /* centre */
    double r;                 /* ball radius */
    double s[3];              /* cuboid side lengths */
} og_shape;

typedef struct {
    int64_t nnz;
    int64_t *colptr;          /* n+1, 1-based */
    int64_t *rowval;          /* nnz, 1-based */
    oc64 *nzval;              /* nnz */
} og_csc;

/* grid + PML tables */
int og_grid_tables(const og_desc *d, double *l_out /* [2][3][N] flattened per axis */,
                   double *dl_out, oc64 *sdl_out);
int og_stretched_dl(const og_desc *d, int gt, int w, oc64 *out /* N[w] */);

/* source discretisation (src/source/source.jl:208-306, pointsrc.jl:64-103) */
int og_distweights(double c, int gt, const double bounds[2], const double *l,
                   const double *dl, int64_t N, int isbloch, int64_t ind[2], double wt[2]);
int og_add_pointsrc(const og_desc *d, int ft, const double c[3], const double p[3],
                    double Idr_re, double Idr_im, oc64 *jarr /* (Nx,Ny,Nz,3) */);
int og_add_planesrc(const og_desc *d, int ft, int nn /* normal axis 0..2 */, double c,
                    const double p[3], double Jdn_re, double Jdn_im, oc64 *jarr /* (Nx,Ny,Nz,3) */);
int og_field_arr2vec(const og_desc *d, const oc64 *jarr, oc64 *vec);

/* sparse assembly; CSC arrays are malloc'ed by the oracle and freed by og_csc_free */
int og_curl(const og_desc *d, int which /* 0 = Ce, 1 = Cm */, og_csc *out);
int og_paramop(const og_desc *d, int which /* 0 = eps, 1 = mu */, const oc64 *parr, og_csc *out);
int og_Mcs(const og_desc *d, int which /* 0 = Mc_e, 1 = Mc_m */, og_csc *out);
int og_create_A(const og_desc *d, double om_re, double om_im, const oc64 *eps_arr,
                const oc64 *mu_arr, og_csc *out);
int og_create_b(const og_desc *d, double om_re, double om_im, const oc64 *mu_arr,
                const oc64 *je, const oc64 *jm, oc64 *b);
/* HH formulation (model.jl:238-240, 266-268); Pε must be diagonal */
int og_create_A_hh(const og_desc *d, double om_re, double om_im, const oc64 *eps_arr,
                   const oc64 *mu_arr, og_csc *out);
int og_create_b_hh(const og_desc *d, double om_re, double om_im, const oc64 *eps_arr,
                   const oc64 *je, const oc64 *jm, oc64 *b);
void og_csc_free(og_csc *m);
int og_spmv(int64_t n, const og_csc *A, const oc64 *x, oc64 *y);

/* material assignment and Kottke subpixel smoothing (src/model/full.jl:16-70) */
int og_matparams(const og_desc *d, int64_t nobj, const og_shape *shapes,
                 const oc64 *eps_t /* nobj*9 col-major */, const oc64 *mu_t,
                 oc64 *eps_arr, oc64 *mu_arr);
double og_volfrac(const double vmin[3], const double vmax[3], const double n[3],
                  const double r0[3]);
void og_kottke(const oc64 *efg, const oc64 *ebg, double r, const double n[3], oc64 *out);

/* FP64 CSC Krylov solve with the same scaling as the device solver */
int og_solve_sqmr(const og_desc *d, const og_csc *A, const oc64 *b, oc64 *x,
                  double tol, int64_t maxit, int nthreads, int64_t *iters, double *relres);

#endif
