// C-ABI entry points of libmfdfd (declared in include/mfdfd.h).
#include "assembly.hpp"
#include "matparams.hpp"
#include "op.hpp"
#include <cmath>
#include <new>

int op_create_impl(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr,
                   const mfd_c64 *mu_arr, mfd_op **out);
int op_destroy_impl(mfd_op *op);
int op_apply_impl(mfd_op *op, const mfd_c64 *x, mfd_c64 *y, int64_t nrhs);
int op_apply_dev_impl(mfd_op *op, const void *x, void *y, int64_t nrhs, int prec);
int op_set_graddiv_impl(mfd_op *op, int enable);
int op_set_components_impl(mfd_op *op, int c0, int nc);
int op_set_tile_impl(mfd_op *op, int prec, int ty, int kz, int xcd);
int op_autotune_impl(mfd_op *op, int64_t nrhs, int iters, int *best64, int *best32, double *ms_out);
int op_bench_apply_impl(mfd_op *op, int64_t nrhs, int prec, int iters, double *ms);
int op_bench_iteration_impl(mfd_op *op, int64_t nrhs, int prec, int iters, double *ms);
void solve_opts_default(mfd_solve_opts *o);
int solve_impl(mfd_op *op, const mfd_c64 *b, mfd_c64 *e, int64_t nrhs, const mfd_solve_opts *o, mfd_solve_stats *st);
int solve_dev_impl(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *o, mfd_solve_stats *st);
int solve_batch_impl(int count, mfd_op *const *ops, const mfd_c64 *const *b, mfd_c64 *const *e,
                     const mfd_solve_opts *o, mfd_solve_stats *st);
int solve_batch_dev_impl(int count, mfd_op *const *ops, const void *const *b, void *const *e,
                         const mfd_solve_opts *o, mfd_solve_stats *st);
int field_convert_impl(mfd_op *op, int which, const mfd_c64 *in, const mfd_c64 *src, mfd_c64 *out);
int solve_adjoint_impl(mfd_op *op, const mfd_c64 *g, mfd_c64 *lam, const mfd_solve_opts *o, mfd_solve_stats *st);
int adjoint_grad_impl(mfd_op *op, const mfd_c64 *e, const mfd_c64 *lam, double *grad);
int create_b_impl(mfd_ctx *ctx, const mfd::DevGrid &G, int ft, mfd_c64 omega, const mfd_c64 *par, const mfd_c64 *je,
                  const mfd_c64 *jm, mfd_c64 *b);
int op_create_csc_impl(mfd_ctx *ctx, const mfd_csc *A, mfd_op **out);
int solve_dev_noalloc_impl(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *o,
                           mfd_solve_stats *st);
namespace mfd {
int noalloc_scratch_init(int device);
}

using namespace mfd;

#define CHECK_CTX(ctx)                                        \
    do {                                                      \
        if (!(ctx)) { set_error("null context"); return MFD_ERR_ARG; } \
        if (hipSetDevice((ctx)->device) != hipSuccess) { set_error("hipSetDevice failed"); return MFD_ERR_HIP; } \
    } while (0)

extern "C" {

const char *mfd_version(void) { return "libmfdfd 0.1.0 (gfx950)"; }
const char *mfd_last_error(void) { return mfd::last_error(); }

int mfd_ctx_create(int device, mfd_ctx **out) {
    if (!out) { set_error("null output"); return MFD_ERR_ARG; }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { set_error("no HIP device visible"); return MFD_ERR_HIP; }
    if (device < 0 || device >= n) { set_error("device index out of range"); return MFD_ERR_ARG; }
    MFD_HIP(hipSetDevice(device));
    // device address of this GPU's scratch array (mfd_solve_dev_noalloc), taken
    // once per device with that device current
    MFD_TRY(noalloc_scratch_init(device));
    mfd_ctx *c = new (std::nothrow) mfd_ctx();
    if (!c) return MFD_ERR_ALLOC;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed");
        return MFD_ERR_HIP;
    }
    *out = c;
    return MFD_OK;
}

int mfd_ctx_destroy(mfd_ctx *ctx) {
    if (!ctx) return MFD_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return MFD_OK;
}

int mfd_stretched_dl(mfd_ctx *ctx, const mfd_grid *g, mfd_c64 *out) {
    CHECK_CTX(ctx);
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    MFD_HIP(hipMemcpy(out, G.sdl64, sizeof(c64) * G.total, hipMemcpyDeviceToHost));
    return MFD_OK;
}

// source.jl:208-306 (indices returned 1-based as in Julia)
int mfd_distweights(double c, int gt, const double bounds[2], const double *l, const double *dl, int64_t N,
                    int isbloch, int64_t ind[2], double wt[2]) {
    if (!(bounds[0] <= c && c <= bounds[1])) { set_error("distweights: c must be within bounds"); return MFD_ERR_ARG; }
    double L = bounds[1] - bounds[0];
    if (!(N > 1 || isbloch)) { set_error("distweights: length(l) must be > 1 for a symmetry boundary"); return MFD_ERR_ARG; }
    bool zeroing = (gt == MFD_PRIM) && !isbloch;   // the boundary zeroes primal sources
    int64_t lo = zeroing ? 1 : 0, hi = N - 1;      // range unaffected by the boundary: l[lo] <= c < l[hi]
    bool free_range = (l[lo] <= c) && (c < l[hi]);
    int64_t i1, i2;
    double dc, span;
    if (free_range) {
        // l is sorted: last index with l[i] <= c (binary search)
        int64_t a = lo, b = hi;   // l[a] <= c < l[b]
        while (b - a > 1) {
            int64_t m = (a + b) / 2;
            if (l[m] - c <= 0) a = m; else b = m;
        }
        i1 = a;
        i2 = a + 1;
        dc = c - l[i1];
        span = l[i2] - l[i1];
    } else {
        double bnd;
        if (c < l[lo]) { i1 = lo; i2 = N - 1; bnd = bounds[0]; }
        else { i1 = hi; i2 = 0; bnd = bounds[1]; }
        dc = fabs(l[i1] - c);
        if (gt == MFD_PRIM) span = fabs(l[i1] - bnd);
        else span = isbloch ? L - fabs(l[i1] - l[i2]) : 2.0 * fabs(l[i1] - bnd);
    }
    double r = dc / span;
    double w1 = 1.0 / dl[i1];
    if (free_range || isbloch || zeroing) w1 *= 1.0 - r;
    double w2;
    if (c == l[i1] || (!free_range && !isbloch)) { i2 = i1; w2 = 0.0; }
    else {
        w2 = 1.0 / dl[i2];
        if (free_range || isbloch) w2 *= r;
    }
    ind[0] = i1 + 1;
    ind[1] = i2 + 1;
    wt[0] = w1;
    wt[1] = w2;
    return MFD_OK;
}

// pointsrc.jl:64-103 for K = Kf = 3: J_w = I∆r p_w Π_u wt_u over the 2x2x2 corners
int mfd_add_pointsrc(const mfd_grid *g, int ft, const double c[3], const double p[3], mfd_c64 Idr, mfd_c64 *jarr) {
    HGrid h;
    MFD_TRY(build_hgrid(g, &h));
    if (ft != MFD_EE && ft != MFD_HH) { set_error("ft must be EE or HH"); return MFD_ERR_ARG; }
    double nrm = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (!(nrm > 0)) { set_error("point source polarisation must be nonzero"); return MFD_ERR_ARG; }
    double pn[3] = {p[0] / nrm, p[1] / nrm, p[2] / nrm};
    for (int w = 0; w < 3; ++w) {
        int64_t ind[3][2];
        double wt[3][2];
        for (int u = 0; u < 3; ++u) {
            int gt = gtype(h.bft, ft, w, u);
            double b2[2] = {h.bnd[0][u], h.bnd[1][u]};
            MFD_TRY(mfd_distweights(c[u], gt, b2, h.l[gt][u].data(), h.dl[gt][u].data(), h.N[u], h.bloch[u], ind[u], wt[u]));
        }
        double ir = Idr.re * pn[w], ii = Idr.im * pn[w];
        for (int kz = 0; kz < 2; ++kz)
            for (int ky = 0; ky < 2; ++ky)
                for (int kx = 0; kx < 2; ++kx) {
                    double wv = ((1.0 * wt[0][kx]) * wt[1][ky]) * wt[2][kz];
                    int64_t q = (ind[0][kx] - 1) + h.N[0] * ((ind[1][ky] - 1) + h.N[1] * (ind[2][kz] - 1));
                    mfd_c64 &J = jarr[q + h.M * w];
                    J.re += ir * wv;
                    J.im += ii * wv;
                }
    }
    return MFD_OK;
}

// planesrc.jl:41-79 for K = Kf = 3: along the normal axis nn, distweights of
// the intercept c pick two planes of J_w's grid type; J_w = J∆n p_w wt is added
// over each whole plane
int mfd_add_planesrc(const mfd_grid *g, int ft, int nn, double c, const double p[3], mfd_c64 Jdn, mfd_c64 *jarr) {
    HGrid h;
    MFD_TRY(build_hgrid(g, &h));
    if (ft != MFD_EE && ft != MFD_HH) { set_error("ft must be EE or HH"); return MFD_ERR_ARG; }
    if (nn < 0 || nn > 2) { set_error("plane normal must be a Cartesian axis 0..2"); return MFD_ERR_ARG; }
    double nrm = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (!(nrm > 0)) { set_error("plane source polarisation must be nonzero"); return MFD_ERR_ARG; }
    if (!jarr) { set_error("null current array"); return MFD_ERR_ARG; }
    double pn[3] = {p[0] / nrm, p[1] / nrm, p[2] / nrm};
    for (int w = 0; w < 3; ++w) {
        int gt = gtype(h.bft, ft, w, nn);
        double b2[2] = {h.bnd[0][nn], h.bnd[1][nn]};
        int64_t ind[2];
        double wt[2];
        MFD_TRY(mfd_distweights(c, gt, b2, h.l[gt][nn].data(), h.dl[gt][nn].data(), h.N[nn], h.bloch[nn], ind, wt));
        double jr = Jdn.re * pn[w], ji = Jdn.im * pn[w];
        for (int k = 0; k < 2; ++k) {
            int64_t lo[3] = {0, 0, 0}, hi[3] = {h.N[0] - 1, h.N[1] - 1, h.N[2] - 1};
            lo[nn] = hi[nn] = ind[k] - 1;
            double vr = jr * wt[k], vi = ji * wt[k];
            for (int64_t z = lo[2]; z <= hi[2]; ++z)
                for (int64_t y = lo[1]; y <= hi[1]; ++y)
                    for (int64_t x = lo[0]; x <= hi[0]; ++x) {
                        mfd_c64 &J = jarr[x + h.N[0] * (y + h.N[1] * z) + h.M * w];
                        J.re += vr;
                        J.im += vi;
                    }
        }
    }
    return MFD_OK;
}

int mfd_curl_nnz(mfd_ctx *ctx, const mfd_grid *g, int which, int64_t *nnz) {
    CHECK_CTX(ctx);
    if (which != 0 && which != 1) { set_error("which must be 0 (Cₑ) or 1 (Cₘ)"); return MFD_ERR_ARG; }
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    return curl_csc(ctx, G, which, nnz, nullptr, nullptr, nullptr);
}

int mfd_curl_csc(mfd_ctx *ctx, const mfd_grid *g, int which, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval) {
    CHECK_CTX(ctx);
    if (which != 0 && which != 1) { set_error("which must be 0 (Cₑ) or 1 (Cₘ)"); return MFD_ERR_ARG; }
    if (!colptr || !rowval || !nzval) { set_error("null output buffer"); return MFD_ERR_ARG; }
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    int64_t nnz;
    return curl_csc(ctx, G, which, &nnz, colptr, rowval, nzval);
}

int mfd_Mcs_nnz(mfd_ctx *ctx, const mfd_grid *g, int which, int64_t *nnz) {
    CHECK_CTX(ctx);
    if (which != 0 && which != 1) { set_error("which must be 0 (Mcₑ) or 1 (Mcₘ)"); return MFD_ERR_ARG; }
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    return mean_csc(ctx, G, which, nnz, nullptr, nullptr, nullptr);
}

int mfd_Mcs_csc(mfd_ctx *ctx, const mfd_grid *g, int which, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval) {
    CHECK_CTX(ctx);
    if (which != 0 && which != 1) { set_error("which must be 0 (Mcₑ) or 1 (Mcₘ)"); return MFD_ERR_ARG; }
    if (!colptr || !rowval || !nzval) { set_error("null output buffer"); return MFD_ERR_ARG; }
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    int64_t nnz;
    return mean_csc(ctx, G, which, &nnz, colptr, rowval, nzval);
}

int mfd_matparams(mfd_ctx *ctx, const mfd_grid *g, int64_t nobj, const mfd_shape *shapes, const mfd_c64 *eps_t,
                  const mfd_c64 *mu_t, mfd_c64 *eps_arr, mfd_c64 *mu_arr) {
    CHECK_CTX(ctx);
    return matparams(ctx, g, nobj, shapes, eps_t, mu_t, eps_arr, mu_arr);
}

int mfd_paramop_nnz(mfd_ctx *ctx, const mfd_grid *g, int which, const mfd_c64 *parr, int64_t *nnz) {
    CHECK_CTX(ctx);
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    return param_csc(ctx, G, which, parr, nnz, nullptr, nullptr, nullptr);
}

int mfd_paramop_csc(mfd_ctx *ctx, const mfd_grid *g, int which, const mfd_c64 *parr, int64_t *colptr,
                    int64_t *rowval, mfd_c64 *nzval) {
    CHECK_CTX(ctx);
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    int64_t nnz;
    return param_csc(ctx, G, which, parr, &nnz, colptr, rowval, nzval);
}

static int ft_check(int ft) {
    if (ft == MFD_EE || ft == MFD_HH) return MFD_OK;
    set_error("ft is unsupported");   // model.jl:241-242, 269-270
    return MFD_ERR_UNSUPPORTED;
}

int mfd_A_nnz(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr, const mfd_c64 *mu_arr,
              int64_t *nnz) {
    CHECK_CTX(ctx);
    MFD_TRY(ft_check(ft));
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    return A_csc(ctx, G, ft, omega, eps_arr, mu_arr, nnz, nullptr, nullptr, nullptr);
}

int mfd_A_csc(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr, const mfd_c64 *mu_arr,
              int64_t *colptr, int64_t *rowval, mfd_c64 *nzval) {
    CHECK_CTX(ctx);
    MFD_TRY(ft_check(ft));
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    int64_t nnz;
    return A_csc(ctx, G, ft, omega, eps_arr, mu_arr, &nnz, colptr, rowval, nzval);
}

int mfd_create_b(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr,
                 const mfd_c64 *mu_arr, const mfd_c64 *je, const mfd_c64 *jm, mfd_c64 *b) {
    CHECK_CTX(ctx);
    MFD_TRY(ft_check(ft));
    DevGrid G;
    MFD_TRY(G.create(g, ctx->stream));
    return create_b_impl(ctx, G, ft, omega, ft == MFD_EE ? mu_arr : eps_arr, je, jm, b);
}

int mfd_op_create(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr,
                  const mfd_c64 *mu_arr, mfd_op **out) {
    CHECK_CTX(ctx);
    if (!out || !eps_arr || !mu_arr) { set_error("null argument"); return MFD_ERR_ARG; }
    return op_create_impl(ctx, g, ft, omega, eps_arr, mu_arr, out);
}
int mfd_op_destroy(mfd_op *op) { return op_destroy_impl(op); }
int mfd_op_create_csc(mfd_ctx *ctx, const mfd_csc *A, mfd_op **out) {
    CHECK_CTX(ctx);
    if (!out || !A) { set_error("null argument"); return MFD_ERR_ARG; }
    return op_create_csc_impl(ctx, A, out);
}
int mfd_A_from_csc_nnz(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
                       const mfd_csc *Cm, int64_t *nnz) {
    CHECK_CTX(ctx);
    if (!nnz) { set_error("null argument"); return MFD_ERR_ARG; }
    return A_from_csc(ctx, ft, omega, Pe, Pm, Ce, Cm, nnz, nullptr, nullptr, nullptr);
}
int mfd_A_from_csc(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
                   const mfd_csc *Cm, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval) {
    CHECK_CTX(ctx);
    if (!colptr || !rowval || !nzval) { set_error("null output"); return MFD_ERR_ARG; }
    int64_t nnz = 0;
    return A_from_csc(ctx, ft, omega, Pe, Pm, Ce, Cm, &nnz, colptr, rowval, nzval);
}
int mfd_b_from_csc(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
                   const mfd_csc *Cm, const mfd_c64 *je, const mfd_c64 *jm, mfd_c64 *b) {
    CHECK_CTX(ctx);
    return b_from_csc(ctx, ft, omega, Pe, Pm, Ce, Cm, je, jm, b);
}
int mfd_h_from_e_csc(mfd_ctx *ctx, const mfd_c64 *e, mfd_c64 omega, const mfd_csc *Pm, const mfd_csc *Ce,
                     const mfd_c64 *jm, mfd_c64 *h) {
    CHECK_CTX(ctx);
    return field_from_csc(ctx, 0, omega, Pm, Ce, e, jm, h);
}
int mfd_e_from_h_csc(mfd_ctx *ctx, const mfd_c64 *h, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Cm,
                     const mfd_c64 *je, mfd_c64 *e) {
    CHECK_CTX(ctx);
    return field_from_csc(ctx, 1, omega, Pe, Cm, h, je, e);
}
int mfd_op_apply(mfd_op *op, const mfd_c64 *x, mfd_c64 *y, int64_t nrhs) {
    if (!op) { set_error("null operator"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return op_apply_impl(op, x, y, nrhs);
}
int mfd_op_apply_dev(mfd_op *op, const void *x, void *y, int64_t nrhs, int prec) {
    if (!op) { set_error("null operator"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return op_apply_dev_impl(op, x, y, nrhs, prec);
}
int mfd_op_set_graddiv(mfd_op *op, int enable) {
    if (!op) { set_error("null operator"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return op_set_graddiv_impl(op, enable);
}
int mfd_op_set_components(mfd_op *op, int c0, int nc) {
    if (!op) { set_error("null operator"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return op_set_components_impl(op, c0, nc);
}
int mfd_op_set_tile(mfd_op *op, int prec, int ty, int kz, int xcd_order) {
    if (!op) { set_error("null operator"); return MFD_ERR_ARG; }
    return op_set_tile_impl(op, prec, ty, kz, xcd_order);
}
int mfd_op_autotune(mfd_op *op, int64_t nrhs, int iters, int tile64[3], int tile32[3], double ms[2]) {
    if (!op || nrhs < 1 || iters < 1) { set_error("bad argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return op_autotune_impl(op, nrhs, iters, tile64, tile32, ms);
}
int mfd_op_bench_apply(mfd_op *op, int64_t nrhs, int prec, int iters, double *ms) {
    if (!op || !ms || iters < 1) { set_error("bad argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return op_bench_apply_impl(op, nrhs, prec, iters, ms);
}

int mfd_op_bench_iteration(mfd_op *op, int64_t nrhs, int prec, int iters, double ms[3]) {
    if (!op || !ms) { set_error("bad argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return op_bench_iteration_impl(op, nrhs, prec, iters, ms);
}
void mfd_solve_opts_default(mfd_solve_opts *o) { solve_opts_default(o); }
int mfd_solve(mfd_op *op, const mfd_c64 *b, mfd_c64 *e, int64_t nrhs, const mfd_solve_opts *o, mfd_solve_stats *st) {
    if (!op || !b || !e) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return solve_impl(op, b, e, nrhs, o, st);
}
int mfd_solve_dev(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *o, mfd_solve_stats *st) {
    if (!op || !b || !e) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return solve_dev_impl(op, b, e, nrhs, o, st);
}
int mfd_solve_dev_noalloc(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *o,
                          mfd_solve_stats *st) {
    if (!op || !b || !e) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return solve_dev_noalloc_impl(op, b, e, nrhs, o, st);
}
int mfd_solve_batch(int count, mfd_op *const *ops, const mfd_c64 *const *b, mfd_c64 *const *e,
                    const mfd_solve_opts *o, mfd_solve_stats *st) {
    if (count < 1 || !ops || !ops[0]) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(ops[0]->ctx);
    return solve_batch_impl(count, ops, b, e, o, st);
}
int mfd_solve_batch_dev(int count, mfd_op *const *ops, const void *const *b, void *const *e,
                        const mfd_solve_opts *o, mfd_solve_stats *st) {
    if (count < 1 || !ops || !ops[0]) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(ops[0]->ctx);
    return solve_batch_dev_impl(count, ops, b, e, o, st);
}
int mfd_solve_adjoint(mfd_op *op, const mfd_c64 *g, mfd_c64 *lam, const mfd_solve_opts *o, mfd_solve_stats *st) {
    if (!op || !g || !lam) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return solve_adjoint_impl(op, g, lam, o, st);
}
int mfd_adjoint_grad(mfd_op *op, const mfd_c64 *e, const mfd_c64 *lam, double *grad) {
    if (!op || !e || !lam || !grad) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return adjoint_grad_impl(op, e, lam, grad);
}

}  // extern "C"

int mfd_h_from_e(mfd_op *op, const mfd_c64 *e, const mfd_c64 *jm, mfd_c64 *h) {
    if (!op || !e || !h) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return field_convert_impl(op, 0, e, jm, h);
}
int mfd_e_from_h(mfd_op *op, const mfd_c64 *h, const mfd_c64 *je, mfd_c64 *e) {
    if (!op || !h || !e) { set_error("null argument"); return MFD_ERR_ARG; }
    CHECK_CTX(op->ctx);
    return field_convert_impl(op, 1, h, je, e);
}
