/*
 * CPU restatement of the MaxwellFDFD assembly path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, the smoke check and bench.py's cpu_baseline leg load this file,
 * and only as the checker.  The product library never links or calls it.
 *
 * What each part follows, and what pins it:
 *   grid tables        Grid conventions used by test/source.jl:6-27 and
 *                      pointsrc.jl:66-68 (primal ghost at the positive end, dual
 *                      ghost at the negative end, ∆l[PRIM] = diff(dual incl.
 *                      ghost), ∆l[DUAL] = diff(primal incl. ghost)).
 *                      Pinned by the test/source.jl:17-27 grid and closed forms.
 *   stretched ∆l       model.jl:122-139 -> MaxwellBase create_stretched_∆l
 *                      (absent); restated as SC-PML, Shin & Fan, J. Comput.
 *                      Phys. 231 (2012) 3406: s = 1 + σ/(iω), σ = σmax d^m,
 *                      σmax = -(m+1) ln R / (2 Lpml).  Pinned by the closed form
 *                      and the reflection invariant (tests); parity vs MaxwellBase
 *                      itself unpinned.
 *   distweights        source.jl:208-306, line by line.  Pinned by the golden
 *                      vectors of test/source.jl:17-133.
 *   point source       pointsrc.jl:64-103.  Pinned by test/source.jl:184-242.
 *   plane source       planesrc.jl:41-79.  Pinned by test/source.jl:123-170
 *                      (peak J = K/∆z on the 1 and 0.5 grids, zero Jy/Jz,
 *                      resolution-independent total current).
 *   curl               model.jl:160-175 -> MaxwellBase create_curl (absent);
 *                      Yee forward/backward differences.  Pinned by div∘curl=0,
 *                      Cm = S_E^-1 Ce^T S_H, the periodic plane-wave eigenvalue
 *                      and brute-force differences on tiny grids (tests).
 *   param operator     model.jl:141-158 -> MaxwellBase create_paramop (absent).
 *                      Pinned by the diagonal case and S-symmetry (tests).
 *   corner mean        model.jl:287-306 -> MaxwellBase create_mean (absent).
 *                      Pinned by the uniform-grid weights (1/2, 1/2), the
 *                      partition of unity on nonuniform real grids and the
 *                      Bloch-wave eigenrelation (tests); parity vs MaxwellBase
 *                      itself unpinned.
 *   create_A/create_b  model.jl:225-274 literally (SpGEMM, sparse minus that
 *                      drops numerical zeros, as Julia SparseArrays does).
 *   matparams          full.jl:16-70 -> assign_param!/smooth_param! (absent);
 *                      Kottke, Farjadpour & Johnson, PRE 77, 036611 (2008).
 *                      Pinned by brute-force fill fractions, the planar closed
 *                      form and tensor invariants (tests).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { PRIM = 0, DUAL = 1, EE = 0, HH = 1 };

typedef struct {
    int64_t N[3], M;
    double *l[2][3], *lg[2][3], *tlg[2][3], *dl[2][3];
    oc64 *sdl[2][3];
    double bnd[2][3], L[3];
    int bloch[3], bft[3], cmpfirst;
    oc64 ph[3];
    int64_t npml[2][3];
    double om[2], m, lnR;
} G;

static void g_free(G *g) {
    for (int t = 0; t < 2; ++t)
        for (int w = 0; w < 3; ++w) {
            free(g->l[t][w]); free(g->lg[t][w]); free(g->tlg[t][w]);
            free(g->dl[t][w]); free(g->sdl[t][w]);
        }
}

static inline int alt(int gt) { return 1 - gt; }

/* Stretch factor s at location u on axis w (SC-PML, engineer's exp(+iωt)). */
static oc64 stretch(const G *g, int w, double u) {
    const double *lp = g->lg[PRIM][w];
    int64_t N = g->N[w];
    double sigma = 0.0;
    if (g->npml[0][w] > 0) {
        double lpml = lp[g->npml[0][w]], Lp = lpml - lp[0];
        if (u < lpml) {
            double d = (lpml - u) / Lp;
            sigma = (-(g->m + 1.0) * g->lnR / (2.0 * Lp)) * pow(d, g->m);
        }
    }
    if (g->npml[1][w] > 0) {
        double lpml = lp[N - g->npml[1][w]], Lp = lp[N] - lpml;
        if (u > lpml) {
            double d = (u - lpml) / Lp;
            sigma = (-(g->m + 1.0) * g->lnR / (2.0 * Lp)) * pow(d, g->m);
        }
    }
    double a = g->om[0], b = g->om[1], den = a * a + b * b;
    /* 1 + σ/(iω) = 1 - σ b/|ω|^2 - i σ a/|ω|^2 */
    double re = 1.0 - sigma * b / den, im = -(sigma * a / den);
    if (sigma == 0.0) { re = 1.0; im = 0.0; }
    return re + im * I;
}

static int g_build(const og_desc *d, G *g) {
    memset(g, 0, sizeof(*g));
    g->M = 1;
    for (int w = 0; w < 3; ++w) {
        int64_t N = d->N[w];
        if (N < 1) return -1;
        g->N[w] = N; g->M *= N;
        g->bloch[w] = d->isbloch[w] != 0;
        g->bft[w] = d->boundft[w];
        g->npml[0][w] = d->npml[0][w]; g->npml[1][w] = d->npml[1][w];
        const double *lp = d->lprim[w];
        for (int t = 0; t < 2; ++t) {
            g->l[t][w] = malloc(sizeof(double) * N);
            g->lg[t][w] = malloc(sizeof(double) * (N + 1));
            g->tlg[t][w] = malloc(sizeof(double) * (N + 1));
            g->dl[t][w] = malloc(sizeof(double) * N);
            g->sdl[t][w] = malloc(sizeof(oc64) * N);
        }
        double L = lp[N] - lp[0];
        g->bnd[0][w] = lp[0]; g->bnd[1][w] = lp[N]; g->L[w] = L;
        for (int64_t i = 0; i <= N; ++i) g->lg[PRIM][w][i] = lp[i];
        for (int64_t i = 0; i < N; ++i) {
            g->l[PRIM][w][i] = lp[i];
            g->l[DUAL][w][i] = 0.5 * (lp[i] + lp[i + 1]);   /* movingavg */
            g->lg[DUAL][w][i + 1] = g->l[DUAL][w][i];
        }
        /* dual ghost at the negative end: periodic image (Bloch) or mirror image */
        g->lg[DUAL][w][0] = g->bloch[w] ? g->l[DUAL][w][N - 1] - L : 2.0 * lp[0] - g->l[DUAL][w][0];
        memcpy(g->tlg[PRIM][w], g->lg[PRIM][w], sizeof(double) * (N + 1));
        memcpy(g->tlg[DUAL][w], g->lg[DUAL][w], sizeof(double) * (N + 1));
        g->tlg[PRIM][w][N] = g->bloch[w] ? lp[0] : lp[N];
        g->tlg[DUAL][w][0] = g->bloch[w] ? g->l[DUAL][w][N - 1] : g->l[DUAL][w][0];
        for (int64_t i = 0; i < N; ++i) {
            g->dl[PRIM][w][i] = g->lg[DUAL][w][i + 1] - g->lg[DUAL][w][i];
            g->dl[DUAL][w][i] = g->lg[PRIM][w][i + 1] - g->lg[PRIM][w][i];
        }
        double kl = d->kbloch[w] * L;
        g->ph[w] = cos(kl) - sin(kl) * I;
    }
    g->cmpfirst = d->order_cmpfirst != 0;
    g->om[0] = d->omega_pml[0]; g->om[1] = d->omega_pml[1];
    g->m = d->pml_m; g->lnR = d->pml_lnR;
    int need = 0;
    for (int w = 0; w < 3; ++w) need |= (g->npml[0][w] > 0) | (g->npml[1][w] > 0);
    if (need && g->om[0] == 0.0 && g->om[1] == 0.0) { g_free(g); return -2; }
    for (int w = 0; w < 3; ++w) {
        for (int64_t i = 0; i < g->N[w]; ++i) {
            oc64 sp = need ? stretch(g, w, g->l[PRIM][w][i]) : 1.0;
            oc64 sd = need ? stretch(g, w, g->l[DUAL][w][i]) : 1.0;
            g->sdl[PRIM][w][i] = sp * g->dl[PRIM][w][i];
            g->sdl[DUAL][w][i] = sd * g->dl[DUAL][w][i];
        }
    }
    return 0;
}

static inline int64_t dof(const G *g, int c, int64_t p) {
    return g->cmpfirst ? (int64_t)c + 3 * p : p + g->M * c;
}
static inline int64_t cell(const G *g, int64_t i, int64_t j, int64_t k) {
    return i + g->N[0] * (j + g->N[1] * k);
}
static inline void unc(const G *g, int64_t p, int64_t q[3]) {
    q[0] = p % g->N[0]; q[1] = (p / g->N[0]) % g->N[1]; q[2] = p / (g->N[0] * g->N[1]);
}
/* grid type along axis a of component w of field type ft */
static inline int gtype(const G *g, int ft, int w, int a) {
    int g0 = (ft == EE) ? (g->bft[a] == EE ? PRIM : DUAL) : (g->bft[a] == HH ? PRIM : DUAL);
    return a == w ? alt(g0) : g0;
}
static inline int g0type(const G *g, int ft, int a) {
    return (ft == EE) ? (g->bft[a] == EE ? PRIM : DUAL) : (g->bft[a] == HH ? PRIM : DUAL);
}

int og_grid_tables(const og_desc *d, double *l_out, double *dl_out, oc64 *sdl_out) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    int64_t off = 0;
    for (int t = 0; t < 2; ++t)
        for (int w = 0; w < 3; ++w) {
            if (l_out) memcpy(l_out + off, g.l[t][w], sizeof(double) * g.N[w]);
            if (dl_out) memcpy(dl_out + off, g.dl[t][w], sizeof(double) * g.N[w]);
            if (sdl_out) memcpy(sdl_out + off, g.sdl[t][w], sizeof(oc64) * g.N[w]);
            off += g.N[w];
        }
    g_free(&g);
    return 0;
}

int og_stretched_dl(const og_desc *d, int gt, int w, oc64 *out) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    memcpy(out, g.sdl[gt][w], sizeof(oc64) * g.N[w]);
    g_free(&g);
    return 0;
}

/* ---------------------------------------------------------------- sources */

/* source.jl:208-306.  Indices returned 0-based. */
int og_distweights(double c, int gt, const double bounds[2], const double *l,
                   const double *dl, int64_t N, int isbloch, int64_t ind[2], double wt[2]) {
    if (!(bounds[0] <= c && c <= bounds[1])) return -1;          /* :214 */
    double L = bounds[1] - bounds[0];                              /* :216 */
    if (!(N > 1 || isbloch)) return -2;                            /* :230 */
    int zeroing_bc = (gt == PRIM) && !isbloch;                     /* :243 */
    int64_t indn = zeroing_bc ? 1 : 0;                             /* :244 (0-based) */
    int64_t indp = N - 1;                                          /* :245 */
    int bc_noeff = (l[indn] <= c) && (c < l[indp]);                /* :250 */
    int64_t i1, i2; double dc1, dlc;
    if (bc_noeff) {
        i1 = 0;
        for (int64_t i = 0; i < N; ++i) if (l[i] - c <= 0) i1 = i; /* :252 findlast */
        i2 = i1 + 1;
        dc1 = c - l[i1];
        dlc = l[i2] - l[i1];
    } else {
        double bnd;
        if (c < l[indn]) { i1 = indn; i2 = N - 1; bnd = bounds[0]; }  /* :257-260 */
        else { i1 = indp; i2 = 0; bnd = bounds[1]; }                   /* :261-264 */
        dc1 = fabs(l[i1] - c);
        if (gt == PRIM) dlc = fabs(l[i1] - bnd);
        else dlc = isbloch ? L - fabs(l[i1] - l[i2]) : 2.0 * fabs(l[i1] - bnd);
    }
    double r = dc1 / dlc;                                          /* :275 */
    double w1 = 1.0 / dl[i1];                                      /* :282 */
    if (bc_noeff || isbloch || zeroing_bc) w1 *= 1.0 - r;          /* :283 */
    double w2;
    if (c == l[i1] || (!bc_noeff && !isbloch)) { i2 = i1; w2 = 0.0; }  /* :289-294 */
    else { w2 = 1.0 / dl[i2]; if (bc_noeff || isbloch) w2 *= r; }      /* :298-299 */
    ind[0] = i1; ind[1] = i2; wt[0] = w1; wt[1] = w2;
    return 0;
}

/* pointsrc.jl:64-103 for K = Kf = 3 (isfield˔shp = false).  p must be normalised
 * by the caller as PointSrc's constructor does (pointsrc.jl:58). */
int og_add_pointsrc(const og_desc *d, int ft, const double c[3], const double p[3],
                    double Idr_re, double Idr_im, oc64 *jarr) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    for (int nw = 0; nw < 3; ++nw) {
        int64_t ind[3][2]; double wt[3][2];
        for (int nu = 0; nu < 3; ++nu) {
            int gt = gtype(&g, ft, nw, nu);
            double b2[2] = {g.bnd[0][nu], g.bnd[1][nu]};
            rc = og_distweights(c[nu], gt, b2, g.l[gt][nu], g.dl[gt][nu], g.N[nu],
                                g.bloch[nu], ind[nu], wt[nu]);
            if (rc) { g_free(&g); return rc; }
        }
        /* Iw∆w = I∆r * p[nw]; J = Iw∆w .* (((1*wx)*wy)*wz) (pointsrc.jl:87,94-95) */
        double iwr = Idr_re * p[nw], iwi = Idr_im * p[nw];
        for (int kz = 0; kz < 2; ++kz)
            for (int ky = 0; ky < 2; ++ky)
                for (int kx = 0; kx < 2; ++kx) {
                    double wv = ((1.0 * wt[0][kx]) * wt[1][ky]) * wt[2][kz];
                    int64_t q = cell(&g, ind[0][kx], ind[1][ky], ind[2][kz]);
                    jarr[q + g.M * nw] += iwr * wv + (iwi * wv) * I;
                }
    }
    g_free(&g);
    return 0;
}

/* planesrc.jl:41-79 for K = Kf = 3 (field and shape spaces coincide, so
 * J_w sits at gt_w(w, gt0)): along the normal axis nn, distweights of the
 * intercept c give two planes and weights; J_w = J∆n p_w wt is added over the
 * whole of each plane (planesrc.jl:71-77). */
int og_add_planesrc(const og_desc *d, int ft, int nn, double c, const double p[3],
                    double Jdn_re, double Jdn_im, oc64 *jarr) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    for (int nw = 0; nw < 3; ++nw) {
        int gt = gtype(&g, ft, nw, nn);
        double b2[2] = {g.bnd[0][nn], g.bnd[1][nn]};
        int64_t ind[2]; double wt[2];
        rc = og_distweights(c, gt, b2, g.l[gt][nn], g.dl[gt][nn], g.N[nn], g.bloch[nn], ind, wt);
        if (rc) { g_free(&g); return rc; }
        double jr = Jdn_re * p[nw], ji = Jdn_im * p[nw];
        for (int k = 0; k < 2; ++k) {
            oc64 v = jr * wt[k] + (ji * wt[k]) * I;
            int64_t lo[3] = {0, 0, 0}, hi[3] = {g.N[0] - 1, g.N[1] - 1, g.N[2] - 1};
            lo[nn] = hi[nn] = ind[k];   /* 0-based */
            for (int64_t z = lo[2]; z <= hi[2]; ++z)
                for (int64_t y = lo[1]; y <= hi[1]; ++y)
                    for (int64_t x = lo[0]; x <= hi[0]; ++x) jarr[cell(&g, x, y, z) + g.M * nw] += v;
        }
    }
    g_free(&g);
    return 0;
}

int og_field_arr2vec(const og_desc *d, const oc64 *jarr, oc64 *vec) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    for (int c = 0; c < 3; ++c)
        for (int64_t p = 0; p < g.M; ++p) vec[dof(&g, c, p)] = jarr[p + g.M * c];
    g_free(&g);
    return 0;
}

/* ------------------------------------------------------------- triplets */

typedef struct { int64_t r, c; oc64 v; } trip;
typedef struct { trip *t; int64_t n, cap; } tlist;
static void tpush(tlist *L, int64_t r, int64_t c, oc64 v) {
    if (L->n == L->cap) { L->cap = L->cap ? 2 * L->cap : 1024; L->t = realloc(L->t, sizeof(trip) * L->cap); }
    L->t[L->n].r = r; L->t[L->n].c = c; L->t[L->n].v = v; L->n++;
}

/* sparse(I,J,V): duplicates summed in input order, explicit zeros kept. */
static int csc_from_trip(int64_t n, tlist *L, og_csc *out) {
    int64_t *cnt = calloc(n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < L->n; ++e) cnt[L->t[e].c + 1]++;
    for (int64_t j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
    trip *s = malloc(sizeof(trip) * (L->n ? L->n : 1));
    int64_t *pos = malloc(sizeof(int64_t) * (n + 1));
    memcpy(pos, cnt, sizeof(int64_t) * (n + 1));
    for (int64_t e = 0; e < L->n; ++e) s[pos[L->t[e].c]++] = L->t[e];   /* stable by column */
    int64_t *colptr = malloc(sizeof(int64_t) * (n + 1));
    int64_t *rowval = malloc(sizeof(int64_t) * (L->n ? L->n : 1));
    oc64 *nzval = malloc(sizeof(oc64) * (L->n ? L->n : 1));
    int64_t nz = 0;
    colptr[0] = 1;
    for (int64_t j = 0; j < n; ++j) {
        int64_t a = cnt[j], b = cnt[j + 1];
        for (int64_t x = a + 1; x < b; ++x) {           /* stable insertion sort by row */
            trip t = s[x]; int64_t y = x - 1;
            while (y >= a && s[y].r > t.r) { s[y + 1] = s[y]; --y; }
            s[y + 1] = t;
        }
        for (int64_t x = a; x < b;) {
            oc64 v = s[x].v; int64_t r = s[x].r; ++x;
            while (x < b && s[x].r == r) { v += s[x].v; ++x; }
            rowval[nz] = r + 1; nzval[nz] = v; ++nz;
        }
        colptr[j + 1] = nz + 1;
    }
    out->nnz = nz; out->colptr = colptr; out->rowval = rowval; out->nzval = nzval;
    free(cnt); free(s); free(pos);
    return 0;
}

void og_csc_free(og_csc *m) {
    free(m->colptr); free(m->rowval); free(m->nzval);
    m->colptr = NULL; m->rowval = NULL; m->nzval = NULL; m->nnz = 0;
}

/* ----------------------------------------------------------------- curl */

/* Ce (which = 0): H = curl E, forward differences along axes with boundft EE,
 * divided by the stretched ∆l at H-field planes (model.jl:168,171).
 * Cm (which = 1): E = curl H, backward differences for EE, divided by the
 * stretched ∆l at E-field planes (model.jl:169,172).
 * Ghost rules: Bloch wraps with e^{-ikL} (forward) or 1/e^{-ikL} (backward);
 * a non-Bloch boundary takes the ghost value as zero. */
static void curl_trip(const G *g, int which, tlist *L) {
    int ftin = which == 0 ? EE : HH, ftout = which == 0 ? HH : EE;
    (void)ftin;
    for (int64_t p = 0; p < g->M; ++p) {
        int64_t q[3]; unc(g, p, q);
        for (int u = 0; u < 3; ++u) {
            int64_t row = dof(g, u, p);
            for (int t = 0; t < 2; ++t) {
                int v = (u + 1 + t) % 3, w = (u + 2 - t) % 3;
                double sg = t == 0 ? 1.0 : -1.0;     /* H_u = ∂_v E_w - ∂_w E_v */
                int isfwd = which == 0 ? (g->bft[v] == EE) : (g->bft[v] == HH);
                int gi = g0type(g, ftout, v);          /* ∆l at output planes */
                oc64 dv = sg / g->sdl[gi][v][q[v]];
                int64_t Nv = g->N[v], st = v == 0 ? 1 : (v == 1 ? g->N[0] : g->N[0] * g->N[1]);
                if (isfwd) {
                    tpush(L, row, dof(g, w, p), -dv);
                    if (q[v] < Nv - 1) tpush(L, row, dof(g, w, p + st), dv);
                    else if (g->bloch[v]) tpush(L, row, dof(g, w, p - (Nv - 1) * st), dv * g->ph[v]);
                } else {
                    tpush(L, row, dof(g, w, p), dv);
                    if (q[v] > 0) tpush(L, row, dof(g, w, p - st), -dv);
                    else if (g->bloch[v]) tpush(L, row, dof(g, w, p + (Nv - 1) * st), -dv / g->ph[v]);
                }
            }
        }
    }
}

int og_curl(const og_desc *d, int which, og_csc *out) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    tlist L = {0};
    curl_trip(&g, which, &L);
    csc_from_trip(3 * g.M, &L, out);
    free(L.t); g_free(&g);
    return 0;
}

/* ------------------------------------------------------- param operator */

static int has_offdiag(const G *g, const oc64 *parr) {
    for (int v = 0; v < 3; ++v)
        for (int w = 0; w < 3; ++w) {
            if (v == w) continue;
            const oc64 *a = parr + g->M * (v + 3 * w);
            for (int64_t p = 0; p < g->M; ++p) if (a[p] != 0) return 1;
        }
    return 0;
}

/* Pε (which = 0) / Pμ (which = 1), model.jl:149-155.  Diagonal entries map
 * F_w to F_w.  An off-diagonal entry (v,w), stored at the voxel corners of
 * the field's corner grid g0, takes F_w to the corners with a ∆l-weighted
 * average along w (direction isfwd_in[w], weights s∆l at the field planes over
 * 2 s∆l at the corner planes), multiplies by the tensor entry, and takes the
 * corner values to F_v with an unweighted half-half average along v in the
 * opposite direction.  Off-diagonal blocks are stored (explicit zeros
 * included) only when some off-diagonal entry of the array is nonzero. */
static void paramop_trip(const G *g, int which, const oc64 *parr, tlist *L) {
    int ft = which == 0 ? EE : HH, full = has_offdiag(g, parr);
    int isfwd_in[3];
    for (int a = 0; a < 3; ++a) isfwd_in[a] = which == 0 ? (g->bft[a] != EE) : (g->bft[a] != HH);
    for (int64_t p = 0; p < g->M; ++p) {
        int64_t q[3]; unc(g, p, q);
        for (int v = 0; v < 3; ++v) {
            int64_t row = dof(g, v, p);
            tpush(L, row, row, parr[p + g->M * (v + 3 * v)]);
            if (!full) continue;
            for (int w = 0; w < 3; ++w) {
                if (w == v) continue;
                int gc = g0type(g, ft, w), gf = alt(gc);   /* corner / field grid type along w */
                int64_t stv = v == 0 ? 1 : (v == 1 ? g->N[0] : g->N[0] * g->N[1]);
                int64_t stw = w == 0 ? 1 : (w == 1 ? g->N[0] : g->N[0] * g->N[1]);
                for (int oc = 0; oc < 2; ++oc) {
                    int64_t cp = p; oc64 fo = 0.5;
                    if (oc == 1) {
                        if (!isfwd_in[v]) {        /* output average forward along v */
                            if (q[v] < g->N[v] - 1) cp = p + stv;
                            else if (g->bloch[v]) { cp = p - (g->N[v] - 1) * stv; fo = 0.5 * g->ph[v]; }
                            else continue;
                        } else {
                            if (q[v] > 0) cp = p - stv;
                            else if (g->bloch[v]) { cp = p + (g->N[v] - 1) * stv; fo = 0.5 / g->ph[v]; }
                            else continue;
                        }
                    }
                    int64_t cq[3]; unc(g, cp, cq);
                    oc64 ec = parr[cp + g->M * (v + 3 * w)];
                    oc64 inv = 1.0 / g->sdl[gc][w][cq[w]];
                    for (int ic = 0; ic < 2; ++ic) {
                        int64_t fp = cp, fi = cq[w]; oc64 ph = 1.0;
                        if (ic == 1) {
                            if (!isfwd_in[w]) {
                                if (cq[w] > 0) { fp = cp - stw; fi = cq[w] - 1; }
                                else if (g->bloch[w]) { fp = cp + (g->N[w] - 1) * stw; fi = g->N[w] - 1; ph = 1.0 / g->ph[w]; }
                                else continue;
                            } else {
                                if (cq[w] < g->N[w] - 1) { fp = cp + stw; fi = cq[w] + 1; }
                                else if (g->bloch[w]) { fp = cp - (g->N[w] - 1) * stw; fi = 0; ph = g->ph[w]; }
                                else continue;
                            }
                        }
                        oc64 wi = (0.5 * g->sdl[gf][w][fi] * inv) * ph;
                        tpush(L, row, dof(g, w, fp), ec * (fo * wi));
                    }
                }
            }
        }
    }
}

/* create_Mcs (model.jl:287-306) -> MaxwellBase create_mean (absent):
 * which 0 = Mcₑ (isfwd = boundft .!= EE, ∆l = s∆lₘ, ∆l′⁻¹ = s∆lₑ⁻¹),
 * which 1 = Mcₘ (isfwd = boundft .!= HH, ∆l = s∆lₑ, ∆l′⁻¹ = s∆lₘ⁻¹).
 * Row (w, corner c) averages F_w over the two field points beside c along w
 * (c − 1 and c going backward, c and c + 1 going forward), each weighted by
 * its s∆l over 2 s∆l of the corner plane; a Bloch wrap carries e⁻ⁱᵏᴸ (or its
 * inverse), a non-Bloch ghost drops.  Formulated per row (corner), unlike the
 * device's per-column enumeration. */
int og_Mcs(const og_desc *d, int which, og_csc *out) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    int ft = which == 0 ? EE : HH;
    tlist L = {0};
    for (int64_t c = 0; c < g.M; ++c) {
        int64_t q[3]; unc(&g, c, q);
        for (int w = 0; w < 3; ++w) {
            int isfwd = which == 0 ? (g.bft[w] != EE) : (g.bft[w] != HH);
            int gc = g0type(&g, ft, w), gf = alt(gc);
            int64_t st = w == 0 ? 1 : (w == 1 ? g.N[0] : g.N[0] * g.N[1]);
            oc64 inv = 1.0 / g.sdl[gc][w][q[w]];
            int64_t row = dof(&g, w, c);
            /* the field point at the corner's own index */
            tpush(&L, row, row, 0.5 * g.sdl[gf][w][q[w]] * inv);
            /* the other field point: index q-1 (backward) or q+1 (forward) */
            int64_t nb = q[w] + (isfwd ? 1 : -1);
            if (nb >= 0 && nb < g.N[w]) {
                tpush(&L, row, dof(&g, w, c + (isfwd ? st : -st)), 0.5 * g.sdl[gf][w][nb] * inv);
            } else if (g.bloch[w]) {
                int64_t wi = isfwd ? 0 : g.N[w] - 1;
                oc64 ph = isfwd ? g.ph[w] : 1.0 / g.ph[w];
                tpush(&L, row, dof(&g, w, c + (wi - q[w]) * st), 0.5 * g.sdl[gf][w][wi] * inv * ph);
            }
        }
    }
    csc_from_trip(3 * g.M, &L, out);
    free(L.t); g_free(&g);
    return 0;
}

int og_paramop(const og_desc *d, int which, const oc64 *parr, og_csc *out) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    tlist L = {0};
    paramop_trip(&g, which, parr, &L);
    csc_from_trip(3 * g.M, &L, out);
    free(L.t); g_free(&g);
    return 0;
}

/* -------------------------------------------------------- A and b */

/* C = A*B (Gustavson, structural entries kept as Julia's spmatmul does). */
static void spgemm(int64_t n, const og_csc *A, const og_csc *B, og_csc *C) {
    int64_t cap = B->nnz * 4 + 16, nz = 0;
    int64_t *colptr = malloc(sizeof(int64_t) * (n + 1));
    int64_t *rowval = malloc(sizeof(int64_t) * cap);
    oc64 *nzval = malloc(sizeof(oc64) * cap);
    int64_t *mark = malloc(sizeof(int64_t) * n);
    oc64 *acc = calloc(n, sizeof(oc64));
    for (int64_t i = 0; i < n; ++i) mark[i] = -1;
    colptr[0] = 1;
    for (int64_t j = 0; j < n; ++j) {
        int64_t start = nz;
        for (int64_t kb = B->colptr[j] - 1; kb < B->colptr[j + 1] - 1; ++kb) {
            int64_t k = B->rowval[kb] - 1; oc64 bv = B->nzval[kb];
            for (int64_t ka = A->colptr[k] - 1; ka < A->colptr[k + 1] - 1; ++ka) {
                int64_t i = A->rowval[ka] - 1;
                if (mark[i] != j) {
                    mark[i] = j;
                    if (nz == cap) { cap *= 2; rowval = realloc(rowval, sizeof(int64_t) * cap); nzval = realloc(nzval, sizeof(oc64) * cap); }
                    rowval[nz++] = i; acc[i] = A->nzval[ka] * bv;
                } else acc[i] += A->nzval[ka] * bv;
            }
        }
        for (int64_t x = start + 1; x < nz; ++x) {      /* sort rows of this column */
            int64_t r = rowval[x], y = x - 1;
            while (y >= start && rowval[y] > r) { rowval[y + 1] = rowval[y]; --y; }
            rowval[y + 1] = r;
        }
        for (int64_t x = start; x < nz; ++x) { nzval[x] = acc[rowval[x]]; rowval[x] += 1; }
        colptr[j + 1] = nz + 1;
    }
    C->nnz = nz; C->colptr = colptr; C->rowval = rowval; C->nzval = nzval;
    free(mark); free(acc);
}

/* C = A - s*B, storing only nonzero results (Julia's sparse map drops zeros). */
static void sub_dropzeros(int64_t n, const og_csc *A, oc64 s, const og_csc *B, og_csc *C) {
    int64_t cap = A->nnz + B->nnz + 1, nz = 0;
    int64_t *colptr = malloc(sizeof(int64_t) * (n + 1));
    int64_t *rowval = malloc(sizeof(int64_t) * cap);
    oc64 *nzval = malloc(sizeof(oc64) * cap);
    colptr[0] = 1;
    for (int64_t j = 0; j < n; ++j) {
        int64_t a = A->colptr[j] - 1, ae = A->colptr[j + 1] - 1;
        int64_t b = B->colptr[j] - 1, be = B->colptr[j + 1] - 1;
        while (a < ae || b < be) {
            int64_t ra = a < ae ? A->rowval[a] : INT64_MAX, rb = b < be ? B->rowval[b] : INT64_MAX;
            oc64 v; int64_t r;
            if (ra == rb) { v = A->nzval[a] - s * B->nzval[b]; r = ra; ++a; ++b; }
            else if (ra < rb) { v = A->nzval[a] - 0.0; r = ra; ++a; }
            else { v = 0.0 - s * B->nzval[b]; r = rb; ++b; }
            if (v != 0) { rowval[nz] = r; nzval[nz] = v; ++nz; }
        }
        colptr[j + 1] = nz + 1;
    }
    C->nnz = nz; C->colptr = colptr; C->rowval = rowval; C->nzval = nzval;
}

/* diagonal of a diagonal parameter operator in DOF order (Pμ for EE, Pε for HH) */
static int mu_diag(const G *g, const oc64 *mu_arr, oc64 *md) {
    if (has_offdiag(g, mu_arr)) return -3;   /* P \ C needs a diagonal P (model.jl:236, 239) */
    for (int c = 0; c < 3; ++c)
        for (int64_t p = 0; p < g->M; ++p) md[dof(g, c, p)] = mu_arr[p + g->M * (c + 3 * c)];
    return 0;
}

/* model.jl:235-237: A = Cₘ*(Pμ\Cₑ); iszero(ω) || (A -= ω^2*Pε) */
int og_create_A(const og_desc *d, double om_re, double om_im, const oc64 *eps_arr,
                const oc64 *mu_arr, og_csc *out) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    int64_t n = 3 * g.M;
    oc64 *md = malloc(sizeof(oc64) * n);
    rc = mu_diag(&g, mu_arr, md);
    if (rc) { free(md); g_free(&g); return rc; }
    tlist L = {0};
    og_csc Ce, Cm, C;
    curl_trip(&g, 0, &L); csc_from_trip(n, &L, &Ce); L.n = 0;
    curl_trip(&g, 1, &L); csc_from_trip(n, &L, &Cm); L.n = 0;
    for (int64_t e = 0; e < Ce.nnz; ++e) Ce.nzval[e] = Ce.nzval[e] / md[Ce.rowval[e] - 1];
    spgemm(n, &Cm, &Ce, &C);
    oc64 om = om_re + om_im * I;
    if (om != 0) {
        og_csc Pe;
        paramop_trip(&g, 0, eps_arr, &L); csc_from_trip(n, &L, &Pe);
        sub_dropzeros(n, &C, om * om, &Pe, out);
        og_csc_free(&Pe); og_csc_free(&C);
    } else *out = C;
    og_csc_free(&Ce); og_csc_free(&Cm);
    free(L.t); free(md); g_free(&g);
    return 0;
}

/* model.jl:238-240: A = Cₑ*(Pε\Cₘ); iszero(ω) || (A -= ω^2*Pμ) (needs a diagonal Pε) */
int og_create_A_hh(const og_desc *d, double om_re, double om_im, const oc64 *eps_arr,
                   const oc64 *mu_arr, og_csc *out) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    int64_t n = 3 * g.M;
    oc64 *ed = malloc(sizeof(oc64) * n);
    rc = mu_diag(&g, eps_arr, ed);
    if (rc) { free(ed); g_free(&g); return rc; }
    tlist L = {0};
    og_csc Ce, Cm, C;
    curl_trip(&g, 0, &L); csc_from_trip(n, &L, &Ce); L.n = 0;
    curl_trip(&g, 1, &L); csc_from_trip(n, &L, &Cm); L.n = 0;
    for (int64_t e = 0; e < Cm.nnz; ++e) Cm.nzval[e] = Cm.nzval[e] / ed[Cm.rowval[e] - 1];
    spgemm(n, &Ce, &Cm, &C);
    oc64 om = om_re + om_im * I;
    if (om != 0) {
        og_csc Pm;
        paramop_trip(&g, 1, mu_arr, &L); csc_from_trip(n, &L, &Pm);
        sub_dropzeros(n, &C, om * om, &Pm, out);
        og_csc_free(&Pm); og_csc_free(&C);
    } else *out = C;
    og_csc_free(&Ce); og_csc_free(&Cm);
    free(L.t); free(ed); g_free(&g);
    return 0;
}

int og_spmv(int64_t n, const og_csc *A, const oc64 *x, oc64 *y) {
    for (int64_t i = 0; i < n; ++i) y[i] = 0;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = A->colptr[j] - 1; k < A->colptr[j + 1] - 1; ++k)
            y[A->rowval[k] - 1] += A->nzval[k] * x[j];
    return 0;
}

/* model.jl:262-265: b = Cₘ*(Pμ\jₘ); b .*= -1; iszero(ω) || (b .-= (im*ω).*jₑ) */
int og_create_b(const og_desc *d, double om_re, double om_im, const oc64 *mu_arr,
                const oc64 *je, const oc64 *jm, oc64 *b) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    int64_t n = 3 * g.M;
    oc64 *md = malloc(sizeof(oc64) * n), *y = malloc(sizeof(oc64) * n);
    rc = mu_diag(&g, mu_arr, md);
    if (rc) { free(md); free(y); g_free(&g); return rc; }
    tlist L = {0}; og_csc Cm;
    curl_trip(&g, 1, &L); csc_from_trip(n, &L, &Cm);
    for (int64_t i = 0; i < n; ++i) y[i] = jm[i] / md[i];
    og_spmv(n, &Cm, y, b);
    oc64 iw = (om_re + om_im * I) * I;
    for (int64_t i = 0; i < n; ++i) {
        b[i] = b[i] * -1.0;
        if (om_re != 0 || om_im != 0) b[i] = b[i] - iw * je[i];
    }
    og_csc_free(&Cm); free(L.t); free(md); free(y); g_free(&g);
    return 0;
}

/* model.jl:266-268: b = Cₑ*(Pε\jₑ); iszero(ω) || (b .-= (im*ω).*jₘ) */
int og_create_b_hh(const og_desc *d, double om_re, double om_im, const oc64 *eps_arr,
                   const oc64 *je, const oc64 *jm, oc64 *b) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    int64_t n = 3 * g.M;
    oc64 *ed = malloc(sizeof(oc64) * n), *y = malloc(sizeof(oc64) * n);
    rc = mu_diag(&g, eps_arr, ed);
    if (rc) { free(ed); free(y); g_free(&g); return rc; }
    tlist L = {0}; og_csc Ce;
    curl_trip(&g, 0, &L); csc_from_trip(n, &L, &Ce);
    for (int64_t i = 0; i < n; ++i) y[i] = je[i] / ed[i];
    og_spmv(n, &Ce, y, b);
    oc64 iw = (om_re + om_im * I) * I;
    if (om_re != 0 || om_im != 0)
        for (int64_t i = 0; i < n; ++i) b[i] = b[i] - iw * jm[i];
    og_csc_free(&Ce); free(L.t); free(ed); free(y); g_free(&g);
    return 0;
}

/* ------------------------------------------------- material parameters */

static int inside(const og_shape *s, const double x[3]) {
    if (s->kind == 0) {
        double a = x[0] - s->c[0], b = x[1] - s->c[1], c = x[2] - s->c[2];
        return a * a + b * b + c * c <= s->r * s->r;
    }
    if (s->kind == 2) {   /* cylinder along z, radius r, height s[2] */
        double a = x[0] - s->c[0], b = x[1] - s->c[1];
        return a * a + b * b <= s->r * s->r && fabs(x[2] - s->c[2]) <= 0.5 * s->s[2];
    }
    for (int a = 0; a < 3; ++a) if (!(fabs(x[a] - s->c[a]) <= 0.5 * s->s[a])) return 0;
    return 1;
}

/* nearest surface point r0 and outward unit normal n of a shape near x */
static void surfpt(const og_shape *s, const double x[3], double r0[3], double n[3]) {
    if (s->kind == 0) {
        double d[3] = {x[0] - s->c[0], x[1] - s->c[1], x[2] - s->c[2]};
        double nd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (nd == 0.0) { n[0] = 1; n[1] = 0; n[2] = 0; }
        else { n[0] = d[0] / nd; n[1] = d[1] / nd; n[2] = d[2] / nd; }
        for (int a = 0; a < 3; ++a) r0[a] = s->c[a] + s->r * n[a];
        return;
    }
    if (s->kind == 2) {   /* the nearer of the lateral surface and an end cap */
        double d0 = x[0] - s->c[0], d1 = x[1] - s->c[1], dz = x[2] - s->c[2];
        double nd = sqrt(d0 * d0 + d1 * d1);
        if (nd - s->r >= fabs(dz) - 0.5 * s->s[2]) {
            if (nd == 0.0) { n[0] = 1; n[1] = 0; } else { n[0] = d0 / nd; n[1] = d1 / nd; }
            n[2] = 0;
            r0[0] = s->c[0] + s->r * n[0]; r0[1] = s->c[1] + s->r * n[1]; r0[2] = x[2];
        } else {
            double sg = dz < 0 ? -1.0 : 1.0;
            n[0] = n[1] = 0; n[2] = sg;
            r0[0] = x[0]; r0[1] = x[1]; r0[2] = s->c[2] + sg * 0.5 * s->s[2];
        }
        return;
    }
    int ax = 0; double best = -INFINITY;
    for (int a = 0; a < 3; ++a) {
        double e = fabs(x[a] - s->c[a]) - 0.5 * s->s[a];
        if (e > best) { best = e; ax = a; }
    }
    double sg = (x[ax] - s->c[ax]) < 0 ? -1.0 : 1.0;
    for (int a = 0; a < 3; ++a) { r0[a] = x[a]; n[a] = 0.0; }
    n[ax] = sg; r0[ax] = s->c[ax] + sg * 0.5 * s->s[ax];
}

static double vf_core(double b, double a1, double a2, double a3) {
    /* fraction of the unit cube with a1 t1 + a2 t2 + a3 t3 <= b, a1>=a2>=a3>=0,
     * 0 < b <= (a1+a2+a3)/2 */
    double v;
    if (a2 <= 1e-12 * a1) v = b / a1;
    else if (a3 <= 1e-12 * a1) {
        double t = b - a2 > 0 ? b - a2 : 0.0;
        v = (b * b - t * t) / (2.0 * a1 * a2);
    } else {
        double t1 = b - a1 > 0 ? b - a1 : 0.0, t2 = b - a2 > 0 ? b - a2 : 0.0, t3 = b - a3 > 0 ? b - a3 : 0.0;
        double t13 = b - a1 - a3 > 0 ? b - a1 - a3 : 0.0, t23 = b - a2 - a3 > 0 ? b - a2 - a3 : 0.0;
        v = (b * b * b - t1 * t1 * t1 - t2 * t2 * t2 - t3 * t3 * t3 + t13 * t13 * t13 + t23 * t23 * t23)
            / (6.0 * a1 * a2 * a3);
    }
    return v < 0 ? 0 : (v > 1 ? 1 : v);
}

/* Volume fraction of the box [vmin,vmax] on the side n·(x-r0) <= 0. */
double og_volfrac(const double vmin[3], const double vmax[3], const double n[3], const double r0[3]) {
    double a[3], b = 0.0;
    for (int i = 0; i < 3; ++i) { a[i] = n[i] * (vmax[i] - vmin[i]); b += n[i] * (r0[i] - vmin[i]); }
    for (int i = 0; i < 3; ++i) if (a[i] < 0) { b -= a[i]; a[i] = -a[i]; }
    /* sort descending */
    if (a[0] < a[1]) { double t = a[0]; a[0] = a[1]; a[1] = t; }
    if (a[1] < a[2]) { double t = a[1]; a[1] = a[2]; a[2] = t; }
    if (a[0] < a[1]) { double t = a[0]; a[0] = a[1]; a[1] = t; }
    double s = a[0] + a[1] + a[2];
    if (s <= 0.0) return b >= 0.0 ? 1.0 : 0.0;
    if (b <= 0.0) return 0.0;
    if (b >= s) return 1.0;
    if (2.0 * b > s) return 1.0 - vf_core(s - b, a[0], a[1], a[2]);
    return vf_core(b, a[0], a[1], a[2]);
}

/* Kottke average of two 3x3 tensors (column-major) with fill fraction r of
 * efg and unit normal n: τ-transform in the frame (n, t1, t2), arithmetic
 * average, inverse τ-transform. */
void og_kottke(const oc64 *efg, const oc64 *ebg, double r, const double n[3], oc64 *out) {
    double B[3][3];   /* B[i][a]: component i of basis vector a */
    int k = 0;
    double an[3] = {fabs(n[0]), fabs(n[1]), fabs(n[2])};
    if (an[1] < an[k]) k = 1;
    if (an[2] < an[k]) k = 2;
    double e[3] = {0, 0, 0}; e[k] = 1.0;
    double t1[3] = {n[1] * e[2] - n[2] * e[1], n[2] * e[0] - n[0] * e[2], n[0] * e[1] - n[1] * e[0]};
    double nt = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    for (int i = 0; i < 3; ++i) t1[i] /= nt;
    double t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
    for (int i = 0; i < 3; ++i) { B[i][0] = n[i]; B[i][1] = t1[i]; B[i][2] = t2[i]; }
    oc64 tau[3][3] = {{0}};
    for (int m = 0; m < 2; ++m) {
        const oc64 *E = m == 0 ? efg : ebg;
        double f = m == 0 ? r : 1.0 - r;
        oc64 P[3][3];   /* P = B^T E B */
        for (int a = 0; a < 3; ++a)
            for (int b2 = 0; b2 < 3; ++b2) {
                oc64 s = 0;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) s += B[i][a] * E[i + 3 * j] * B[j][b2];
                P[a][b2] = s;
            }
        oc64 t[3][3];
        t[0][0] = -1.0 / P[0][0];
        t[0][1] = P[0][1] / P[0][0]; t[0][2] = P[0][2] / P[0][0];
        t[1][0] = P[1][0] / P[0][0]; t[2][0] = P[2][0] / P[0][0];
        for (int a = 1; a < 3; ++a)
            for (int b2 = 1; b2 < 3; ++b2) t[a][b2] = P[a][b2] - P[a][0] * P[0][b2] / P[0][0];
        for (int a = 0; a < 3; ++a)
            for (int b2 = 0; b2 < 3; ++b2) tau[a][b2] += f * t[a][b2];
    }
    oc64 Q[3][3];
    Q[0][0] = -1.0 / tau[0][0];
    Q[0][1] = -tau[0][1] / tau[0][0]; Q[0][2] = -tau[0][2] / tau[0][0];
    Q[1][0] = -tau[1][0] / tau[0][0]; Q[2][0] = -tau[2][0] / tau[0][0];
    for (int a = 1; a < 3; ++a)
        for (int b2 = 1; b2 < 3; ++b2) Q[a][b2] = tau[a][b2] - tau[a][0] * tau[0][b2] / tau[0][0];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            oc64 s = 0;
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2) s += B[i][a] * Q[a][b2] * B[j][b2];
            out[i + 3 * j] = s;
        }
}

/* object index lattice for grid types gt[3] (ghosted sizes N+1), classified
 * at the transformed ghosted coordinates; 0 = no object, the last-added
 * object containing the point wins. */
static void classify(const G *g, const int gt[3], int64_t nobj, const og_shape *sh, int32_t *oind) {
    int64_t Nx = g->N[0] + 1, Ny = g->N[1] + 1, Nz = g->N[2] + 1;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < Nz; ++k)
        for (int64_t j = 0; j < Ny; ++j)
            for (int64_t i = 0; i < Nx; ++i) {
                double x[3] = {g->tlg[gt[0]][0][i], g->tlg[gt[1]][1][j], g->tlg[gt[2]][2][k]};
                int32_t o = 0;
                for (int64_t b = nobj; b >= 1; --b) if (inside(&sh[b - 1], x)) { o = (int32_t)b; break; }
                oind[i + Nx * (j + Ny * k)] = o;
            }
}

/* one parameter-array pass: entries (v,w) listed in ent, located at grid
 * types fgt (field point), smoothed over voxels whose corners are the
 * opposite grid types. */
static void smooth_pass(const G *g, int nent, const int ent[][2], const int fgt[3],
                        const int32_t *oind_ctr, const int32_t *oind_cnr, int64_t nobj,
                        const og_shape *sh, const oc64 *tens, oc64 *arr) {
    int64_t Nx1 = g->N[0] + 1, Ny1 = g->N[1] + 1;
    (void)nobj;
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < g->M; ++p) {
        int64_t q[3]; unc(g, p, q);
        int64_t t[3];
        for (int a = 0; a < 3; ++a) t[a] = fgt[a] == PRIM ? q[a] : q[a] + 1;   /* ghosted index */
        int32_t oc = oind_ctr[t[0] + Nx1 * (t[1] + Ny1 * t[2])];
        int32_t co[8]; int nd = 0; int32_t dist[8];
        for (int c = 0; c < 8; ++c) {
            int64_t ci = q[0] + (c & 1), cj = q[1] + ((c >> 1) & 1), ck = q[2] + ((c >> 2) & 1);
            co[c] = oind_cnr[ci + Nx1 * (cj + Ny1 * ck)];
            int f = 0;
            for (int x = 0; x < nd; ++x) if (dist[x] == co[c]) f = 1;
            if (!f) dist[nd++] = co[c];
        }
        oc64 val[9];
        int smoothed = 0;
        if (nd == 2 && dist[0] > 0 && dist[1] > 0) {
            int32_t fg = dist[0] > dist[1] ? dist[0] : dist[1], bg = dist[0] > dist[1] ? dist[1] : dist[0];
            const oc64 *ef = tens + 9 * (fg - 1), *eb = tens + 9 * (bg - 1);
            int same = 1;
            for (int x = 0; x < 9; ++x) if (ef[x] != eb[x]) same = 0;
            if (!same) {
                double x0[3], vmin[3], vmax[3], r0[3], nn[3];
                for (int a = 0; a < 3; ++a) {
                    x0[a] = g->l[fgt[a]][a][q[a]];
                    vmin[a] = g->lg[alt(fgt[a])][a][q[a]];
                    vmax[a] = g->lg[alt(fgt[a])][a][q[a] + 1];
                }
                surfpt(&sh[fg - 1], x0, r0, nn);
                double r = og_volfrac(vmin, vmax, nn, r0);
                og_kottke(ef, eb, r, nn, val);
                smoothed = 1;
            }
        } else if (nd >= 2) {
            for (int x = 0; x < 9; ++x) {
                oc64 s = 0;
                for (int c = 0; c < 8; ++c) if (co[c] > 0) s += tens[9 * (co[c] - 1) + x];
                val[x] = s / 8.0;
            }
            smoothed = 1;
        }
        for (int e = 0; e < nent; ++e) {
            int v = ent[e][0], w = ent[e][1];
            oc64 out;
            if (smoothed) out = val[v + 3 * w];
            else out = oc > 0 ? tens[9 * (oc - 1) + v + 3 * w] : 0.0;
            arr[p + g->M * (v + 3 * w)] = out;
        }
    }
}

/* full.jl:16-70: assignment + smoothing of the diagonal entries at the F_w
 * locations and the off-diagonal entries at the voxel corners g0. */
int og_matparams(const og_desc *d, int64_t nobj, const og_shape *shapes,
                 const oc64 *eps_t, const oc64 *mu_t, oc64 *eps_arr, oc64 *mu_arr) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
    int64_t nl = (g.N[0] + 1) * (g.N[1] + 1) * (g.N[2] + 1);
    int32_t *lat[8];
    for (int c = 0; c < 8; ++c) {
        int gt[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
        lat[c] = malloc(sizeof(int32_t) * nl);
        classify(&g, gt, nobj, shapes, lat[c]);
    }
    for (int m = 0; m < 2; ++m) {
        int ft = m == 0 ? EE : HH;
        const oc64 *tens = m == 0 ? eps_t : mu_t;
        oc64 *arr = m == 0 ? eps_arr : mu_arr;
        for (int w = 0; w < 3; ++w) {
            int fgt[3], cgt[3];
            for (int a = 0; a < 3; ++a) { fgt[a] = gtype(&g, ft, w, a); cgt[a] = alt(fgt[a]); }
            int ent[1][2] = {{w, w}};
            smooth_pass(&g, 1, ent, fgt, lat[fgt[0] + 2 * fgt[1] + 4 * fgt[2]],
                        lat[cgt[0] + 2 * cgt[1] + 4 * cgt[2]], nobj, shapes, tens, arr);
        }
        int fgt[3], cgt[3];
        for (int a = 0; a < 3; ++a) { fgt[a] = g0type(&g, ft, a); cgt[a] = alt(fgt[a]); }
        int ent[6][2] = {{1, 0}, {2, 0}, {0, 1}, {2, 1}, {0, 2}, {1, 2}};
        smooth_pass(&g, 6, ent, fgt, lat[fgt[0] + 2 * fgt[1] + 4 * fgt[2]],
                    lat[cgt[0] + 2 * cgt[1] + 4 * cgt[2]], nobj, shapes, tens, arr);
    }
    for (int c = 0; c < 8; ++c) free(lat[c]);
    g_free(&g);
    return 0;
}

/* ----------------------------------------------------------- CPU solver */

/* Symmetric QMR (Freund & Nachtigal 1994, no look-ahead) on
 *   P = Q T A T^-1 Q,   T = S_E^{1/2} (separable), Q = diag(A)^{-1/2},
 * which is complex symmetric because A^T = S_E A S_E^-1.  Stops when the
 * true FP64 residual ||b - A x|| / ||b|| <= tol.  The CSC matrix is
 * transposed once to CSR so that the product runs in parallel. */
typedef struct { int64_t n; int64_t *rp; int64_t *ci; oc64 *v; } csr_t;

static void csc2csr(int64_t n, const og_csc *A, csr_t *R) {
    R->n = n;
    R->rp = calloc(n + 1, sizeof(int64_t));
    R->ci = malloc(sizeof(int64_t) * (A->nnz ? A->nnz : 1));
    R->v = malloc(sizeof(oc64) * (A->nnz ? A->nnz : 1));
    for (int64_t k = 0; k < A->nnz; ++k) R->rp[A->rowval[k]]++;
    for (int64_t i = 0; i < n; ++i) R->rp[i + 1] += R->rp[i];
    int64_t *pos = malloc(sizeof(int64_t) * (n + 1));
    memcpy(pos, R->rp, sizeof(int64_t) * (n + 1));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = A->colptr[j] - 1; k < A->colptr[j + 1] - 1; ++k) {
            int64_t i = A->rowval[k] - 1, d = pos[i]++;
            R->ci[d] = j; R->v[d] = A->nzval[k];
        }
    free(pos);
}

static void csr_mv(const csr_t *R, const oc64 *x, oc64 *y) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < R->n; ++i) {
        oc64 s = 0;
        for (int64_t k = R->rp[i]; k < R->rp[i + 1]; ++k) s += R->v[k] * x[R->ci[k]];
        y[i] = s;
    }
}

static double resid(const csr_t *R, const oc64 *b, const oc64 *x, oc64 *w, double bn) {
    csr_mv(R, x, w);
    double s = 0;
    #pragma omp parallel for reduction(+:s)
    for (int64_t i = 0; i < R->n; ++i) { oc64 e = b[i] - w[i]; s += creal(e) * creal(e) + cimag(e) * cimag(e); }
    return sqrt(s) / bn;
}

int og_solve_sqmr(const og_desc *d, const og_csc *A, const oc64 *b, oc64 *x,
                  double tol, int64_t maxit, int nthreads, int64_t *iters, double *relres) {
    G g; int rc = g_build(d, &g); if (rc) return rc;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t n = 3 * g.M;
    csr_t R; csc2csr(n, A, &R);
    oc64 *sc = malloc(sizeof(oc64) * n), *sr = malloc(sizeof(oc64) * n);
    oc64 *z = calloc(n, sizeof(oc64)), *r = malloc(sizeof(oc64) * n);
    oc64 *qv = malloc(sizeof(oc64) * n), *t = malloc(sizeof(oc64) * n);
    oc64 *dd = calloc(n, sizeof(oc64)), *w = malloc(sizeof(oc64) * n);
    for (int64_t p = 0; p < g.M; ++p) {
        int64_t q[3]; unc(&g, p, q);
        for (int cc = 0; cc < 3; ++cc) {
            oc64 T = 1.0;
            for (int a = 0; a < 3; ++a) T *= csqrt(g.sdl[gtype(&g, EE, cc, a)][a][q[a]]);
            int64_t i = dof(&g, cc, p);
            oc64 dg = 0;
            for (int64_t k = R.rp[i]; k < R.rp[i + 1]; ++k) if (R.ci[k] == i) dg = R.v[k];
            oc64 qq = 1.0 / csqrt(dg);
            sc[i] = qq * T; sr[i] = qq / T;
        }
    }
    double bn = 0, tau = 0, rre = 0, rim = 0;
    #pragma omp parallel for reduction(+:bn,tau,rre,rim)
    for (int64_t i = 0; i < n; ++i) {
        bn += creal(b[i]) * creal(b[i]) + cimag(b[i]) * cimag(b[i]);
        oc64 c = sc[i] * b[i]; r[i] = c; qv[i] = c;
        tau += creal(c) * creal(c) + cimag(c) * cimag(c);
        oc64 cc = c * c; rre += creal(cc); rim += cimag(cc);
    }
    bn = sqrt(bn); tau = sqrt(tau);
    oc64 rho = rre + rim * I;
    double cn = tau, theta = 0, rr = 1.0;
    int64_t it = 0, next_check = 0;
    while (it < maxit) {
        #pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) w[i] = sr[i] * qv[i];
        csr_mv(&R, w, t);
        double sre = 0, sim = 0;
        #pragma omp parallel for reduction(+:sre,sim)
        for (int64_t i = 0; i < n; ++i) { t[i] *= sc[i]; oc64 e = qv[i] * t[i]; sre += creal(e); sim += cimag(e); }
        oc64 sig = sre + sim * I;
        if (sig == 0) break;
        oc64 alpha = rho / sig;
        double rn = 0; rre = 0; rim = 0;
        #pragma omp parallel for reduction(+:rn,rre,rim)
        for (int64_t i = 0; i < n; ++i) {
            r[i] -= alpha * t[i];
            rn += creal(r[i]) * creal(r[i]) + cimag(r[i]) * cimag(r[i]);
            oc64 e = r[i] * r[i]; rre += creal(e); rim += cimag(e);
        }
        rn = sqrt(rn);
        oc64 rho1 = rre + rim * I;
        double th1 = rn / tau, cc = 1.0 / sqrt(1.0 + th1 * th1);
        tau = tau * th1 * cc;
        double c2t2 = cc * cc * theta * theta;
        oc64 c2a = cc * cc * alpha;
        oc64 beta = rho1 / rho; rho = rho1; theta = th1;
        #pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) {
            dd[i] = c2t2 * dd[i] + c2a * qv[i];
            z[i] += dd[i];
            qv[i] = r[i] + beta * qv[i];
        }
        ++it;
        if (tau <= tol * cn && it >= next_check) {
            #pragma omp parallel for
            for (int64_t i = 0; i < n; ++i) x[i] = sr[i] * z[i];
            rr = resid(&R, b, x, w, bn);
            if (rr <= tol) break;
            next_check = it + 10;
        }
    }
    #pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) x[i] = sr[i] * z[i];
    rr = resid(&R, b, x, w, bn);
    if (iters) *iters = it;
    if (relres) *relres = rr;
    free(sc); free(sr); free(z); free(r); free(qv); free(t); free(dd); free(w);
    free(R.rp); free(R.ci); free(R.v);
    g_free(&g);
    return rr <= tol ? 0 : 1;
}
