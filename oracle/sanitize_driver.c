/*
 * Drives every oracle entry point on small grids (a PML box, a Bloch grid with
 * a complex phase, nonuniform spacing, component-last order, both boundary
 * field types and both formulations) so that an AddressSanitizer and
 * UndefinedBehaviorSanitizer build of oracle.c reports any out-of-bounds
 * access, leak or undefined operation.  Test infrastructure, like oracle.c;
 * built and run by `make -C oracle sanitize` (tests/test_oracle_sanitize.py).
 * Exit status 0 and "sanitize driver ok" when every call returned 0.
 */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CHECK(x)                                                                \
    do {                                                                        \
        int rc_ = (x);                                                          \
        if (rc_ != 0) { fprintf(stderr, "%s returned %d\n", #x, rc_); return 1; } \
    } while (0)

static int run_case(const int64_t N[3], int bloch, int bft_hh_axis, int cmpfirst, int nonuni, int full) {
    double *lp[3];
    og_desc d;
    memset(&d, 0, sizeof d);
    for (int a = 0; a < 3; ++a) {
        d.N[a] = N[a];
        lp[a] = malloc(sizeof(double) * (N[a] + 1));
        double x = 0.0;
        for (int64_t i = 0; i <= N[a]; ++i) { lp[a][i] = x; x += nonuni ? 0.7 + 0.6 * ((i * 7 + a) % 5) / 4.0 : 1.0; }
        d.lprim[a] = lp[a];
        d.isbloch[a] = bloch;
        d.boundft[a] = a == bft_hh_axis ? 1 : 0;
        d.npml[0][a] = bloch ? 0 : 1;
        d.npml[1][a] = bloch ? 0 : 1;
        d.kbloch[a] = bloch ? 0.3 - 0.1 * a : 0.0;
    }
    d.omega_pml[0] = 0.7;
    d.pml_m = 4.0;
    d.pml_lnR = -16.0;
    d.order_cmpfirst = cmpfirst;
    const int64_t M = N[0] * N[1] * N[2], n = 3 * M;
    oc64 *eps = calloc(9 * M, sizeof(oc64)), *mu = calloc(9 * M, sizeof(oc64));
    for (int64_t p = 0; p < M; ++p)
        for (int c = 0; c < 3; ++c) {
            eps[p + M * (c + 3 * c)] = 2.0 + 0.5 * c + 0.1 * I;
            mu[p + M * (c + 3 * c)] = 1.0 + 0.1 * c;
        }
    if (full)
        for (int64_t p = 0; p < M; ++p) { eps[p + M * 1] = eps[p + M * 3] = 0.2; eps[p + M * 5] = eps[p + M * 7] = -0.1; }
    /* grid tables and stretched ∆l */
    int64_t tot = N[0] + N[1] + N[2];
    double *l = malloc(sizeof(double) * 2 * tot), *dl = malloc(sizeof(double) * 2 * tot);
    oc64 *sdl = malloc(sizeof(oc64) * 2 * tot), *st = malloc(sizeof(oc64) * (N[0] + 1));
    CHECK(og_grid_tables(&d, l, dl, sdl));
    CHECK(og_stretched_dl(&d, 1, 0, st));
    /* sources */
    oc64 *jarr = calloc(n, sizeof(oc64)), *je = calloc(n, sizeof(oc64)), *jm = calloc(n, sizeof(oc64));
    const double c[3] = {0.5 * lp[0][N[0]], 0.4 * lp[1][N[1]], 0.6 * lp[2][N[2]]}, p[3] = {0.3, -0.5, 0.8};
    CHECK(og_add_pointsrc(&d, 0, c, p, 1.0, 0.5, jarr));
    CHECK(og_add_planesrc(&d, 0, 2, c[2], p, 1.0, 0.0, jarr));
    CHECK(og_field_arr2vec(&d, jarr, je));
    int64_t ind[2];
    double wt[2], bnd[2] = {lp[0][0], lp[0][N[0]]};
    CHECK(og_distweights(c[0], 0, bnd, l, dl, N[0], bloch, ind, wt));
    /* sparse operators, both formulations */
    og_csc Ce, Cm, Pe, Pm, Mc, A, Ah;
    CHECK(og_curl(&d, 0, &Ce));
    CHECK(og_curl(&d, 1, &Cm));
    CHECK(og_paramop(&d, 0, eps, &Pe));
    CHECK(og_paramop(&d, 1, mu, &Pm));
    CHECK(og_Mcs(&d, 0, &Mc));
    CHECK(og_create_A(&d, 0.6, 0.01, eps, mu, &A));
    oc64 *b = calloc(n, sizeof(oc64)), *y = calloc(n, sizeof(oc64)), *x = calloc(n, sizeof(oc64));
    CHECK(og_create_b(&d, 0.6, 0.01, mu, je, jm, b));
    CHECK(og_spmv(n, &A, b, y));
    if (!full) {
        CHECK(og_create_A_hh(&d, 0.6, 0.01, eps, mu, &Ah));
        CHECK(og_create_b_hh(&d, 0.6, 0.01, eps, je, jm, b));
        og_csc_free(&Ah);
    }
    int64_t its = 0;
    double rr = 0;
    /* 1 = not converged within maxit: a valid outcome on these tiny grids */
    const int src = og_solve_sqmr(&d, &A, y, x, 1e-8, 200, 1, &its, &rr);
    if (src != 0 && src != 1) { fprintf(stderr, "og_solve_sqmr returned %d\n", src); return 1; }
    og_csc_free(&Ce); og_csc_free(&Cm); og_csc_free(&Pe); og_csc_free(&Pm); og_csc_free(&Mc); og_csc_free(&A);
    free(b); free(y); free(x); free(jarr); free(je); free(jm);
    free(l); free(dl); free(sdl); free(st); free(eps); free(mu);
    for (int a = 0; a < 3; ++a) free(lp[a]);
    return 0;
}

static int run_materials(void) {
    const int64_t N[3] = {7, 6, 5};
    double *lp[3];
    og_desc d;
    memset(&d, 0, sizeof d);
    for (int a = 0; a < 3; ++a) {
        lp[a] = malloc(sizeof(double) * (N[a] + 1));
        for (int64_t i = 0; i <= N[a]; ++i) lp[a][i] = (double)i;
        d.N[a] = N[a];
        d.lprim[a] = lp[a];
    }
    d.omega_pml[0] = 1.0;
    d.pml_m = 4.0;
    d.pml_lnR = -16.0;
    d.order_cmpfirst = 1;
    og_shape sh[4];
    memset(sh, 0, sizeof sh);
    sh[0].kind = 1; sh[0].c[0] = 3.5; sh[0].c[1] = 3; sh[0].c[2] = 2.5; sh[0].s[0] = 7; sh[0].s[1] = 6; sh[0].s[2] = 5;
    sh[1].kind = 0; sh[1].c[0] = 3.2; sh[1].c[1] = 2.9; sh[1].c[2] = 2.4; sh[1].r = 1.8;
    sh[2].kind = 2; sh[2].c[0] = 4.1; sh[2].c[1] = 3.3; sh[2].c[2] = 2.5; sh[2].r = 1.1; sh[2].s[2] = 3.0;
    sh[3].kind = 1; sh[3].c[0] = 1.5; sh[3].c[1] = 1.5; sh[3].c[2] = 1.5; sh[3].s[0] = 1.3; sh[3].s[1] = 2.1; sh[3].s[2] = 1.7;
    oc64 et[36], mt[36];
    memset(et, 0, sizeof et);
    memset(mt, 0, sizeof mt);
    for (int o = 0; o < 4; ++o)
        for (int c = 0; c < 3; ++c) { et[9 * o + 4 * c] = 1.0 + 3.0 * o + c; mt[9 * o + 4 * c] = 1.0; }
    const int64_t M = N[0] * N[1] * N[2];
    oc64 *ea = calloc(9 * M, sizeof(oc64)), *ma = calloc(9 * M, sizeof(oc64));
    CHECK(og_matparams(&d, 4, sh, et, mt, ea, ma));
    const double vmin[3] = {0, 0, 0}, vmax[3] = {1, 1, 1}, nn[3] = {0.6, 0.8, 0.0}, r0[3] = {0.5, 0.5, 0.5};
    volatile double f = og_volfrac(vmin, vmax, nn, r0);
    (void)f;
    oc64 out[9];
    og_kottke(et + 9, et, 0.3, nn, out);
    free(ea); free(ma);
    for (int a = 0; a < 3; ++a) free(lp[a]);
    return 0;
}

int main(void) {
    const int64_t A3[3] = {4, 3, 5}, B3[3] = {1, 4, 3}, C3[3] = {5, 5, 4};
    if (run_case(A3, 0, -1, 1, 0, 0)) return 1;   /* PML box, EE */
    if (run_case(A3, 0, -1, 1, 0, 1)) return 1;   /* full ε */
    if (run_case(B3, 1, -1, 0, 1, 0)) return 1;   /* Bloch with a complex phase, nonuniform, component-last */
    if (run_case(C3, 0, 1, 1, 1, 0)) return 1;    /* boundft = HH on y */
    if (run_materials()) return 1;
    printf("sanitize driver ok\n");
    return 0;
}
