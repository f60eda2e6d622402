// Drives libmfdfd's host-side entry points (grid tables, distweights, point
// and plane source deposits, error reporting, context creation without a GPU)
// in a build whose host code carries AddressSanitizer and
// UndefinedBehaviorSanitizer (`make sanitize-host`).  Needs no GPU.  Prints
// "host sanitize driver ok" when every call behaved as documented.
#include "../../include/mfdfd.h"
#include <cstdio>
#include <cstring>
#include <vector>

static int fail(const char *what) {
    std::fprintf(stderr, "host sanitize driver: %s (%s)\n", what, mfd_last_error());
    return 1;
}

static int grid_case(const int64_t N[3], int bloch, int nonuni, int cmpfirst) {
    std::vector<double> lp[3];
    mfd_grid g;
    std::memset(&g, 0, sizeof g);
    for (int a = 0; a < 3; ++a) {
        double x = -0.5 * N[a];
        for (int64_t i = 0; i <= N[a]; ++i) {
            lp[a].push_back(x);
            x += nonuni ? 0.6 + 0.8 * ((i * 5 + a) % 7) / 6.0 : 1.0;
        }
        g.N[a] = N[a];
        g.lprim[a] = lp[a].data();
        g.isbloch[a] = bloch;
        g.npml[0][a] = g.npml[1][a] = bloch ? 0 : (N[a] > 4 ? 2 : 0);
        g.kbloch[a] = bloch ? 0.2 * a : 0.0;
    }
    g.omega_pml.re = 0.8;
    g.pml_m = 4.0;
    g.pml_lnR = -16.0;
    g.order_cmpfirst = cmpfirst;
    const int64_t M = N[0] * N[1] * N[2];
    std::vector<mfd_c64> j(3 * M);
    const double p[3] = {0.2, -0.7, 0.5};
    const mfd_c64 one = {1.0, 0.25};
    // sources at, between and near the bounds of every axis
    for (int s = 0; s < 5; ++s) {
        const double f = s / 4.0;
        double c[3];
        for (int a = 0; a < 3; ++a) c[a] = lp[a][0] + f * (lp[a][N[a]] - lp[a][0]);
        if (mfd_add_pointsrc(&g, MFD_EE, c, p, one, j.data()) != MFD_OK) return fail("pointsrc");
        if (mfd_add_pointsrc(&g, MFD_HH, c, p, one, j.data()) != MFD_OK) return fail("pointsrc HH");
        for (int nn = 0; nn < 3; ++nn)
            if (mfd_add_planesrc(&g, MFD_EE, nn, c[nn], p, one, j.data()) != MFD_OK) return fail("planesrc");
    }
    // outside the domain: an argument error, nothing written out of bounds
    double out[3] = {lp[0][N[0]] + 1.0, 0.0, 0.0};
    if (mfd_add_pointsrc(&g, MFD_EE, out, p, one, j.data()) != MFD_ERR_ARG) return fail("pointsrc outside accepted");
    for (int a = 0; a < 3; ++a) {
        std::vector<double> dl(N[a]);
        for (int64_t i = 0; i < N[a]; ++i) dl[i] = lp[a][i + 1] - lp[a][i];
        const double bounds[2] = {lp[a][0], lp[a][N[a]]};
        int64_t ind[2];
        double wt[2];
        for (int s = 0; s <= 8; ++s) {
            const double c = bounds[0] + s / 8.0 * (bounds[1] - bounds[0]);
            for (int gt = 0; gt < 2; ++gt)
                if (mfd_distweights(c, gt, bounds, lp[a].data(), dl.data(), N[a], bloch, ind, wt) != MFD_OK)
                    return fail("distweights");
        }
    }
    return 0;
}

int main() {
    const int64_t A[3] = {6, 5, 7}, B[3] = {1, 3, 4}, C[3] = {9, 2, 3};   // a 1-cell axis needs Bloch (B)
    if (grid_case(A, 0, 0, 1) || grid_case(A, 1, 1, 0) || grid_case(B, 1, 0, 1) || grid_case(C, 0, 1, 1)) return 1;
    mfd_ctx *ctx = nullptr;
    const int rc = mfd_ctx_create(0, &ctx);   // without a GPU: an error code and a message
    if (rc == MFD_OK) mfd_ctx_destroy(ctx);
    else if (std::strlen(mfd_last_error()) == 0) return fail("no message for a failed context");
    // operator entry points refuse a null handle before touching the device
    if (mfd_op_set_components(nullptr, 0, 2) != MFD_ERR_ARG) return fail("set_components(null)");
    if (mfd_op_set_tile(nullptr, MFD_FP64, 4, 16, 1) != MFD_ERR_ARG) return fail("set_tile(null)");
    if (mfd_op_set_graddiv(nullptr, 1) != MFD_ERR_ARG) return fail("set_graddiv(null)");
    std::printf("host sanitize driver ok (%s)\n", mfd_version());
    return 0;
}
