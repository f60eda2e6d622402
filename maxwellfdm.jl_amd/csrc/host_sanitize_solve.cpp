// Drives libmfdfd's operator and solver entry points on a GPU in a build whose
// host code carries AddressSanitizer and UndefinedBehaviorSanitizer (device
// code unchanged; `make sanitize-host-solve`, on a machine with a GPU): graph-
// captured SQMR in FP64 and mixed precision, BiCGSTAB, COCG, an operator on the
// reflected grid (boundft = HH axes), the HH formulation and a component-
// windowed 2D operator.  Each solve is checked against the residual of the
// operator's own apply.  Prints "host sanitize solve ok" when every case passed.
#include "../../include/mfdfd.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

struct Case {
    const char *name;
    int64_t N[3];
    int bloch[3], bft[3], cmpfirst, ft;
    int c0, nc;          // component window (nc = 3: none)
    int precision, method;
};

static double norm(const std::vector<mfd_c64> &v) {
    double s = 0;
    for (const mfd_c64 &z : v) s += z.re * z.re + z.im * z.im;
    return std::sqrt(s);
}

static int run(mfd_ctx *ctx, const Case &k) {
    std::vector<double> lp[3];
    mfd_grid g;
    std::memset(&g, 0, sizeof g);
    for (int a = 0; a < 3; ++a) {
        for (int64_t i = 0; i <= k.N[a]; ++i) lp[a].push_back((double)i + 0.1 * ((i * 3 + a) % 4));
        g.N[a] = k.N[a];
        g.lprim[a] = lp[a].data();
        g.isbloch[a] = k.bloch[a];
        g.boundft[a] = k.bft[a];
        g.npml[0][a] = g.npml[1][a] = k.bloch[a] ? 0 : 2;
    }
    g.omega_pml.re = 0.6;
    g.pml_m = 4.0;
    g.pml_lnR = -16.0;
    g.order_cmpfirst = k.cmpfirst;
    const int64_t M = k.N[0] * k.N[1] * k.N[2];
    // 3×3 tensors per cell, entry (r, c) at p + M·(r + 3c); diagonal only
    std::vector<mfd_c64> eps(9 * M), mu(9 * M);
    for (int64_t p = 0; p < M; ++p)
        for (int c = 0; c < 3; ++c) {
            eps[p + M * (c + 3 * c)] = {1.0 + 3.0 * ((p * 7 + c) % 5) / 4.0, 0.02};
            mu[p + M * (c + 3 * c)] = {1.0, 0.0};
        }
    mfd_op *op = nullptr;
    const mfd_c64 om = {0.55, 0.0};
    if (mfd_op_create(ctx, &g, k.ft, om, eps.data(), mu.data(), &op) != MFD_OK) {
        std::fprintf(stderr, "%s: create: %s\n", k.name, mfd_last_error());
        return 1;
    }
    int rc = 0;
    if (k.nc < 3 && mfd_op_set_components(op, k.c0, k.nc) != MFD_OK) {
        std::fprintf(stderr, "%s: window: %s\n", k.name, mfd_last_error());
        rc = 1;
    }
    const int64_t n = (int64_t)k.nc * M;
    std::vector<mfd_c64> b(n), e(n), y(n);
    for (int64_t i = 0; i < n && !rc; ++i) b[i] = {std::cos(0.37 * i), std::sin(0.11 * i)};
    mfd_solve_opts o;
    mfd_solve_opts_default(&o);
    o.tol = 1e-9;
    o.precision = k.precision;
    o.method = k.method;
    mfd_solve_stats st;
    if (!rc && mfd_solve(op, b.data(), e.data(), 1, &o, &st) != MFD_OK) {
        std::fprintf(stderr, "%s: solve: %s\n", k.name, mfd_last_error());
        rc = 1;
    }
    if (!rc && mfd_op_apply(op, e.data(), y.data(), 1) != MFD_OK) {
        std::fprintf(stderr, "%s: apply: %s\n", k.name, mfd_last_error());
        rc = 1;
    }
    if (!rc) {
        for (int64_t i = 0; i < n; ++i) y[i] = {b[i].re - y[i].re, b[i].im - y[i].im};
        const double rel = norm(y) / norm(b);
        std::printf("%-22s %6lld iterations, residual %.2e\n", k.name, (long long)st.iterations, rel);
        if (!(st.converged && rel <= 2e-9)) {
            std::fprintf(stderr, "%s: residual %.3e\n", k.name, rel);
            rc = 1;
        }
    }
    mfd_op_destroy(op);
    return rc;
}

int main() {
    mfd_ctx *ctx = nullptr;
    if (mfd_ctx_create(0, &ctx) != MFD_OK) {
        std::fprintf(stderr, "no GPU context: %s\n", mfd_last_error());
        return 1;
    }
    const Case cases[] = {
        {"ee sqmr fp64", {12, 10, 9}, {0, 0, 0}, {0, 0, 0}, 1, MFD_EE, 0, 3, MFD_FP64, -1},
        {"ee sqmr mixed cmplast", {12, 10, 9}, {0, 1, 0}, {0, 0, 0}, 0, MFD_EE, 0, 3, MFD_MIXED, -1},
        {"ee cocg", {12, 10, 9}, {0, 0, 0}, {0, 0, 0}, 1, MFD_EE, 0, 3, MFD_FP64, MFD_COCG},
        {"ee bicgstab", {12, 10, 9}, {0, 0, 0}, {0, 0, 0}, 1, MFD_EE, 0, 3, MFD_FP64, MFD_BICGSTAB},
        {"boundft hh x, z", {12, 10, 9}, {0, 0, 0}, {1, 0, 1}, 1, MFD_EE, 0, 3, MFD_FP64, -1},
        {"hh formulation", {12, 10, 9}, {0, 0, 0}, {0, 0, 0}, 1, MFD_HH, 0, 3, MFD_FP64, -1},
        {"window te 2d", {20, 17, 1}, {0, 0, 1}, {0, 0, 0}, 0, MFD_EE, 0, 2, MFD_FP64, -1},
        {"window tm 2d mixed", {20, 17, 1}, {0, 0, 1}, {0, 0, 0}, 0, MFD_EE, 2, 1, MFD_MIXED, -1},
    };
    int bad = 0;
    for (const Case &k : cases) bad += run(ctx, k);
    mfd_ctx_destroy(ctx);
    if (bad) return 1;
    std::printf("host sanitize solve ok (%s)\n", mfd_version());
    return 0;
}
