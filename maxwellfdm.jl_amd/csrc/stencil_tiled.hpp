// Tiled 2.5D curl-curl stencil for gfx950 (diagonal ε, μ = 1).
//
// A block owns a 64 × TY column tile in x-y (one 64-lane wavefront per x row,
// so global loads of a row are contiguous) and marches over a chunk of KZ
// z-planes.  Per plane it stages the scaled iterate X = Rf ∘ x of plane k+1
// (tile + one-cell halo ring) into LDS, computes H = Cₑ X of plane k on the
// tile + (−1) halo into LDS, and forms y at plane k from H(k), its −x/−y
// neighbours (LDS) and H(k−1) (registers of the owning thread).  Each iterate
// value is read from HBM once per block plus the halo ring; H is never
// written to memory.  Included by op.hip (needs its helpers).

template <int TY>
struct TileCfg {
    static constexpr int TX = 64;
    static constexpr int NT = TX * TY;        // threads per block
    static constexpr int XW = TX + 2;         // staged row width (halo both sides)
    static constexpr int XH = TY + 2;
    static constexpr int HW = TX + 1;         // H rows: −1 halo on the low side
    static constexpr int HH_ = TY + 1;
};

// a staged X value at global (gi, gj, gk): ghosts via ldX, points beyond the
// one-cell ring are never used and read as zero
template <typename R, int MODE>
__device__ inline cx<R> stage_ld(const DGrid<R> &g, const OpData<R> &od, const cx<R> *__restrict__ x, int c, int gi,
                                 int gj, int gk) {
    if (gi < -1 || gi > g.N[0] || gj < -1 || gj > g.N[1] || gk < -1 || gk > g.N[2]) return czero<R>();
    return ldX<R, MODE>(g, od, x, c, gi, gj, gk);
}

template <typename R, int MODE, int EPI, int TY>
__global__ __launch_bounds__(64 * TY) void k_apply_tiled(DGrid<R> g, OpData<R> od, const cx<R> *__restrict__ xall,
                                                         cx<R> *__restrict__ yall, const c64 *__restrict__ ball,
                                                         int64_t vs, int KZ, int nch, double *part, unsigned *cnt,
                                                         Scal *scal, double *rn) {
    using C = TileCfg<TY>;
    __shared__ cx<R> Xs[2][3][C::XH][C::XW];
    __shared__ cx<R> Hs[3][C::HH_][C::HW];
    const int rhs = blockIdx.z / nch, chunk = blockIdx.z - rhs * nch;
    const cx<R> *x = xall + rhs * vs;
    cx<R> *y = yall + rhs * vs;
    bool skip = false;
    if (EPI == EPI_DOT) {
        if (scal[rhs].done) return;
        skip = scal[rhs].conv != 0;
    }
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    const int i0 = blockIdx.x * C::TX, j0 = blockIdx.y * TY;
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < g.N[0] && j < g.N[1];
    const int k0 = chunk * KZ, k1 = min(k0 + KZ, g.N[2]);
    const int nring = 2 * C::XW + 2 * TY;
    double acc[2] = {0.0, 0.0};

    // stage plane gk into buffer b: interior by its owner (raw x, q kept), ring by the first threads
    auto stage = [&](int b, int gk, cx<R> (&xr)[3], cx<R> (&qr)[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cx<R> v = czero<R>();
            if (own && gk >= 0 && gk < g.N[2]) {
                int64_t p = i + (int64_t)g.N[0] * (j + (int64_t)g.N[1] * gk);
                cx<R> xv = x[dofidx(g, c, p)];
                xr[c] = xv;
                if (MODE == MODE_P) {
                    cx<R> qv = od.q[3 * p + c];
                    qr[c] = qv;
                    v = xv * (qv * tfac(g, c, i, j, gk, true));
                } else v = xv;
            } else if (own) {
                v = stage_ld<R, MODE>(g, od, x, c, i, j, gk);
            }
            Xs[b][c][ty + 1][tx + 1] = v;
        }
        for (int t = tid; t < nring; t += C::NT) {
            int ii, jj;
            if (t < C::XW) { ii = t; jj = 0; }
            else if (t < 2 * C::XW) { ii = t - C::XW; jj = C::XH - 1; }
            else { int u = t - 2 * C::XW; ii = (u & 1) ? C::XW - 1 : 0; jj = 1 + (u >> 1); }
#pragma unroll
            for (int c = 0; c < 3; ++c) Xs[b][c][jj][ii] = stage_ld<R, MODE>(g, od, x, c, i0 - 1 + ii, j0 - 1 + jj, gk);
        }
        // interior points of the tile that fall outside the grid
        if (!own) {
#pragma unroll
            for (int c = 0; c < 3; ++c) Xs[b][c][ty + 1][tx + 1] = stage_ld<R, MODE>(g, od, x, c, i, j, gk);
        }
    };

    // H_c at LDS position (hi, hj) of plane gk from buffers bk (plane gk) and bk1 (gk+1)
    auto hval = [&](int bk, int bk1, int hi, int hj, int gk, cx<R> (&h)[3]) {
        int ri = i0 - 1 + hi, rj = j0 - 1 + hj;
        if (ri >= g.N[0] || rj >= g.N[1] || (ri < 0 && !g.bloch[0]) || (rj < 0 && !g.bloch[1]) ||
            (gk < 0 && !g.bloch[2])) {
            h[0] = h[1] = h[2] = czero<R>();
            return;
        }
        int ix = ri < 0 ? g.N[0] - 1 : ri, iy = rj < 0 ? g.N[1] - 1 : rj, iz = gk < 0 ? g.N[2] - 1 : gk;
        cx<R> ax = g.sdlinv[DUAL][0][ix], ay = g.sdlinv[DUAL][1][iy], az = g.sdlinv[DUAL][2][iz];
        cx<R> X0[3], Xpx[3], Xpy[3], Xpz[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            X0[c] = Xs[bk][c][hj][hi];
            Xpx[c] = Xs[bk][c][hj][hi + 1];
            Xpy[c] = Xs[bk][c][hj + 1][hi];
            Xpz[c] = Xs[bk1][c][hj][hi];
        }
        h[0] = ay * (Xpy[2] - X0[2]) - az * (Xpz[1] - X0[1]);
        h[1] = az * (Xpz[0] - X0[0]) - ax * (Xpx[2] - X0[2]);
        h[2] = ax * (Xpx[1] - X0[1]) - ay * (Xpy[0] - X0[0]);
    };

    cx<R> xc[3], qc[3], xn[3], qn[3];   // raw x and q of the owner at planes k and k+1
    cx<R> hpx = czero<R>(), hpy = czero<R>();
    if (!skip && k0 < k1) {
        // prologue: planes k0−1 and k0, H_x/H_y of plane k0−1 at the owned point
        stage(0, k0 - 1, xn, qn);
        stage(1, k0, xc, qc);
        __syncthreads();
        cx<R> h[3];
        hval(0, 1, tx + 1, ty + 1, k0 - 1, h);
        hpx = h[0];
        hpy = h[1];
        for (int k = k0; k < k1; ++k) {
            const int bk = (k - k0 + 1) & 1, bn = bk ^ 1;
            __syncthreads();                       // everyone is done with buffer bn and Hs
            stage(bn, k + 1, xn, qn);
            __syncthreads();
            for (int t = tid; t < C::HH_ * C::HW; t += C::NT) {
                int hj = t / C::HW, hi = t - hj * C::HW;
                cx<R> hh[3];
                hval(bk, bn, hi, hj, k, hh);
#pragma unroll
                for (int c = 0; c < 3; ++c) Hs[c][hj][hi] = hh[c];
            }
            __syncthreads();
            if (own) {
                cx<R> H0[3], Hxm[3], Hym[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    H0[c] = Hs[c][ty + 1][tx + 1];
                    Hxm[c] = Hs[c][ty + 1][tx];
                    Hym[c] = Hs[c][ty][tx + 1];
                }
                cx<R> ex = g.sdlinv[PRIM][0][i], ey = g.sdlinv[PRIM][1][j], ez = g.sdlinv[PRIM][2][k];
                cx<R> cc[3];
                cc[0] = ey * (H0[2] - Hym[2]) - ez * (H0[1] - hpy);
                cc[1] = ez * (H0[0] - hpx) - ex * (H0[2] - Hxm[2]);
                cc[2] = ex * (H0[1] - Hxm[1]) - ey * (H0[0] - Hym[0]);
                hpx = H0[0];
                hpy = H0[1];
                int64_t p = i + (int64_t)g.N[0] * (j + (int64_t)g.N[1] * k);
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    cx<R> out;
                    if (MODE == MODE_A) {
                        out = cc[u] - od.om2 * (od.epsd[3 * p + u] * xc[u]);
                    } else {
                        cx<R> lf = qc[u] * tfac(g, u, i, j, k, false);
                        cx<R> cd = ccdiag<R, false>(g, od, u, i, j, k);
                        out = lf * cc[u] + (cone<R>() - cd * (qc[u] * qc[u])) * xc[u];
                    }
                    int64_t e = dofidx(g, u, p);
                    if (EPI == EPI_RESID) {
                        c64 bb = ball[rhs * vs + e];
                        cx<R> rr = mk<R>((R)bb.re - out.re, (R)bb.im - out.im);
                        y[e] = rr;
                        acc[0] += (double)rr.re * rr.re + (double)rr.im * rr.im;
                    } else {
                        y[e] = out;
                        if (EPI == EPI_DOT) {
                            acc[0] += (double)xc[u].re * out.re - (double)xc[u].im * out.im;
                            acc[1] += (double)xc[u].re * out.im + (double)xc[u].im * out.re;
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) { xc[c] = xn[c]; qc[c] = qn[c]; }
        }
    }
    if (EPI == EPI_STORE) return;
    const int nb = gridDim.x * gridDim.y * nch;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * chunk);
    double tot[2];
    if (!reduce_last_n<2>(acc, part + (size_t)rhs * NPART * 4, cnt + rhs, tot, bid, nb)) return;
    if (threadIdx.x != 0) return;
    if (EPI == EPI_RESID) { rn[rhs] = tot[0]; return; }
    Scal &S = scal[rhs];
    if (skip) { S.done = 1; return; }
    S.sig_re = tot[0];
    S.sig_im = tot[1];
    if (tot[0] == 0.0 && tot[1] == 0.0) { S.breakdown = 1; S.conv = 1; S.alpha_re = S.alpha_im = 0; return; }
    c64 a = cdiv(mk<double>(S.rho_re, S.rho_im), mk<double>(tot[0], tot[1]));
    S.alpha_re = a.re;
    S.alpha_im = a.im;
}
