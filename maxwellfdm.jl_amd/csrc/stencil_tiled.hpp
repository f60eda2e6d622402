// Tiled 2.5D curl-curl stencil for gfx950 (diagonal ε, μ = 1).
//
// Vector storage precision R (complex FP64 or FP32 iterates in HBM); the
// operator data (ε, Jacobi scaling, grad-div weights), the 1D metric tables,
// the LDS staging and all arithmetic are FP64, so the FP32 solver iterates
// against exactly the operator the FP64 residual uses — only the vectors are
// rounded, once, at the store.  The kernel is HBM-bound: FP64 math is free.
//
// A block owns a TX × TY column tile in x-y (TX = 64: one 64-lane wavefront
// per x row, so a row's loads are contiguous; TX = 32: two rows per
// wavefront, for grids whose x size is not a multiple of 64, where 64-wide
// tiles would leave lanes of the last tile idle) and marches over a chunk of
// KZ z-planes.
// Per plane it stages the scaled iterate X = Rf ∘ x of plane k+1 (tile + a
// one-cell halo ring) in LDS, computes H = Cₑ X of plane k on the tile plus
// its −x/−y halo in LDS, and forms y at plane k from H(k), its −x/−y
// neighbours and H(k−1) (registers of the owning thread).  Every 1D metric
// factor a thread needs (x from its column, y from its row) is loaded once,
// into registers (or, in the full-ε form, into block tables); the z factors
// of a plane are block-uniform.  H is never written to memory.
//
// GD adds the grad-div term of the augmented inner operator
//     A' = Cₘ Cₑ − ω² ε − ε G W Div ε,
// staging Y = ε ∘ X next to X and keeping ψ = W ∘ Div(Y) for planes k and k+1
// in LDS (G: forward differences to E edges, Div: backward differences to the
// primal nodes, the same metric as the curls).  Included by op.hip.

template <int TY, int TXW = 64>
struct TileCfg {
    static constexpr int TX = TXW;       // tile width: 64, or 32 (two rows per wavefront)
    static constexpr int NT = TX * TY;   // threads per block
    static constexpr int XW = TX + 2;    // staged row width (halo both sides)
    static constexpr int XH = TY + 2;
    static constexpr int HW = TX + 1;    // H rows: −1 halo on the low side
    static constexpr int HHt = TY + 1;
    static constexpr int PW = TX + 1;    // ψ rows: +1 halo on the high side
    static constexpr int PHt = TY + 1;
};

// wrap a tile coordinate along axis a: returns false when the point is a
// non-Bloch ghost (value zero) or lies beyond the one-cell ring
template <typename R>
__device__ inline bool wrap_ax(const DGrid<R> &g, int a, int v, int &w, cx<R> &ph) {
    int N = g.N[a];
    if (v < -1 || v > N) return false;
    if (v == -1) {
        if (!g.bloch[a]) return false;
        w = N - 1;
        ph = ph * g.phinv[a];
    } else if (v == N) {
        if (!g.bloch[a]) return false;
        w = 0;
        ph = ph * g.ph[a];
    } else w = v;
    return true;
}

// 1D term of the curl-curl diagonal along axis a at index ia (μ = 1)
template <typename R>
__device__ inline cx<R> axis_term(const DGrid<R> &g, int a, int ia) {
    int Na = g.N[a];
    cx<R> ie = g.sdlinv[PRIM][a][ia];
    if (Na == 1 && g.bloch[a]) {
        cx<R> f = mk<R>((R)2, (R)0) - g.ph[a] - g.phinv[a];
        return (ie * g.sdlinv[DUAL][a][0]) * f;
    }
    if (ia == 0 && !g.bloch[a]) return ie * g.sdlinv[DUAL][a][0];
    int im = ia > 0 ? ia - 1 : Na - 1;
    return ie * (g.sdlinv[DUAL][a][ia] + g.sdlinv[DUAL][a][im]);
}

// Read-only table entry at a block-uniform index, loaded through the constant
// address space as one 16-byte value so that it compiles to a scalar load: its
// wait (lgkmcnt) is independent of the vector-load counter the prefetch uses.
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ inline c64 ldu(const c64 *tab, int idx) {
    const int u = __builtin_amdgcn_readfirstlane(idx);
    const d2v v = *(const __attribute__((address_space(4))) d2v *)(tab + u);
    return mk<double>(v.x, v.y);
}

// axis_term along z at a block-uniform plane index, from scalar loads
__device__ inline c64 axis_term_u(const DGrid<double> &g, int ia) {
    const int Na = g.N[2];
    c64 ie = ldu(g.sdlinv[PRIM][2], ia);
    if (Na == 1 && g.bloch[2]) {
        c64 f = mk<double>(2.0, 0.0) - g.ph[2] - g.phinv[2];
        return (ie * ldu(g.sdlinv[DUAL][2], 0)) * f;
    }
    if (ia == 0 && !g.bloch[2]) return ie * ldu(g.sdlinv[DUAL][2], 0);
    int im = ia > 0 ? ia - 1 : Na - 1;
    return ie * (ldu(g.sdlinv[DUAL][2], ia) + ldu(g.sdlinv[DUAL][2], im));
}

// per-thread cached description of one staged point (fixed x-y position)
template <typename R>
struct StagePt {
    int gi, gj;        // wrapped indices
    bool ok;           // not a zero ghost / out of ring
    cx<R> ph;          // x-y wrap phase
};

template <typename R>
__device__ inline void stage_pt_init(const DGrid<R> &g, int vi, int vj, StagePt<R> &s) {
    s.ph = cone<R>();
    s.ok = wrap_ax(g, 0, vi, s.gi, s.ph) && wrap_ax(g, 1, vj, s.gj, s.ph);
    if (!s.ok) { s.gi = 0; s.gj = 0; }
}

// per-plane (block-uniform) z data
template <typename R>
struct PlaneZ {
    int gk;
    bool ok;
    cx<R> ph;
};

template <typename R>
__device__ inline void plane_init(const DGrid<R> &g, int vk, PlaneZ<R> &z) {
    z.ph = cone<R>();
    z.ok = wrap_ax(g, 2, vk, z.gk, z.ph);
    if (!z.ok) z.gk = 0;
}

// FUSE (COCG step): the operand is p = r + β p_old, formed while staging from
// r (rall) and p_old (xall); the owner of each point also stores p (pnall) and
// applies the previous step's deferred iterate update it += α_prev p_old
// (itall), with β and α_prev taken from the solver scalars.
// Launch bound: the full-ε form is held to two waves per SIMD (256 VGPRs,
// a few spilled) — at TY = 4 that doubles its blocks per CU and beats the
// unbounded allocation (one wave per SIMD with AGPR overflow) by 24%.
// OFFD: the ε tensor has corner off-diagonals (xy, xz, yz); their term is
// formed from the staged planes k−1 (three X_z values kept in registers),
// k and k+1, with the corner ε read from global memory.
// REV: the grid g is the reflection of the operator's grid along the axes of
// g.rmask (boundft HH there, EE in g) and the kernel works in its
// coordinates: every per-cell global access (vectors, qs, ε, 1/μ) goes to the
// caller's cell (index N_a−1−i along a reflected axis), so the vectors keep
// their order, and the E components along reflected axes are negated on
// load and on store.
// WIN: the vectors hold only the components c0 .. c0+nc−1 of g (the 2D and 1D
// models); the others read as zero and are not written.
template <typename R, int MODE, int EPI, int TY, bool GD, bool FUSE = false, bool OFFD = false, int TXW = 64,
          bool MU = false, bool REV = false, bool WIN = false>
__global__ __launch_bounds__(TXW * TY, OFFD ? 2 : 1) void k_apply_tiled(DGrid<double> g, OpData<double> od, c64 om2,
                                                         const cx<R> *__restrict__ xall,
                                                         cx<R> *__restrict__ yall, const c64 *__restrict__ ball,
                                                         int64_t vs, int KZ, int nch, double *part,
                                                         const Scal *scal, const int *__restrict__ mask, int xcd,
                                                         const cx<R> *__restrict__ rall = nullptr,
                                                         cx<R> *__restrict__ pnall = nullptr,
                                                         cx<R> *__restrict__ itall = nullptr) {
    using C = TileCfg<TY, TXW>;
    static_assert(!(REV && (GD || OFFD || FUSE)), "the reflected form serves diagonal ε, unfused, without grad-div");
    static_assert(!(WIN && (GD || FUSE || MU || REV)), "the windowed form serves μ = 1, boundft EE, unfused");
    constexpr int GB = GD ? 2 : 1, GH = GD ? C::XH : 1, GW = GD ? C::XW : 1;
    __shared__ c64 Xs[2][3][C::XH][C::XW];
    __shared__ c64 Ys[GB][3][GH][GW];
    __shared__ c64 Hs[3][C::HHt][C::HW];
    __shared__ c64 Ps[GB][GD ? C::PHt : 1][GD ? C::PW : 1];
    // OFFD: the corner-average weights of the tile's columns (x) and rows (y),
    // w0 = ∆l_dual(i)/(2 ∆l_prim(i)) and w1 = ∆l_dual(i−1)/(2 ∆l_prim(i))
    __shared__ c64 Wx[2][OFFD ? C::TX : 1];
    __shared__ c64 Wy[2][OFFD ? TY : 1];
    // OFFD: the owner's per-column and per-row 1D factors (1/∆l_p, the axis
    // term, ∆l_d, ∆l_p) and the H points' 1/∆l_d, read from block tables every
    // plane instead of held in registers (the full-ε stencil is the one short
    // of registers)
    __shared__ c64 Fx[OFFD ? 4 : 1][OFFD ? C::TX : 1];
    __shared__ c64 Fy[OFFD ? 4 : 1][OFFD ? TY : 1];
    __shared__ c64 Hx[OFFD ? C::HW : 1];
    __shared__ c64 Hy[OFFD ? C::HHt : 1];
    // Tile of this block.  Workgroups are dispatched to the 8 XCDs round-robin
    // (dispatch index L goes to XCD L mod 8); with xcd != 0 each XCD gets one
    // contiguous run of tiles in x-then-y order, so the y halo rows a tile
    // stages are read by its neighbour on the same XCD and hit that XCD's L2.
    // The right-hand side is the fastest-varying index of that run: the blocks
    // of one tile for every RHS are neighbours in it, so they run together on
    // one XCD and all but the first find the tile's operator data (qs, corner
    // ε) and its halo rows in that XCD's L2.
    const unsigned nbx = gridDim.x, nby = gridDim.y;
    const unsigned T = nbx * nby * gridDim.z;
    unsigned L = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z), lt = L;
    if (xcd) {
        const unsigned q = T / 8, rem = T % 8, x = L % 8;
        lt = x * q + (x < rem ? x : rem) + L / 8;
    }
    const unsigned nrhs_l = gridDim.z / (unsigned)nch, tile = lt / nrhs_l;
    const int rhs = (int)(lt - tile * nrhs_l);
    const int bx = (int)(tile % nbx), by = (int)((tile / nbx) % nby), chunk = (int)(tile / (nbx * nby));
    const cx<R> *x = xall + rhs * vs;
    cx<R> *y = yall + rhs * vs;
    const cx<R> *rr = FUSE ? rall + rhs * vs : nullptr;
    cx<R> *pn = FUSE ? pnall + rhs * vs : nullptr;
    cx<R> *itv = FUSE ? itall + rhs * vs : nullptr;
    c64 fbeta = czero<double>(), fapend = czero<double>();
    if (FUSE) {
        fbeta = mk<double>(scal[rhs].beta_re, scal[rhs].beta_im);
        fapend = mk<double>(scal[rhs].apend_re, scal[rhs].apend_im);
    }
    bool skip = false;
    if (EPI == EPI_RESID && mask && !mask[rhs]) return;
    if (EPI == EPI_DOT) {
        if (scal[rhs].done) return;
        skip = scal[rhs].conv != 0;
    }
    const int tid = threadIdx.x, tx = tid % C::TX, ty = tid / C::TX;
    const int i0 = bx * C::TX, j0 = by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < g.N[0] && j < g.N[1];
    const int k0 = chunk * KZ, k1 = min(k0 + KZ, g.N[2]);
    const int64_t NxNy = (int64_t)g.N[0] * g.N[1];
    double acc[2] = {0.0, 0.0};
    // the caller's cell of cell (ci, cj, ck) of g: x-y part and z part
    auto gxy = [&](int ci, int cj) -> int64_t {
        if (REV && (g.rmask & 1)) ci = g.N[0] - 1 - ci;
        if (REV && (g.rmask & 2)) cj = g.N[1] - 1 - cj;
        return ci + (int64_t)g.N[0] * cj;
    };
    auto gz = [&](int ck) -> int64_t { return NxNy * (REV && (g.rmask & 4) ? g.N[2] - 1 - ck : ck); };
    // REV: E_c is negated along a reflected axis c
    auto rsg = [&](int c, c64 v) -> c64 { return REV && ((g.rmask >> c) & 1) ? -v : v; };
    // component window of the vectors (the others read as zero, are not written)
    auto inw = [&](int c) -> bool { return !WIN || (unsigned)(c - g.c0) < (unsigned)g.nc; };

    // staged points of this thread: its interior point (slot 0) and its share
    // of the halo ring (slots 1..NRS; the ring has 2·XW + 2·TY points)
    constexpr int nring = 2 * C::XW + 2 * TY;
    constexpr int NRS = (nring + C::NT - 1) / C::NT;
    StagePt<double> sp[1 + NRS];
    int sjj[1 + NRS], sii[1 + NRS];
    bool sact[1 + NRS];
    sii[0] = tx + 1;
    sjj[0] = ty + 1;
    sact[0] = true;
    stage_pt_init(g, i, j, sp[0]);
#pragma unroll
    for (int q = 0; q < NRS; ++q) {
        int t = tid + q * C::NT, ii, jj;
        sact[1 + q] = t < nring;
        if (!sact[1 + q]) t = 0;
        if (t < C::XW) { ii = t; jj = 0; }
        else if (t < 2 * C::XW) { ii = t - C::XW; jj = C::XH - 1; }
        else { int u = t - 2 * C::XW; ii = (u & 1) ? C::XW - 1 : 0; jj = 1 + (u >> 1); }
        sii[1 + q] = ii;
        sjj[1 + q] = jj;
        stage_pt_init(g, i0 - 1 + ii, j0 - 1 + jj, sp[1 + q]);
    }
    // owner's output factors (x from the column, y from the row)
    const int io = own ? i : 0, jo = own ? j : 0;
    c64 exo = cone<double>(), eyo = cone<double>(), tmx = czero<double>(), tmy = czero<double>();
    c64 gdx = czero<double>(), gdy = czero<double>();
    // x-y part of T² (T = S_E^{1/2}: the product of the stretched ∆l)
    c64 T2xy[3] = {cone<double>(), cone<double>(), cone<double>()};
    if (!OFFD) {
        exo = g.sdlinv[PRIM][0][io];
        eyo = g.sdlinv[PRIM][1][jo];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            T2xy[c] = g.sdl[c == 0 ? DUAL : PRIM][0][io] * g.sdl[c == 1 ? DUAL : PRIM][1][jo];
        tmx = axis_term(g, 0, io);
        tmy = axis_term(g, 1, jo);
        gdx = g.sdlinv[DUAL][0][io];
        gdy = g.sdlinv[DUAL][1][jo];
    }
    if (OFFD) {   // published by the prologue's barrier
        if (tid < C::TX) {
            const int ii = min(i0 + tid, g.N[0] - 1);
            Fx[0][tid] = g.sdlinv[PRIM][0][ii];
            Fx[1][tid] = axis_term(g, 0, ii);
            Fx[2][tid] = g.sdl[DUAL][0][ii];
            Fx[3][tid] = g.sdl[PRIM][0][ii];
        }
        if (tid < TY) {
            const int jj = min(j0 + tid, g.N[1] - 1);
            Fy[0][tid] = g.sdlinv[PRIM][1][jj];
            Fy[1][tid] = axis_term(g, 1, jj);
            Fy[2][tid] = g.sdl[DUAL][1][jj];
            Fy[3][tid] = g.sdl[PRIM][1][jj];
        }
        for (int t = tid; t < C::HW; t += C::NT) {   // HW = TX + 1 exceeds the 64 threads of TY = 1
            const int ri = i0 - 1 + t;
            Hx[t] = g.sdlinv[DUAL][0][ri < 0 ? g.N[0] - 1 : (ri < g.N[0] ? ri : 0)];
        }
        if (tid < C::HHt) {
            const int rj = j0 - 1 + tid;
            Hy[tid] = g.sdlinv[DUAL][1][rj < 0 ? g.N[1] - 1 : (rj < g.N[1] ? rj : 0)];
        }
        const c64 h5 = mk<double>(0.5, 0);
        if (tid < C::TX) {
            const int ii = min(i0 + tid, g.N[0] - 1), im = ii > 0 ? ii - 1 : g.N[0] - 1;
            const c64 ix = g.sdlinv[PRIM][0][ii];
            Wx[0][tid] = (h5 * g.sdl[DUAL][0][ii]) * ix;
            Wx[1][tid] = (h5 * g.sdl[DUAL][0][im]) * ix;
        }
        if (tid < TY) {
            const int jj = min(j0 + tid, g.N[1] - 1), jm = jj > 0 ? jj - 1 : g.N[1] - 1;
            const c64 iy = g.sdlinv[PRIM][1][jj];
            Wy[0][tid] = (h5 * g.sdl[DUAL][1][jj]) * iy;
            Wy[1][tid] = (h5 * g.sdl[DUAL][1][jm]) * iy;
        }
    }

    // H points of this thread (fixed across planes): LDS (hi, hj) -> global (i0-1+hi, j0-1+hj)
    constexpr int NHP = (C::HHt * C::HW + C::NT - 1) / C::NT;
    c64 hax[NHP], hay[NHP];
    bool hok[NHP];
    int64_t hxy[MU ? NHP : 1];   // MU: wrapped x-y cell of each H point
#pragma unroll
    for (int r = 0; r < NHP; ++r) {
        hok[r] = false;
        hax[r] = hay[r] = czero<double>();
        if (OFFD) continue;   // from the tables and the position, in hplane
        int t = tid + r * C::NT;
        int hj = t / C::HW, hi = t - hj * C::HW;
        int ri = i0 - 1 + hi, rj = j0 - 1 + hj;
        hok[r] = t < C::HHt * C::HW && ri < g.N[0] && rj < g.N[1] && (ri >= 0 || g.bloch[0]) && (rj >= 0 || g.bloch[1]);
        int ix = ri < 0 ? g.N[0] - 1 : (ri < g.N[0] ? ri : 0), iy = rj < 0 ? g.N[1] - 1 : (rj < g.N[1] ? rj : 0);
        hax[r] = g.sdlinv[DUAL][0][ix];
        hay[r] = g.sdlinv[DUAL][1][iy];
        if (MU) hxy[r] = gxy(ix, iy);
    }
    // ψ nodes of this thread: LDS (pi, pj) -> global (i0+pi, j0+pj), pi in [0,64], pj in [0,TY]
    constexpr int NPP = GD ? (C::PHt * C::PW + C::NT - 1) / C::NT : 1;
    c64 pex[NPP], pey[NPP], pph[NPP];
    int pidx[NPP];
    bool pok[NPP];
    if (GD) {
#pragma unroll
        for (int r = 0; r < NPP; ++r) {
            int t = tid + r * C::NT;
            int pj = t / C::PW, pi = t - pj * C::PW;
            int gi = 0, gj = 0;
            c64 ph = cone<double>();
            pok[r] = t < C::PHt * C::PW && wrap_ax(g, 0, i0 + pi, gi, ph) && wrap_ax(g, 1, j0 + pj, gj, ph);
            if (!pok[r]) { gi = 0; gj = 0; }
            pidx[r] = gi + g.N[0] * gj;
            pph[r] = ph;
            pex[r] = g.sdlinv[PRIM][0][gi];
            pey[r] = g.sdlinv[PRIM][1][gj];
        }
    }

    // stage plane vk into buffer b; the owner keeps raw x, q and ε of its point
    auto stage = [&](int b, int vk, c64 (&xr)[3], c64 (&qr)[3], c64 (&er)[3]) {
        PlaneZ<double> z;
        plane_init(g, vk, z);
#pragma unroll
        for (int s = 0; s < 1 + NRS; ++s) {
            if (!sact[s]) continue;
            const StagePt<double> &P = sp[s];
            bool ok = P.ok && z.ok;
            const int64_t p = gxy(P.gi, P.gj) + gz(z.gk);
            c64 ph = P.ph * z.ph;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                c64 v = czero<double>(), yv = czero<double>();
                if (ok) {
                    const int64_t dof = dofidx(g, c, p);
                    c64 xv;
                    if (FUSE) {
                        c64 po = cvt<double>(x[dof]);
                        // rounded to the storage precision, so that the operator
                        // sees exactly the p that is stored
                        xv = cvt<double>(cvt<R>(cvt<double>(rr[dof]) + fbeta * po));
                        if (s == 0 && own && vk >= k0 && vk < k1) {   // this block's point
                            pn[dof] = cvt<R>(xv);
                            itv[dof] = cvt<R>(cvt<double>(itv[dof]) + fapend * po);
                        }
                    } else xv = inw(c) ? rsg(c, cvt<double>(x[dof])) : czero<double>();
                    c64 qv = cone<double>();
                    if (MODE == MODE_P) {
                        if (inw(c)) qv = od.qs[3 * p + c];   // q·T⁻¹ (outside the window x is zero)
                        v = xv * qv;
                    } else v = xv;
                    c64 ev = cone<double>();
                    if (GD || ((MODE == MODE_A || MU) && s == 0)) ev = cvt<double>(od.epsd[3 * p + c]);
                    if (GD) yv = ev * v;
                    if (s == 0) { xr[c] = xv; qr[c] = qv; er[c] = ev; }
                    v = v * ph;
                    yv = yv * ph;
                }
                Xs[b][c][sjj[s]][sii[s]] = v;
                if (GD) Ys[b][c][sjj[s]][sii[s]] = yv;
            }
        }
    };

    // Software pipeline (no grad-div, fused or off-diagonal term, whose extra
    // per-thread loads would wait on it): the global loads of plane k+2 are
    // issued before plane k's H and output and land in registers while those
    // run; commit() scales them into LDS one pass later.  Nothing between a
    // fetch and its commit issues a per-thread global load.
    constexpr bool PIPE = !GD && !FUSE && !OFFD;
    constexpr int NPF = PIPE ? 1 + NRS : 1;
    cx<R> pfx[NPF][3];
    c64 pfq[MODE == MODE_P ? NPF : 1][3];
    c64 pfe[3];
    auto fetch = [&](int vk) {
        PlaneZ<double> z;
        plane_init(g, vk, z);
#pragma unroll
        for (int s = 0; s < NPF; ++s) {
            if (!sact[s]) continue;
            const bool ok = sp[s].ok && z.ok;
            const int64_t p = gxy(sp[s].gi, sp[s].gj) + gz(z.gk);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pfx[s][c] = ok && inw(c) ? x[dofidx(g, c, p)] : mk<R>((R)0, (R)0);
                if (MODE == MODE_P) pfq[s][c] = ok ? od.qs[3 * p + c] : czero<double>();
                if ((MODE == MODE_A || MU) && s == 0) pfe[c] = ok ? od.epsd[3 * p + c] : cone<double>();
            }
        }
    };
    auto commit = [&](int b, int vk, c64 (&xr)[3], c64 (&qr)[3], c64 (&er)[3]) {
        PlaneZ<double> z;
        plane_init(g, vk, z);
#pragma unroll
        for (int s = 0; s < NPF; ++s) {
            if (!sact[s]) continue;
            const c64 ph = sp[s].ph * z.ph;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const c64 xv = rsg(c, cvt<double>(pfx[s][c]));
                c64 v = xv;
                if (MODE == MODE_P) v = xv * pfq[s][c];
                if (s == 0) {
                    xr[c] = xv;
                    qr[c] = MODE == MODE_P ? pfq[s][c] : cone<double>();
                    er[c] = (MODE == MODE_A || MU) ? pfe[c] : cone<double>();
                }
                Xs[b][c][sjj[s]][sii[s]] = v * ph;
            }
        }
    };

    // H of an interior plane (buffers bk: that plane, bn: the next) for all H
    // points; az is 1/s∆l_z of the plane
    // MU: 1/μ at this thread's H points of the plane being processed, loaded at
    // the top of the plane (ahead of the prefetch, so waiting for them does
    // not wait for it)
    c64 hmu[MU ? NHP : 1][3];
    auto load_mu = [&](int vk) {
        PlaneZ<double> z;
        plane_init(g, vk, z);
#pragma unroll
        for (int r = 0; r < (MU ? NHP : 0); ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                hmu[r][c] = hok[r] && z.ok ? od.muinv[3 * (hxy[r] + gz(z.gk)) + c] : czero<double>();
    };
    // slab: a one-cell z axis, whose plane k+1 is plane k times pzn (the
    // Bloch phase, or 0 past a non-Bloch end) and is not staged
    auto hplane = [&](int bk, int bn, c64 az, bool slab, c64 pzn) {
#pragma unroll
        for (int r = 0; r < NHP; ++r) {
            int t = tid + r * C::NT;
            if (t >= C::HHt * C::HW) break;
            int hj = t / C::HW, hi = t - hj * C::HW;
            c64 h0 = czero<double>(), h1 = czero<double>(), h2 = czero<double>();
            bool okh = hok[r];
            c64 ax = hax[r], ay = hay[r];
            if (OFFD) {
                const int ri = i0 - 1 + hi, rj = j0 - 1 + hj;
                okh = ri < g.N[0] && rj < g.N[1] && (ri >= 0 || g.bloch[0]) && (rj >= 0 || g.bloch[1]);
                ax = Hx[hi];
                ay = Hy[hj];
            }
            if (okh) {
                c64 X0x = Xs[bk][0][hj][hi], X0y = Xs[bk][1][hj][hi], X0z = Xs[bk][2][hj][hi];
                const c64 Xnx = slab ? pzn * X0x : Xs[bn][0][hj][hi];
                const c64 Xny = slab ? pzn * X0y : Xs[bn][1][hj][hi];
                h0 = ay * (Xs[bk][2][hj + 1][hi] - X0z) - az * (Xny - X0y);
                h1 = az * (Xnx - X0x) - ax * (Xs[bk][2][hj][hi + 1] - X0z);
                h2 = ax * (Xs[bk][1][hj][hi + 1] - X0y) - ay * (Xs[bk][0][hj + 1][hi] - X0x);
            }
            if (MU) {   // H = μ⁻¹ Cₑ X (zero points stay zero)
                h0 = h0 * hmu[r][0];
                h1 = h1 * hmu[r][1];
                h2 = h2 * hmu[r][2];
            }
            Hs[0][hj][hi] = h0;
            Hs[1][hj][hi] = h1;
            Hs[2][hj][hi] = h2;
        }
    };

    // ψ = W Div(Y) of node plane vk (Y buffers bk: plane vk, bm: plane vk−1) into Ps[pb]
    auto pplane = [&](int bk, int bm, int vk, int pb) {
        int gk = 0;
        c64 phz = cone<double>();
        bool zok = wrap_ax(g, 2, vk, gk, phz);
        const c64 ez = g.sdlinv[PRIM][2][gk];
#pragma unroll
        for (int r = 0; r < NPP; ++r) {
            int t = tid + r * C::NT;
            if (t >= C::PHt * C::PW) break;
            int pj = t / C::PW, pi = t - pj * C::PW;
            c64 psi = czero<double>();
            if (pok[r] && zok) {
                int ii = pi + 1, jj = pj + 1;   // node (i0+pi, j0+pj) in staged coordinates
                c64 dv = pex[r] * (Ys[bk][0][jj][ii] - Ys[bk][0][jj][ii - 1]) +
                           pey[r] * (Ys[bk][1][jj][ii] - Ys[bk][1][jj - 1][ii]) +
                           ez * (Ys[bk][2][jj][ii] - Ys[bm][2][jj][ii]);
                // the staged Y already carries the wrap phase of a ghost node
                psi = cvt<double>(od.wnode[pidx[r] + (int64_t)g.N[0] * g.N[1] * gk]) * dv;
            }
            Ps[pb][pj][pi] = psi;
        }
    };

    c64 xc[3], qc[3], ec[3], xn[3], qn[3], en[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { xc[c] = xn[c] = czero<double>(); qc[c] = qn[c] = ec[c] = en[c] = cone<double>(); }
    c64 hpx = czero<double>(), hpy = czero<double>();
    c64 zkm0 = czero<double>(), zkmx = czero<double>(), zkmy = czero<double>();
    if (!skip && k0 < k1) {
        // a one-cell z axis (an extruded 2D model): its single plane is staged
        // once; its neighbours are that plane times the Bloch phase, or zero
        const bool slab = (PIPE || OFFD) && g.N[2] == 1;
        const c64 pzn = g.bloch[2] ? g.ph[2] : czero<double>();
        // the owner's factors of H_x, H_y at plane k0−1 (loaded ahead of the
        // prefetch, as every per-thread load in the pipelined loop is)
        const bool hpv = own && (k0 > 0 || g.bloch[2]);
        c64 pax = czero<double>(), pay = czero<double>(), paz = czero<double>();
        c64 pmx = cone<double>(), pmy = cone<double>();   // MU: 1/μ of H_x, H_y at plane k0−1
        if (hpv) {
            paz = g.sdlinv[DUAL][2][k0 > 0 ? k0 - 1 : g.N[2] - 1];
            pax = g.sdlinv[DUAL][0][i];
            pay = g.sdlinv[DUAL][1][j];
            if (MU) {
                const int64_t pm1 = gxy(i, j) + gz(k0 > 0 ? k0 - 1 : g.N[2] - 1);
                pmx = od.muinv[3 * pm1 + 0];
                pmy = od.muinv[3 * pm1 + 1];
            }
        }
        if (slab) {   // one plane, staged once into buffer 1
            if (PIPE) {
                fetch(0);
                commit(1, 0, xc, qc, ec);
            } else stage(1, 0, xc, qc, ec);
        } else if (PIPE) {
            fetch(k0 - 1);
            commit(0, k0 - 1, xn, qn, en);
            fetch(k0);
            commit(1, k0, xc, qc, ec);
            fetch(k0 + 1);
        } else {
            stage(0, k0 - 1, xn, qn, en);
            stage(1, k0, xc, qc, ec);
        }
        __syncthreads();
        // H_x, H_y of plane k0−1 at the owned point (zero across a non-Bloch z ghost)
        if (hpv && slab) {   // plane −1 is plane 0 times e^{+ikL}, the inverse phase
            int hi = tx + 1, hj = ty + 1;
            const c64 pm = g.phinv[2], dm = cone<double>() - pm;
            c64 X0x = Xs[1][0][hj][hi], X0y = Xs[1][1][hj][hi], X0z = Xs[1][2][hj][hi];
            hpx = pay * (pm * (Xs[1][2][hj + 1][hi] - X0z)) - paz * (dm * X0y);
            hpy = paz * (dm * X0x) - pax * (pm * (Xs[1][2][hj][hi + 1] - X0z));
        } else if (hpv) {
            int hi = tx + 1, hj = ty + 1;
            c64 X0x = Xs[0][0][hj][hi], X0y = Xs[0][1][hj][hi], X0z = Xs[0][2][hj][hi];
            hpx = pay * (Xs[0][2][hj + 1][hi] - X0z) - paz * (Xs[1][1][hj][hi] - X0y);
            hpy = paz * (Xs[1][0][hj][hi] - X0x) - pax * (Xs[0][2][hj][hi + 1] - X0z);
        }
        if (MU) {
            hpx = hpx * pmx;
            hpy = hpy * pmy;
        }
        if (GD) pplane(1, 0, k0, k0 & 1);
        if (OFFD && own) {   // X_z of plane k0−1 at (i, j), (i+1, j), (i, j+1)
            // (slab: plane 0 times e^{+ikL}, or zero past a non-Bloch end)
            const c64 pm = slab ? (g.bloch[2] ? g.phinv[2] : czero<double>()) : cone<double>();
            const int b0 = slab ? 1 : 0;
            zkm0 = pm * Xs[b0][2][ty + 1][tx + 1];
            zkmx = pm * Xs[b0][2][ty + 1][tx + 2];
            zkmy = pm * Xs[b0][2][ty + 2][tx + 1];
        }
        for (int k = k0; k < k1; ++k) {
            const int bk = (k - k0 + 1) & 1, bn = bk ^ 1;
            // buffer bn, Hs and Ps[(k+1)&1] are free.  Without the grad-div
            // and off-diagonal terms the output phase reads only Hs, and the
            // last reads of buffer bn (the previous plane's H) lie before the
            // barrier that published Hs, so only the prologue's reads of the
            // first two planes need this one
            if (GD || OFFD || k == k0) __syncthreads();
            if (MU) load_mu(k);
            // z factors of plane k (scalar loads: waiting for them does not
            // wait for the prefetch)
            const c64 az_k = ldu(g.sdlinv[DUAL][2], k), ez_k = ldu(g.sdlinv[PRIM][2], k);
            const c64 tzm = axis_term_u(g, k);
            const c64 sdD = ldu(g.sdl[DUAL][2], k), sdP = ldu(g.sdl[PRIM][2], k);
            if (slab) {
            } else if (PIPE) {
                commit(bn, k + 1, xn, qn, en);
                // past the chunk: an out-of-ring plane, which loads nothing
                fetch(k + 2 <= k1 ? k + 2 : -2);
            } else stage(bn, k + 1, xn, qn, en);
            __syncthreads();
            hplane(bk, bn, az_k, slab, pzn);
            if (GD) pplane(bn, bk, k + 1, (k + 1) & 1);
            __syncthreads();
            if (own) {
                c64 H0[3], Hxm[3], Hym[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    H0[c] = Hs[c][ty + 1][tx + 1];
                    Hxm[c] = Hs[c][ty + 1][tx];
                    Hym[c] = Hs[c][ty][tx + 1];
                }
                if (OFFD) {   // the owner's 1D factors, from the block tables
                    exo = Fx[0][tx];
                    eyo = Fy[0][ty];
                    tmx = Fx[1][tx];
                    tmy = Fy[1][ty];
                    T2xy[0] = Fx[2][tx] * Fy[3][ty];
                    T2xy[1] = Fx[3][tx] * Fy[2][ty];
                    T2xy[2] = Fx[3][tx] * Fy[3][ty];
                }
                const c64 ez = ez_k;
                c64 cc[3];
                cc[0] = eyo * (H0[2] - Hym[2]) - ez * (H0[1] - hpy);
                cc[1] = ez * (H0[0] - hpx) - exo * (H0[2] - Hxm[2]);
                cc[2] = exo * (H0[1] - Hxm[1]) - eyo * (H0[0] - Hym[0]);
                hpx = H0[0];
                hpy = H0[1];
                if (GD) {   // cc −= ε G ψ (forward differences of ψ to the E edges)
                    const int pc = k & 1, pn = (k + 1) & 1;
                    c64 p0 = Ps[pc][ty][tx];
                    cc[0] = cc[0] - ec[0] * (gdx * (Ps[pc][ty][tx + 1] - p0));
                    cc[1] = cc[1] - ec[1] * (gdy * (Ps[pc][ty + 1][tx] - p0));
                    cc[2] = cc[2] - ec[2] * (az_k * (Ps[pn][ty][tx] - p0));
                }
                int64_t p = i + (int64_t)g.N[0] * j + NxNy * k;
                c64 term[3] = {czero<double>(), czero<double>(), czero<double>()};
                if (OFFD) {
                    // ∆l weights of the corner averages along each axis: x and y
                    // from the block's LDS tables, z from the plane's scalar loads
                    const c64 h5 = mk<double>(0.5, 0);
                    const int km = k > 0 ? k - 1 : g.N[2] - 1;
                    const c64 wx0 = Wx[0][tx], wx1 = Wx[1][tx], wy0 = Wy[0][ty], wy1 = Wy[1][ty];
                    const c64 wz0 = (h5 * sdD) * ez_k, wz1 = (h5 * ldu(g.sdl[DUAL][2], km)) * ez_k;
                    const int hi = tx + 1, hj = ty + 1;
                    // corner cells p + û (wrapped), or -1 past a non-Bloch end
                    int64_t cpx = i + 1 < g.N[0] ? p + 1 : (g.bloch[0] ? p - (g.N[0] - 1) : -1);
                    int64_t cpy = j + 1 < g.N[1] ? p + g.N[0] : (g.bloch[1] ? p - (int64_t)(g.N[1] - 1) * g.N[0] : -1);
                    int64_t cpz = k + 1 < g.N[2] ? p + NxNy : (g.bloch[2] ? p - (int64_t)(g.N[2] - 1) * NxNy : -1);
                    // u = x: ww = y (xy) and ww = z (xz), corners p and p + x̂
                    // (a component outside the window is not written: no term)
                    if (inw(0)) {
                        c64 s = (h5 * od.epso[3 * p + 0]) * (wy0 * Xs[bk][1][hj][hi] + wy1 * Xs[bk][1][hj - 1][hi]) +
                                (h5 * od.epso[3 * p + 1]) * (wz0 * Xs[bk][2][hj][hi] + wz1 * zkm0);
                        if (cpx >= 0)
                            s = s + (h5 * od.epso[3 * cpx + 0]) * (wy0 * Xs[bk][1][hj][hi + 1] + wy1 * Xs[bk][1][hj - 1][hi + 1]) +
                                (h5 * od.epso[3 * cpx + 1]) * (wz0 * Xs[bk][2][hj][hi + 1] + wz1 * zkmx);
                        term[0] = s;
                    }
                    // u = y: ww = x (xy) and ww = z (yz), corners p and p + ŷ
                    if (inw(1)) {
                        c64 s = (h5 * od.epso[3 * p + 0]) * (wx0 * Xs[bk][0][hj][hi] + wx1 * Xs[bk][0][hj][hi - 1]) +
                                (h5 * od.epso[3 * p + 2]) * (wz0 * Xs[bk][2][hj][hi] + wz1 * zkm0);
                        if (cpy >= 0)
                            s = s + (h5 * od.epso[3 * cpy + 0]) * (wx0 * Xs[bk][0][hj + 1][hi] + wx1 * Xs[bk][0][hj + 1][hi - 1]) +
                                (h5 * od.epso[3 * cpy + 2]) * (wz0 * Xs[bk][2][hj + 1][hi] + wz1 * zkmy);
                        term[1] = s;
                    }
                    // u = z: ww = x (xz) and ww = y (yz), corners p and p + ẑ
                    if (inw(2)) {
                        c64 s = (h5 * od.epso[3 * p + 1]) * (wx0 * Xs[bk][0][hj][hi] + wx1 * Xs[bk][0][hj][hi - 1]) +
                                (h5 * od.epso[3 * p + 2]) * (wy0 * Xs[bk][1][hj][hi] + wy1 * Xs[bk][1][hj - 1][hi]);
                        if (cpz >= 0) {   // plane k+1 (slab: plane k times the phase)
                            c64 ax_, ay_;
                            if (slab) {
                                ax_ = pzn * (wx0 * Xs[bk][0][hj][hi] + wx1 * Xs[bk][0][hj][hi - 1]);
                                ay_ = pzn * (wy0 * Xs[bk][1][hj][hi] + wy1 * Xs[bk][1][hj - 1][hi]);
                            } else {
                                ax_ = wx0 * Xs[bn][0][hj][hi] + wx1 * Xs[bn][0][hj][hi - 1];
                                ay_ = wy0 * Xs[bn][1][hj][hi] + wy1 * Xs[bn][1][hj - 1][hi];
                            }
                            s = s + (h5 * od.epso[3 * cpz + 1]) * ax_ + (h5 * od.epso[3 * cpz + 2]) * ay_;
                        }
                        term[2] = s;
                    }
                    // X_z of this plane, for the next one
                    zkm0 = Xs[bk][2][hj][hi];
                    zkmx = Xs[bk][2][hj][hi + 1];
                    zkmy = Xs[bk][2][hj + 1][hi];
                }
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    c64 out;
                    if (MODE == MODE_A) {
                        out = cc[u] - om2 * (ec[u] * xc[u]);
                        if (OFFD) out = out - om2 * term[u];
                    } else {
                        // (Q T A T⁻¹ Q x)_u with q = qs·T: the output factor
                        // q T = qs T² =: w and the mass term −ω² ε q² x =
                        // (1 − c q²) x give x + w (cc − c qs x); grad-div keeps
                        // ε: w (cc − ω² ε qs x)
                        const c64 w = qc[u] * (T2xy[u] * (u == 2 ? sdD : sdP));
                        const c64 v = xc[u] * qc[u];
                        c64 in;
                        if (GD || MU) in = cc[u] - om2 * (ec[u] * v);
                        else {
                            const c64 cd = u == 0 ? tmy + tzm : (u == 1 ? tzm + tmx : tmx + tmy);
                            in = cc[u] - cd * v;
                        }
                        if (OFFD) in = in - om2 * term[u];
                        out = w * in;
                        if (!GD && !MU) out = out + xc[u];
                    }
                    if (!inw(u)) continue;
                    int64_t e = dofidx(g, u, REV ? gxy(i, j) + gz(k) : p);
                    if (EPI == EPI_RESID) {
                        const c64 oc = rsg(u, out);   // in the caller's frame
                        c64 bb = ball[rhs * vs + e];
                        c64 rr = mk<double>(bb.re - oc.re, bb.im - oc.im);
                        if (yall) y[e] = cvt<R>(rr);   // null output: the residual norm only
                        acc[0] += (double)rr.re * rr.re + (double)rr.im * rr.im;
                    } else {
                        // the dot product xᵀy is the same in either frame
                        y[e] = cvt<R>(rsg(u, out));
                        if (EPI == EPI_DOT) {
                            acc[0] += (double)xc[u].re * out.re - (double)xc[u].im * out.im;
                            acc[1] += (double)xc[u].re * out.im + (double)xc[u].im * out.re;
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) { xc[c] = xn[c]; qc[c] = qn[c]; ec[c] = en[c]; }
        }
    }
    if (EPI == EPI_STORE) return;
    const int bid = bx + (int)nbx * (by + (int)nby * chunk);
    block_partial<2>(acc, part + (size_t)rhs * NPART * 4, bid);
}
