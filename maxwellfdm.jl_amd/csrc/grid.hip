// Grid tables and the PML-stretched ∆l (SC-PML) on the device.
//
// Replaces MaxwellBase create_stretched_∆l / invert_∆l / t_ind as called from
// create_stretched_∆ls (reference src/model/model.jl:122-139).  Grid
// conventions follow the reference's use of Grid (test/source.jl:6-27,
// pointsrc.jl:66-68): the primal ghost sits at the positive bound, the dual
// ghost at the negative end; ∆l[PRIM] = diff(dual incl. ghost),
// ∆l[DUAL] = diff(primal incl. ghost).
#include "grid.hpp"
#include <cmath>
#include <cstring>
#include <mutex>

namespace mfd {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

int build_hgrid(const mfd_grid *g, HGrid *h) {
    if (!g) { set_error("null grid"); return MFD_ERR_ARG; }
    h->M = 1;
    for (int w = 0; w < 3; ++w) {
        int64_t N = g->N[w];
        if (N < 1 || !g->lprim[w]) { set_error("grid: N must be >= 1 with lprim given"); return MFD_ERR_ARG; }
        h->N[w] = N;
        h->M *= N;
        h->bloch[w] = g->isbloch[w] != 0;
        h->bft[w] = g->boundft[w];
        if (h->bft[w] != EE && h->bft[w] != HH) { set_error("boundft must be EE or HH"); return MFD_ERR_ARG; }
        h->npml[0][w] = g->npml[0][w];
        h->npml[1][w] = g->npml[1][w];
        if (g->npml[0][w] < 0 || g->npml[1][w] < 0 || g->npml[0][w] + g->npml[1][w] > N) {
            set_error("Npml out of range");
            return MFD_ERR_ARG;
        }
        h->kb[w] = g->kbloch[w];
        const double *lp = g->lprim[w];
        for (int64_t i = 0; i < N; ++i)
            if (!(lp[i + 1] > lp[i])) { set_error("lprim must be strictly increasing"); return MFD_ERR_ARG; }
        for (int t = 0; t < 2; ++t) {
            h->l[t][w].resize(N);
            h->lg[t][w].resize(N + 1);
            h->tlg[t][w].resize(N + 1);
            h->dl[t][w].resize(N);
        }
        double L = lp[N] - lp[0];
        h->bnd[0][w] = lp[0];
        h->bnd[1][w] = lp[N];
        h->L[w] = L;
        for (int64_t i = 0; i <= N; ++i) h->lg[PRIM][w][i] = lp[i];
        for (int64_t i = 0; i < N; ++i) {
            h->l[PRIM][w][i] = lp[i];
            h->l[DUAL][w][i] = 0.5 * (lp[i] + lp[i + 1]);
            h->lg[DUAL][w][i + 1] = h->l[DUAL][w][i];
        }
        // dual ghost: periodic image for Bloch, mirror image about the bound otherwise
        h->lg[DUAL][w][0] = h->bloch[w] ? h->l[DUAL][w][N - 1] - L : 2.0 * lp[0] - h->l[DUAL][w][0];
        h->tlg[PRIM][w] = h->lg[PRIM][w];
        h->tlg[DUAL][w] = h->lg[DUAL][w];
        h->tlg[PRIM][w][N] = h->bloch[w] ? lp[0] : lp[N];
        h->tlg[DUAL][w][0] = h->bloch[w] ? h->l[DUAL][w][N - 1] : h->l[DUAL][w][0];
        for (int64_t i = 0; i < N; ++i) {
            h->dl[PRIM][w][i] = h->lg[DUAL][w][i + 1] - h->lg[DUAL][w][i];
            h->dl[DUAL][w][i] = h->lg[PRIM][w][i + 1] - h->lg[PRIM][w][i];
        }
    }
    h->cmpfirst = g->order_cmpfirst != 0;
    h->om[0] = g->omega_pml.re;
    h->om[1] = g->omega_pml.im;
    h->m = g->pml_m;
    h->lnR = g->pml_lnR;
    bool need = false;
    for (int w = 0; w < 3; ++w) need |= (h->npml[0][w] > 0) || (h->npml[1][w] > 0);
    if (need && h->om[0] == 0.0 && h->om[1] == 0.0) {
        set_error("PML requested with ωpml = 0; call set_ωpml! first");
        return MFD_ERR_ARG;
    }
    return MFD_OK;
}

struct StretchArgs {
    const double *pos[2][3];
    const double *dl[2][3];
    const double *lprim[3];
    int64_t N[3];
    int64_t npml[2][3];
    int64_t off[2][3];
    double om_re, om_im, m, lnR;
    int any_pml;
};

// K1: one thread per table entry; s = 1 + σ/(iωpml), σ = σmax d^m.
__global__ void k_stretch(StretchArgs a, int64_t total, c64 *sdl, c64 *sdlinv, c64 *ssq, c64 *ssqinv) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= total) return;
    int t = 0, w = 0;
    for (int tt = 0; tt < 2; ++tt)
        for (int ww = 0; ww < 3; ++ww)
            if (e >= a.off[tt][ww] && e < a.off[tt][ww] + a.N[ww]) { t = tt; w = ww; }
    int64_t i = e - a.off[t][w];
    double u = a.pos[t][w][i], dl = a.dl[t][w][i];
    const double *lp = a.lprim[w];
    int64_t N = a.N[w];
    double sigma = 0.0;
    if (a.npml[0][w] > 0) {
        double lpml = lp[a.npml[0][w]], Lp = lpml - lp[0];
        if (u < lpml) {
            double d = (lpml - u) / Lp;
            sigma = (-(a.m + 1.0) * a.lnR / (2.0 * Lp)) * pow(d, a.m);
        }
    }
    if (a.npml[1][w] > 0) {
        double lpml = lp[N - a.npml[1][w]], Lp = lp[N] - lpml;
        if (u > lpml) {
            double d = (u - lpml) / Lp;
            sigma = (-(a.m + 1.0) * a.lnR / (2.0 * Lp)) * pow(d, a.m);
        }
    }
    c64 s = mk<double>(1.0, 0.0);
    if (a.any_pml && sigma != 0.0) {
        double den = a.om_re * a.om_re + a.om_im * a.om_im;
        s = mk<double>(1.0 - sigma * a.om_im / den, -(sigma * a.om_re / den));
    }
    c64 v = mk<double>(s.re * dl, s.im * dl);
    sdl[e] = v;
    sdlinv[e] = cinv(v);
    c64 r = csqrt_(v);
    ssq[e] = r;
    ssqinv[e] = cinv(r);
}

__global__ void k_to32(const c64 *a, c32 *b, int64_t n) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < n) b[e] = mk<float>((float)a[e].re, (float)a[e].im);
}

DevGrid::~DevGrid() { release(); }

void DevGrid::release() {
    if (buf) hipFree(buf);
    buf = nullptr;
    if (rbuf) hipFree(rbuf);
    rbuf = nullptr;
}

// reflected tables: dst[t][w][i] = src[1−t][w][N_w−1−i] for the four kinds
// along the axes of mask, dst = src along the others
struct ReflectArgs {
    int64_t off[2][3], N[3];
    int mask;
};
__global__ void k_reflect_tables(ReflectArgs a, int64_t tot, const c64 *s0, const c64 *s1, const c64 *s2,
                                 const c64 *s3, c64 *d0, c64 *d1, c64 *d2, c64 *d3) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= tot) return;
    int t = 0, w = 0;
    for (int tt = 0; tt < 2; ++tt)
        for (int ww = 0; ww < 3; ++ww)
            if (e >= a.off[tt][ww] && e < a.off[tt][ww] + a.N[ww]) { t = tt; w = ww; }
    const int64_t i = e - a.off[t][w];
    const int64_t src = ((a.mask >> w) & 1) ? a.off[1 - t][w] + (a.N[w] - 1 - i) : e;
    d0[e] = s0[src];
    d1[e] = s1[src];
    d2[e] = s2[src];
    d3[e] = s3[src];
}

int DevGrid::reflect(int mask, hipStream_t stream) {
    if (rbuf) return MFD_OK;
    MFD_HIP(hipMalloc(&rbuf, sizeof(c64) * 4 * total));
    c64 *dst[4];
    for (int k = 0; k < 4; ++k) dst[k] = (c64 *)rbuf + k * total;
    ReflectArgs a;
    int64_t off = 0;
    for (int t = 0; t < 2; ++t)
        for (int w = 0; w < 3; ++w) {
            a.off[t][w] = off;
            off += h.N[w];
        }
    for (int w = 0; w < 3; ++w) a.N[w] = h.N[w];
    a.mask = mask;
    // the four kinds of d64, in the order fill_view lays them out
    const c64 *src[4] = {d64.sdl[PRIM][0], d64.sdlinv[PRIM][0], d64.ssq[PRIM][0], d64.ssqinv[PRIM][0]};
    const int bs = 256;
    k_reflect_tables<<<(unsigned)((total + bs - 1) / bs), bs, 0, stream>>>(a, total, src[0], src[1], src[2], src[3],
                                                                           dst[0], dst[1], dst[2], dst[3]);
    MFD_HIP(hipGetLastError());
    r64 = d64;
    for (int t = 0; t < 2; ++t)
        for (int w = 0; w < 3; ++w) {
            r64.sdl[t][w] = dst[0] + a.off[t][w];
            r64.sdlinv[t][w] = dst[1] + a.off[t][w];
            r64.ssq[t][w] = dst[2] + a.off[t][w];
            r64.ssqinv[t][w] = dst[3] + a.off[t][w];
        }
    for (int w = 0; w < 3; ++w) {
        r64.bft[w] = EE;
        if ((mask >> w) & 1) {
            r64.ph[w] = d64.phinv[w];
            r64.phinv[w] = d64.ph[w];
        }
    }
    r64.tdof = r64.tdofinv = nullptr;
    r64.rmask = mask;
    return MFD_OK;
}

int DevGrid::create(const mfd_grid *g, hipStream_t stream) {
    MFD_TRY(build_hgrid(g, &h));
    int64_t tot = 2 * (h.N[0] + h.N[1] + h.N[2]);
    total = tot;
    // device layout: positions, dl, lprim (double); 4 complex double tables; 4 complex float tables
    int64_t nl = h.N[0] + h.N[1] + h.N[2] + 3;
    size_t bytes = sizeof(double) * (2 * tot + nl) + sizeof(c64) * 4 * tot + sizeof(c32) * 4 * tot + 256;
    MFD_HIP(hipMalloc(&buf, bytes));
    char *p = (char *)buf;
    double *pos = (double *)p; p += sizeof(double) * tot;
    double *dl = (double *)p; p += sizeof(double) * tot;
    double *lpd = (double *)p; p += sizeof(double) * nl;
    c64 *t64[4];
    for (int k = 0; k < 4; ++k) { t64[k] = (c64 *)p; p += sizeof(c64) * tot; }
    c32 *t32[4];
    for (int k = 0; k < 4; ++k) { t32[k] = (c32 *)p; p += sizeof(c32) * tot; }
    std::vector<double> hpos(tot), hdl(tot), hlp(nl);
    StretchArgs a;
    int64_t off = 0, lo = 0;
    for (int w = 0; w < 3; ++w) {
        for (int64_t i = 0; i <= h.N[w]; ++i) hlp[lo + i] = h.lg[PRIM][w][i];
        a.lprim[w] = lpd + lo;
        lo += h.N[w] + 1;
    }
    for (int t = 0; t < 2; ++t)
        for (int w = 0; w < 3; ++w) {
            a.off[t][w] = off;
            for (int64_t i = 0; i < h.N[w]; ++i) {
                hpos[off + i] = h.l[t][w][i];
                hdl[off + i] = h.dl[t][w][i];
            }
            a.pos[t][w] = pos + off;
            a.dl[t][w] = dl + off;
            off += h.N[w];
        }
    for (int w = 0; w < 3; ++w) {
        a.N[w] = h.N[w];
        a.npml[0][w] = h.npml[0][w];
        a.npml[1][w] = h.npml[1][w];
    }
    a.om_re = h.om[0];
    a.om_im = h.om[1];
    a.m = h.m;
    a.lnR = h.lnR;
    a.any_pml = 0;
    for (int w = 0; w < 3; ++w) a.any_pml |= (h.npml[0][w] > 0) || (h.npml[1][w] > 0);
    MFD_HIP(hipMemcpyAsync(pos, hpos.data(), sizeof(double) * tot, hipMemcpyHostToDevice, stream));
    MFD_HIP(hipMemcpyAsync(dl, hdl.data(), sizeof(double) * tot, hipMemcpyHostToDevice, stream));
    MFD_HIP(hipMemcpyAsync(lpd, hlp.data(), sizeof(double) * nl, hipMemcpyHostToDevice, stream));
    int bs = 256;
    k_stretch<<<(unsigned)((tot + bs - 1) / bs), bs, 0, stream>>>(a, tot, t64[0], t64[1], t64[2], t64[3]);
    MFD_HIP(hipGetLastError());
    for (int k = 0; k < 4; ++k) k_to32<<<(unsigned)((tot + bs - 1) / bs), bs, 0, stream>>>(t64[k], t32[k], tot);
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipStreamSynchronize(stream));
    sdl64 = t64[0];
    fill(d64, t64);
    fill(d32, t32);
    return MFD_OK;
}

template <typename R>
static void fill_view(const HGrid &h, DGrid<R> &d, cx<R> *const *tab) {
    int64_t off = 0;
    for (int w = 0; w < 3; ++w) d.N[w] = (int)h.N[w];
    d.M = h.M;
    for (int w = 0; w < 3; ++w) {
        d.bloch[w] = h.bloch[w];
        d.bft[w] = h.bft[w];
        double kl = h.kb[w] * h.L[w];
        c64 ph = mk<double>(cos(kl), -sin(kl));
        c64 pi = cinv(ph);
        d.ph[w] = mk<R>((R)ph.re, (R)ph.im);
        d.phinv[w] = mk<R>((R)pi.re, (R)pi.im);
    }
    d.cmpfirst = h.cmpfirst;
    for (int t = 0; t < 2; ++t)
        for (int w = 0; w < 3; ++w) {
            d.sdl[t][w] = tab[0] + off;
            d.sdlinv[t][w] = tab[1] + off;
            d.ssq[t][w] = tab[2] + off;
            d.ssqinv[t][w] = tab[3] + off;
            off += h.N[w];
        }
}

void DevGrid::fill(DGrid<double> &d, c64 *const *t) { fill_view(h, d, t); }
void DevGrid::fill(DGrid<float> &d, c32 *const *t) { fill_view(h, d, t); }

}  // namespace mfd
