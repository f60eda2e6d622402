// Material assignment and Kottke subpixel smoothing on the device.
//
// Replaces calc_matparams! for ModelFull (reference src/model/full.jl:16-70),
// whose assign_param!/smooth_param! live in MaxwellBase:
//   - object indices are computed once on each of the 8 staggered lattices
//     (ghosted, (N+1)^3), at the transformed ghost coordinates; the
//     last-added object containing a point wins (0 = none);
//   - the diagonal entry (w,w) sits at the F_w locations, the off-diagonal
//     entries at the voxel corners g0; a voxel's corners are the 8 points of
//     the opposite grid types around it (full.jl:58-67 cross-wires the ε and
//     μ index arrays for exactly this reason);
//   - a voxel whose corners see exactly two objects with different tensors is
//     smoothed with Kottke's τ-average (Kottke, Farjadpour & Johnson, PRE 77,
//     036611 (2008)): fill fraction from the planar cut of the voxel at the
//     foreground shape's nearest surface point, normal from the same point;
//     three or more objects give the plain mean of the 8 corner tensors.
#include "matparams.hpp"

namespace mfd {

struct DShape {
    int kind;
    double c[3], r, s[3], lo[3], hi[3];   // bounding box lo/hi
};

__device__ inline bool inside(const DShape &s, const double x[3]) {
    if (x[0] < s.lo[0] || x[0] > s.hi[0] || x[1] < s.lo[1] || x[1] > s.hi[1] || x[2] < s.lo[2] || x[2] > s.hi[2])
        return false;
    if (s.kind == 0) {
        double a = x[0] - s.c[0], b = x[1] - s.c[1], c = x[2] - s.c[2];
        return a * a + b * b + c * c <= s.r * s.r;
    }
    if (s.kind == 2) {   // cylinder along z: radius r, height s[2]
        double a = x[0] - s.c[0], b = x[1] - s.c[1];
        return a * a + b * b <= s.r * s.r && fabs(x[2] - s.c[2]) <= 0.5 * s.s[2];
    }
    for (int a = 0; a < 3; ++a)
        if (!(fabs(x[a] - s.c[a]) <= 0.5 * s.s[a])) return false;
    return true;
}

__device__ inline void surfpt(const DShape &s, const double x[3], double r0[3], double n[3]) {
    if (s.kind == 0) {
        double d[3] = {x[0] - s.c[0], x[1] - s.c[1], x[2] - s.c[2]};
        double nd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (nd == 0.0) { n[0] = 1; n[1] = 0; n[2] = 0; }
        else { n[0] = d[0] / nd; n[1] = d[1] / nd; n[2] = d[2] / nd; }
        for (int a = 0; a < 3; ++a) r0[a] = s.c[a] + s.r * n[a];
        return;
    }
    if (s.kind == 2) {   // the nearer of the lateral surface and an end cap
        double d0 = x[0] - s.c[0], d1 = x[1] - s.c[1], dz = x[2] - s.c[2];
        double nd = sqrt(d0 * d0 + d1 * d1);
        if (nd - s.r >= fabs(dz) - 0.5 * s.s[2]) {
            if (nd == 0.0) { n[0] = 1; n[1] = 0; }
            else { n[0] = d0 / nd; n[1] = d1 / nd; }
            n[2] = 0;
            r0[0] = s.c[0] + s.r * n[0];
            r0[1] = s.c[1] + s.r * n[1];
            r0[2] = x[2];
        } else {
            double sg = dz < 0 ? -1.0 : 1.0;
            n[0] = n[1] = 0;
            n[2] = sg;
            r0[0] = x[0];
            r0[1] = x[1];
            r0[2] = s.c[2] + sg * 0.5 * s.s[2];
        }
        return;
    }
    int ax = 0;
    double best = -1e308;
    for (int a = 0; a < 3; ++a) {
        double e = fabs(x[a] - s.c[a]) - 0.5 * s.s[a];
        if (e > best) { best = e; ax = a; }
    }
    double sg = (x[ax] - s.c[ax]) < 0 ? -1.0 : 1.0;
    for (int a = 0; a < 3; ++a) { r0[a] = x[a]; n[a] = 0.0; }
    n[ax] = sg;
    r0[ax] = s.c[ax] + sg * 0.5 * s.s[ax];
}

// fraction of the unit cube with a1 t1 + a2 t2 + a3 t3 <= b (a1>=a2>=a3>=0, b <= sum/2)
__device__ inline double vf_core(double b, double a1, double a2, double a3) {
    double v;
    if (a2 <= 1e-12 * a1) v = b / a1;
    else if (a3 <= 1e-12 * a1) {
        double t = b - a2 > 0 ? b - a2 : 0.0;
        v = (b * b - t * t) / (2.0 * a1 * a2);
    } else {
        double t1 = b - a1 > 0 ? b - a1 : 0.0, t2 = b - a2 > 0 ? b - a2 : 0.0, t3 = b - a3 > 0 ? b - a3 : 0.0;
        double t13 = b - a1 - a3 > 0 ? b - a1 - a3 : 0.0, t23 = b - a2 - a3 > 0 ? b - a2 - a3 : 0.0;
        v = (b * b * b - t1 * t1 * t1 - t2 * t2 * t2 - t3 * t3 * t3 + t13 * t13 * t13 + t23 * t23 * t23) /
            (6.0 * a1 * a2 * a3);
    }
    return v < 0 ? 0 : (v > 1 ? 1 : v);
}

// volume fraction of the box [vmin,vmax] where n·(x − r0) <= 0
__device__ inline double volfrac(const double vmin[3], const double vmax[3], const double n[3], const double r0[3]) {
    double a[3], b = 0.0;
    for (int i = 0; i < 3; ++i) { a[i] = n[i] * (vmax[i] - vmin[i]); b += n[i] * (r0[i] - vmin[i]); }
    for (int i = 0; i < 3; ++i)
        if (a[i] < 0) { b -= a[i]; a[i] = -a[i]; }
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

// Kottke τ-average of two column-major 3x3 tensors in the frame (n, t1, t2)
__device__ void kottke(const c64 *efg, const c64 *ebg, double r, const double n[3], c64 *out) {
    double B[3][3];
    int k = 0;
    double an[3] = {fabs(n[0]), fabs(n[1]), fabs(n[2])};
    if (an[1] < an[k]) k = 1;
    if (an[2] < an[k]) k = 2;
    double e[3] = {0, 0, 0};
    e[k] = 1.0;
    double t1[3] = {n[1] * e[2] - n[2] * e[1], n[2] * e[0] - n[0] * e[2], n[0] * e[1] - n[1] * e[0]};
    double nt = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    for (int i = 0; i < 3; ++i) t1[i] /= nt;
    double t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
    for (int i = 0; i < 3; ++i) { B[i][0] = n[i]; B[i][1] = t1[i]; B[i][2] = t2[i]; }
    c64 tau[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) tau[a][b] = mk<double>(0, 0);
    for (int m = 0; m < 2; ++m) {
        const c64 *E = m == 0 ? efg : ebg;
        double f = m == 0 ? r : 1.0 - r;
        c64 P[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                c64 s = mk<double>(0, 0);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) s = s + B[j][b] * (B[i][a] * E[i + 3 * j]);
                P[a][b] = s;
            }
        c64 t[3][3];
        t[0][0] = cdiv(mk<double>(-1.0, 0.0), P[0][0]);
        t[0][1] = cdiv(P[0][1], P[0][0]);
        t[0][2] = cdiv(P[0][2], P[0][0]);
        t[1][0] = cdiv(P[1][0], P[0][0]);
        t[2][0] = cdiv(P[2][0], P[0][0]);
        for (int a = 1; a < 3; ++a)
            for (int b = 1; b < 3; ++b) t[a][b] = P[a][b] - cdiv(P[a][0] * P[0][b], P[0][0]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) tau[a][b] = tau[a][b] + f * t[a][b];
    }
    c64 Q[3][3];
    Q[0][0] = cdiv(mk<double>(-1.0, 0.0), tau[0][0]);
    Q[0][1] = cdiv(-tau[0][1], tau[0][0]);
    Q[0][2] = cdiv(-tau[0][2], tau[0][0]);
    Q[1][0] = cdiv(-tau[1][0], tau[0][0]);
    Q[2][0] = cdiv(-tau[2][0], tau[0][0]);
    for (int a = 1; a < 3; ++a)
        for (int b = 1; b < 3; ++b) Q[a][b] = tau[a][b] - cdiv(tau[a][0] * tau[0][b], tau[0][0]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            c64 s = mk<double>(0, 0);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) s = s + B[j][b] * (B[i][a] * Q[a][b]);
            out[i + 3 * j] = s;
        }
}

struct LatArgs {
    const double *tlg[2][3];   // transformed ghosted coordinates, N+1 each
    const double *lg[2][3];    // ghosted coordinates
    const double *l[2][3];     // non-ghost coordinates
    int N[3];
};

__global__ void k_classify(LatArgs A, int gx, int gy, int gz, const DShape *sh, int nobj, int *oind) {
    int64_t Nx = A.N[0] + 1, Ny = A.N[1] + 1, Nz = A.N[2] + 1;
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= Nx * Ny * Nz) return;
    int64_t i = e % Nx, j = (e / Nx) % Ny, k = e / (Nx * Ny);
    double x[3] = {A.tlg[gx][0][i], A.tlg[gy][1][j], A.tlg[gz][2][k]};
    int o = 0;
    for (int b = nobj; b >= 1; --b)
        if (inside(sh[b - 1], x)) { o = b; break; }
    oind[e] = o;
}

// one pass over the cells: entries ent[0..nent) of the tensor at grid types fgt
struct PassArgs {
    int fgt[3];
    int nent;
    int ent[6][2];
};

__global__ void k_smooth(LatArgs A, PassArgs P, const int *oind_ctr, const int *oind_cnr, const DShape *sh,
                         const c64 *tens, c64 *arr) {
    int64_t M = (int64_t)A.N[0] * A.N[1] * A.N[2];
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= M) return;
    int q[3] = {(int)(p % A.N[0]), (int)((p / A.N[0]) % A.N[1]), (int)(p / ((int64_t)A.N[0] * A.N[1]))};
    int64_t Nx1 = A.N[0] + 1, Ny1 = A.N[1] + 1;
    int t[3];
    for (int a = 0; a < 3; ++a) t[a] = P.fgt[a] == PRIM ? q[a] : q[a] + 1;
    int oc = oind_ctr[t[0] + Nx1 * (t[1] + Ny1 * t[2])];
    int co[8], dist[8], nd = 0;
    for (int c = 0; c < 8; ++c) {
        int64_t ci = q[0] + (c & 1), cj = q[1] + ((c >> 1) & 1), ck = q[2] + ((c >> 2) & 1);
        co[c] = oind_cnr[ci + Nx1 * (cj + Ny1 * ck)];
        bool f = false;
        for (int x = 0; x < nd; ++x)
            if (dist[x] == co[c]) f = true;
        if (!f) dist[nd++] = co[c];
    }
    c64 val[9];
    bool smoothed = false;
    if (nd == 2 && dist[0] > 0 && dist[1] > 0) {
        int fg = dist[0] > dist[1] ? dist[0] : dist[1], bg = dist[0] > dist[1] ? dist[1] : dist[0];
        const c64 *ef = tens + 9 * (fg - 1), *eb = tens + 9 * (bg - 1);
        bool same = true;
        for (int x = 0; x < 9; ++x)
            if (ef[x].re != eb[x].re || ef[x].im != eb[x].im) same = false;
        if (!same) {
            double x0[3], vmin[3], vmax[3], r0[3], nn[3];
            for (int a = 0; a < 3; ++a) {
                x0[a] = A.l[P.fgt[a]][a][q[a]];
                vmin[a] = A.lg[1 - P.fgt[a]][a][q[a]];
                vmax[a] = A.lg[1 - P.fgt[a]][a][q[a] + 1];
            }
            surfpt(sh[fg - 1], x0, r0, nn);
            double r = volfrac(vmin, vmax, nn, r0);
            kottke(ef, eb, r, nn, val);
            smoothed = true;
        }
    } else if (nd >= 2) {
        for (int x = 0; x < 9; ++x) {
            c64 s = mk<double>(0, 0);
            for (int c = 0; c < 8; ++c)
                if (co[c] > 0) s = s + tens[9 * (co[c] - 1) + x];
            val[x] = mk<double>(s.re / 8.0, s.im / 8.0);
        }
        smoothed = true;
    }
    for (int e = 0; e < P.nent; ++e) {
        int v = P.ent[e][0], w = P.ent[e][1];
        c64 out;
        if (smoothed) out = val[v + 3 * w];
        else out = oc > 0 ? tens[9 * (oc - 1) + v + 3 * w] : mk<double>(0, 0);
        arr[p + M * (v + 3 * w)] = out;
    }
}

int matparams(mfd_ctx *ctx, const mfd_grid *g, int64_t nobj, const mfd_shape *shapes, const mfd_c64 *eps_t,
              const mfd_c64 *mu_t, mfd_c64 *eps_arr, mfd_c64 *mu_arr) {
    HGrid h;
    MFD_TRY(build_hgrid(g, &h));
    if (nobj < 0 || (nobj > 0 && (!shapes || !eps_t || !mu_t))) { set_error("matparams: bad object list"); return MFD_ERR_ARG; }
    hipStream_t s = ctx->stream;
    int64_t M = h.M;
    int64_t nl = (h.N[0] + 1) * (h.N[1] + 1) * (h.N[2] + 1);
    std::vector<DShape> ds(nobj > 0 ? nobj : 1);
    for (int64_t o = 0; o < nobj; ++o) {
        const mfd_shape &a = shapes[o];
        DShape &d = ds[o];
        d.kind = a.kind;
        if (a.kind != 0 && a.kind != 1 && a.kind != 2) {
            set_error("matparams: shape kind must be 0 (ball), 1 (cuboid) or 2 (cylinder along z)");
            return MFD_ERR_ARG;
        }
        for (int k = 0; k < 3; ++k) {
            d.c[k] = a.c[k];
            d.s[k] = a.s[k];
            // conservative box: the exact test below decides
            double half = a.kind == 0 ? a.r : (a.kind == 2 && k < 2 ? a.r : 0.5 * a.s[k]);
            double pad = 1e-9 * (fabs(a.c[k]) + fabs(half)) + 1e-300;
            d.lo[k] = a.c[k] - half - pad;
            d.hi[k] = a.c[k] + half + pad;
        }
        d.r = a.r;
    }
    // coordinate tables
    size_t ncoord = 0;
    for (int w = 0; w < 3; ++w) ncoord += 2 * (3 * (h.N[w] + 1));
    size_t bytes = sizeof(double) * ncoord + sizeof(DShape) * ds.size() + sizeof(c64) * 18 * (nobj > 0 ? nobj : 1) +
                   sizeof(int) * 8 * nl + sizeof(c64) * 9 * M + 1024;
    void *buf = nullptr;
    MFD_HIP(hipMalloc(&buf, bytes));
    char *p = (char *)buf;
    std::vector<double> hc;
    hc.reserve(ncoord);
    LatArgs A;
    size_t off = 0;
    for (int t = 0; t < 2; ++t)
        for (int w = 0; w < 3; ++w) {
            A.tlg[t][w] = (const double *)p + off;
            hc.insert(hc.end(), h.tlg[t][w].begin(), h.tlg[t][w].end());
            off += h.N[w] + 1;
            A.lg[t][w] = (const double *)p + off;
            hc.insert(hc.end(), h.lg[t][w].begin(), h.lg[t][w].end());
            off += h.N[w] + 1;
            A.l[t][w] = (const double *)p + off;
            hc.insert(hc.end(), h.l[t][w].begin(), h.l[t][w].end());
            hc.push_back(0.0);
            off += h.N[w] + 1;
        }
    for (int w = 0; w < 3; ++w) A.N[w] = (int)h.N[w];
    int rc = MFD_OK;
    hipError_t e = hipSuccess;
    char *q = p + sizeof(double) * ncoord;
    DShape *dsh = (DShape *)q; q += sizeof(DShape) * ds.size();
    c64 *et = (c64 *)q; q += sizeof(c64) * 9 * (nobj > 0 ? nobj : 1);
    c64 *mt = (c64 *)q; q += sizeof(c64) * 9 * (nobj > 0 ? nobj : 1);
    int *lat = (int *)q; q += sizeof(int) * 8 * nl;
    q = (char *)(((uintptr_t)q + 255) & ~(uintptr_t)255);
    c64 *arr = (c64 *)q;
    e = hipMemcpyAsync(p, hc.data(), sizeof(double) * hc.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dsh, ds.data(), sizeof(DShape) * ds.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nobj > 0) e = hipMemcpyAsync(et, eps_t, sizeof(c64) * 9 * nobj, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nobj > 0) e = hipMemcpyAsync(mt, mu_t, sizeof(c64) * 9 * nobj, hipMemcpyHostToDevice, s);
    for (int c = 0; c < 8 && e == hipSuccess; ++c) {
        k_classify<<<(unsigned)((nl + 255) / 256), 256, 0, s>>>(A, c & 1, (c >> 1) & 1, (c >> 2) & 1, dsh, (int)nobj,
                                                               lat + c * nl);
        e = hipGetLastError();
    }
    for (int m = 0; m < 2 && e == hipSuccess; ++m) {
        int ft = m == 0 ? EE : HH;
        const c64 *tens = m == 0 ? et : mt;
        for (int pass = 0; pass < 4 && e == hipSuccess; ++pass) {
            PassArgs P;
            int cgt[3];
            if (pass < 3) {
                int w = pass;
                for (int a = 0; a < 3; ++a) { P.fgt[a] = gtype(h.bft, ft, w, a); cgt[a] = 1 - P.fgt[a]; }
                P.nent = 1;
                P.ent[0][0] = w;
                P.ent[0][1] = w;
            } else {
                for (int a = 0; a < 3; ++a) { P.fgt[a] = g0type(h.bft, ft, a); cgt[a] = 1 - P.fgt[a]; }
                const int ent[6][2] = {{1, 0}, {2, 0}, {0, 1}, {2, 1}, {0, 2}, {1, 2}};
                P.nent = 6;
                for (int x = 0; x < 6; ++x) { P.ent[x][0] = ent[x][0]; P.ent[x][1] = ent[x][1]; }
            }
            const int *ctr = lat + (P.fgt[0] + 2 * P.fgt[1] + 4 * P.fgt[2]) * nl;
            const int *cnr = lat + (cgt[0] + 2 * cgt[1] + 4 * cgt[2]) * nl;
            k_smooth<<<(unsigned)((M + 255) / 256), 256, 0, s>>>(A, P, ctr, cnr, dsh, tens, arr);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(m == 0 ? eps_arr : mu_arr, arr, sizeof(c64) * 9 * M, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) {
        set_error(std::string("matparams: ") + hipGetErrorString(e));
        rc = MFD_ERR_HIP;
    }
    hipStreamSynchronize(s);
    hipFree(buf);
    return rc;
}

}  // namespace mfd
