// Sparse CSC assembly on the device: curl operators, material operators and
// A = Cₘ (Pμ \ Cₑ) − ω² Pε, in Julia SparseMatrixCSC{ComplexF64,Int64} form
// (1-based, rows ascending per column, duplicates summed, explicit zeros kept
// as sparse(I,J,V) does; the final subtraction drops numerical zeros as
// Julia's sparse `-` does).
//
// Every kernel assigns one thread per column and enumerates that column's
// entries in closed form (no triplet sort, no general SpGEMM): two passes,
// count then fill, with a device prefix sum in between.
//
// Reference call sites: create_curls (src/model/model.jl:160-175),
// create_paramops (:141-158), create_A (:225-246).
#include "assembly.hpp"
#include <hipcub/hipcub.hpp>

namespace mfd {

__device__ inline void unc3(const DGrid<double> &g, int64_t p, int q[3]) {
    q[0] = (int)(p % g.N[0]);
    q[1] = (int)((p / g.N[0]) % g.N[1]);
    q[2] = (int)(p / ((int64_t)g.N[0] * g.N[1]));
}
__device__ inline int64_t stride3(const DGrid<double> &g, int v) {
    return v == 0 ? 1 : (v == 1 ? (int64_t)g.N[0] : (int64_t)g.N[0] * g.N[1]);
}
__device__ inline void col2cp(const DGrid<double> &g, int64_t col, int &c, int64_t &p) {
    if (g.cmpfirst) { c = (int)(col % 3); p = col / 3; }
    else { c = (int)(col / g.M); p = col % g.M; }
}

// small sorted entry list
struct Ent {
    int64_t row;
    int key;
    c64 v;
};
template <int CAP>
struct EList {
    Ent e[CAP];
    int n = 0;
    __device__ void push(int64_t r, int key, c64 v) { e[n].row = r; e[n].key = key; e[n].v = v; ++n; }
    // stable sort by (row, key) then merge equal rows summing in key order
    __device__ void sort_merge() {
        for (int x = 1; x < n; ++x) {
            Ent t = e[x];
            int y = x - 1;
            while (y >= 0 && (e[y].row > t.row || (e[y].row == t.row && e[y].key > t.key))) { e[y + 1] = e[y]; --y; }
            e[y + 1] = t;
        }
        int m = 0;
        for (int x = 0; x < n;) {
            Ent t = e[x++];
            while (x < n && e[x].row == t.row) { t.v = t.v + e[x].v; ++x; }
            e[m++] = t;
        }
        n = m;
    }
};

// Column (w, p) of Cₑ (which = 0) or Cₘ (which = 1).  Row generation of a
// curl row H_u(r) = ∂_v F_w − ∂_w F_v, v = u+1, w = u+2: diagonal term first,
// then the neighbour term; a non-Bloch ghost is zero.
__device__ void curl_col(const DGrid<double> &g, int which, int w, int64_t p, EList<4> &L) {
    int q[3];
    unc3(g, p, q);
    int ftout = which == 0 ? HH : EE;
    for (int t = 0; t < 2; ++t) {
        int u = t == 0 ? (w + 1) % 3 : (w + 2) % 3;
        int v = t == 0 ? (w + 2) % 3 : (w + 1) % 3;
        double sg = t == 0 ? 1.0 : -1.0;
        int isfwd = which == 0 ? (g.bft[v] == EE) : (g.bft[v] == HH);
        int gi = g0type(g.bft, ftout, v);
        const c64 *inv = g.sdlinv[gi][v];
        int Nv = g.N[v];
        int64_t st = stride3(g, v);
        int key = 0;
        if (isfwd) {
            c64 dg = -(sg * inv[q[v]]);
            if (q[v] > 0) {
                L.push(dofidx(g, u, p), key, dg);
                L.push(dofidx(g, u, p - st), key + 1, sg * inv[q[v] - 1]);
            } else if (g.bloch[v]) {
                c64 val = (sg * inv[Nv - 1]) * g.ph[v];
                int64_t r = p + (int64_t)(Nv - 1) * st;
                if (Nv == 1) L.push(dofidx(g, u, p), key, dg + val);
                else { L.push(dofidx(g, u, p), key, dg); L.push(dofidx(g, u, r), key + 1, val); }
            } else L.push(dofidx(g, u, p), key, dg);
        } else {
            c64 dg = sg * inv[q[v]];
            if (q[v] < Nv - 1) {
                L.push(dofidx(g, u, p), key, dg);
                L.push(dofidx(g, u, p + st), key + 1, -(sg * inv[q[v] + 1]));
            } else if (g.bloch[v]) {
                c64 val = (-(sg * inv[0])) * g.phinv[v];
                int64_t r = p - (int64_t)(Nv - 1) * st;
                if (Nv == 1) L.push(dofidx(g, u, p), key, dg + val);
                else { L.push(dofidx(g, u, p), key, dg); L.push(dofidx(g, u, r), key + 1, val); }
            } else L.push(dofidx(g, u, p), key, dg);
        }
    }
    L.sort_merge();
}

// Corners that the field value F_w(qc) contributes to, with their weights:
// the ∆l-weighted average along w (direction isfwd_in[w]; weights s∆l at the
// field planes over 2 s∆l at the corner planes) that create_mean performs for
// Pε/Pμ and for create_Mcs (model.jl:287-306).  Returns the entry count (1 or 2).
__device__ inline int corner_entries(const DGrid<double> &g, int which, int w, int64_t qc, int64_t cor[2],
                                     c64 wt[2]) {
    int ft = which == 0 ? EE : HH;
    int q[3];
    unc3(g, qc, q);
    int gc = g0type(g.bft, ft, w), gf = 1 - gc;
    int fwd_in_w = which == 0 ? (g.bft[w] != EE) : (g.bft[w] != HH);
    int Nw = g.N[w];
    int64_t stw = stride3(g, w);
    int n = 0;
    for (int ic = 0; ic < 2; ++ic) {
        int64_t c = qc;
        int cw = q[w], fi = q[w];
        c64 phin = mk<double>(1.0, 0.0);
        if (ic == 1) {
            if (!fwd_in_w) {   // corner c takes fields c and c - ŵ: q is the c - ŵ of c = q + ŵ
                if (q[w] < Nw - 1) { c = qc + stw; cw = q[w] + 1; }
                else if (g.bloch[w]) { c = qc - (int64_t)(Nw - 1) * stw; cw = 0; phin = g.phinv[w]; }
                else continue;
            } else {           // corner c takes fields c and c + ŵ
                if (q[w] > 0) { c = qc - stw; cw = q[w] - 1; }
                else if (g.bloch[w]) { c = qc + (int64_t)(Nw - 1) * stw; cw = Nw - 1; phin = g.ph[w]; }
                else continue;
            }
        }
        cor[n] = c;
        wt[n] = ((0.5 * g.sdl[gf][w][fi]) * g.sdlinv[gc][w][cw]) * phin;
        ++n;
    }
    return n;
}

// Column (w, q) of the corner-mean operator Mcₑ (which = 0) or Mcₘ (which = 1):
// F_w at q goes to the w-component slot of its corners.
__device__ void mean_col(const DGrid<double> &g, int which, int w, int64_t qc, EList<2> &L) {
    int64_t cor[2];
    c64 wt[2];
    int n = corner_entries(g, which, w, qc, cor, wt);
    for (int x = 0; x < n; ++x) L.push(dofidx(g, w, cor[x]), x, wt[x]);
    L.sort_merge();
}

// Column (w, q) of Pε (which = 0) or Pμ (which = 1).  The off-diagonal block
// (v, w) takes F_w to the corners of grid g0 with a ∆l-weighted average along
// w (direction isfwd_in[w]; weights s∆l at the field planes over 2 s∆l at the
// corner planes), multiplies by parr(v,w) at the corner, and takes the corner
// values to F_v by a half-half average along v in the other direction.
__device__ void param_col(const DGrid<double> &g, int which, const c64 *parr, int full, int w, int64_t qc,
                          EList<9> &L) {
    int64_t M = g.M;
    L.push(dofidx(g, w, qc), 0, parr[qc + M * (w + 3 * w)]);
    if (!full) return;
    int64_t cor[2];
    c64 wt[2];
    const int ncor = corner_entries(g, which, w, qc, cor, wt);
    for (int ic = 0; ic < ncor; ++ic) {
        const int64_t c = cor[ic];
        const c64 wi = wt[ic];
        int cq[3];
        unc3(g, c, cq);
        for (int v = 0; v < 3; ++v) {
            if (v == w) continue;
            int fwd_in_v = which == 0 ? (g.bft[v] != EE) : (g.bft[v] != HH);
            int Nv = g.N[v];
            int64_t stv = stride3(g, v);
            c64 ec = parr[c + M * (v + 3 * w)];
            for (int oc = 0; oc < 2; ++oc) {
                int64_t pr = c;
                c64 fo = mk<double>(0.5, 0.0);
                if (oc == 1) {
                    if (!fwd_in_v) {   // row p takes corners p and p + v̂
                        if (cq[v] > 0) pr = c - stv;
                        else if (g.bloch[v]) { pr = c + (int64_t)(Nv - 1) * stv; fo = 0.5 * g.ph[v]; }
                        else continue;
                    } else {           // row p takes corners p and p - v̂
                        if (cq[v] < Nv - 1) pr = c + stv;
                        else if (g.bloch[v]) { pr = c - (int64_t)(Nv - 1) * stv; fo = 0.5 * g.phinv[v]; }
                        else continue;
                    }
                }
                L.push(dofidx(g, v, pr), 1 + 2 * oc + ic, ec * (fo * wi));
            }
        }
    }
    L.sort_merge();
}

template <bool FILL>
__global__ void k_curl(DGrid<double> g, int which, int64_t n, int64_t *cnt, const int64_t *colptr,
                       int64_t *rowval, c64 *nzval) {
    int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (col >= n) return;
    int w;
    int64_t p;
    col2cp(g, col, w, p);
    EList<4> L;
    curl_col(g, which, w, p, L);
    if (!FILL) { cnt[col] = L.n; return; }
    int64_t o = colptr[col];
    for (int x = 0; x < L.n; ++x) { rowval[o + x] = L.e[x].row + 1; nzval[o + x] = L.e[x].v; }
}

template <bool FILL>
__global__ void k_mean(DGrid<double> g, int which, int64_t n, int64_t *cnt, const int64_t *colptr,
                       int64_t *rowval, c64 *nzval) {
    int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (col >= n) return;
    int w;
    int64_t p;
    col2cp(g, col, w, p);
    EList<2> L;
    mean_col(g, which, w, p, L);
    if (!FILL) { cnt[col] = L.n; return; }
    int64_t o = colptr[col];
    for (int x = 0; x < L.n; ++x) { rowval[o + x] = L.e[x].row + 1; nzval[o + x] = L.e[x].v; }
}

template <bool FILL>
__global__ void k_param(DGrid<double> g, int which, const c64 *parr, const int *full, int64_t n, int64_t *cnt,
                        const int64_t *colptr, int64_t *rowval, c64 *nzval) {
    int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (col >= n) return;
    int w;
    int64_t p;
    col2cp(g, col, w, p);
    EList<9> L;
    param_col(g, which, parr, *full, w, p, L);
    if (!FILL) { cnt[col] = L.n; return; }
    int64_t o = colptr[col];
    for (int x = 0; x < L.n; ++x) { rowval[o + x] = L.e[x].row + 1; nzval[o + x] = L.e[x].v; }
}

// Column of A = Cₘ (Pμ \ Cₑ) − ω² Pε (ft = EE, model.jl:235-237) or
// A = Cₑ (Pε \ Cₘ) − ω² Pμ (ft = HH, model.jl:238-240).  The product is
// accumulated in the order of Gustavson's algorithm (inner-factor column
// entries by row, then outer-factor column entries by row), then rows are
// sorted; the ω² term is subtracted with zero results dropped (skipped when
// ω = 0, as the reference does).  `dv` is the diagonal divisor array (μ for
// EE, ε for HH) and `pw` the parameter array of the ω² term.
template <bool FILL>
__global__ void k_A(DGrid<double> g, int ft, const c64 *pw, const c64 *dv, const int *full, c64 om2, int omz,
                    int64_t n, int64_t *cnt, const int64_t *colptr, int64_t *rowval, c64 *nzval) {
    int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (col >= n) return;
    const int inner = ft == EE ? 0 : 1;   // Cₑ first for EE, Cₘ first for HH
    int w;
    int64_t p;
    col2cp(g, col, w, p);
    EList<4> ce;
    curl_col(g, inner, w, p, ce);
    EList<16> acc;
    for (int k = 0; k < ce.n; ++k) {
        int u;
        int64_t r;
        col2cp(g, ce.e[k].row, u, r);
        c64 b = cdiv(ce.e[k].v, dv[r + g.M * (u + 3 * u)]);
        EList<4> cm;
        curl_col(g, 1 - inner, u, r, cm);
        for (int x = 0; x < cm.n; ++x) {
            c64 prod = cm.e[x].v * b;
            int f = -1;
            for (int y = 0; y < acc.n; ++y)
                if (acc.e[y].row == cm.e[x].row) { f = y; break; }
            if (f >= 0) acc.e[f].v = acc.e[f].v + prod;
            else acc.push(cm.e[x].row, acc.n, prod);
        }
    }
    acc.sort_merge();
    int64_t o = FILL ? colptr[col] : 0;
    int m = 0;
    if (omz) {
        for (int x = 0; x < acc.n; ++x) {
            if (FILL) { rowval[o + m] = acc.e[x].row + 1; nzval[o + m] = acc.e[x].v; }
            ++m;
        }
    } else {
        EList<9> pe;
        param_col(g, inner, pw, *full, w, p, pe);
        int a = 0, b = 0;
        while (a < acc.n || b < pe.n) {
            int64_t ra = a < acc.n ? acc.e[a].row : INT64_MAX, rb = b < pe.n ? pe.e[b].row : INT64_MAX;
            c64 v;
            int64_t r;
            if (ra == rb) { v = acc.e[a].v - om2 * pe.e[b].v; r = ra; ++a; ++b; }
            else if (ra < rb) { v = acc.e[a].v - mk<double>(0.0, 0.0); r = ra; ++a; }
            else { v = mk<double>(0.0, 0.0) - om2 * pe.e[b].v; r = rb; ++b; }
            if (!iszero(v)) {
                if (FILL) { rowval[o + m] = r + 1; nzval[o + m] = v; }
                ++m;
            }
        }
    }
    if (!FILL) cnt[col] = m;
}

// any nonzero off-diagonal entry in a (Nx,Ny,Nz,3,3) array
__global__ void k_offdiag(const c64 *parr, int64_t M, int *flag) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= 9 * M) return;
    int64_t vw = e / M;
    int v = (int)(vw % 3), w = (int)(vw / 3);
    if (v != w && !iszero(parr[e])) atomicOr(flag, 1);
}

__global__ void k_colptr1(int64_t *colptr, int64_t n) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e <= n) colptr[e] += 1;
}

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// count -> scan -> fill -> copy out (1-based) driver shared by all operators
template <typename CountF, typename FillF>
static int csc_drive(hipStream_t s, int64_t n, int64_t *nnz_out, int64_t *colptr_h, int64_t *rowval_h,
                     mfd_c64 *nzval_h, CountF count, FillF fill) {
    int64_t *cnt = nullptr, *colptr = nullptr, *rowval = nullptr;
    c64 *nzval = nullptr;
    void *tmp = nullptr;
    size_t tmpb = 0;
    int rc = MFD_OK;
    auto fail = [&](hipError_t e) {
        set_error(std::string("HIP error in CSC assembly: ") + hipGetErrorString(e));
        rc = MFD_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipMalloc(&cnt, sizeof(int64_t) * (n + 1))) != hipSuccess) { fail(e); goto done; }
    if ((e = hipMalloc(&colptr, sizeof(int64_t) * (n + 1))) != hipSuccess) { fail(e); goto done; }
    count(cnt);
    if ((e = hipGetLastError()) != hipSuccess) { fail(e); goto done; }
    hipcub::DeviceScan::ExclusiveSum(nullptr, tmpb, cnt, colptr, n + 1, s);
    if ((e = hipMalloc(&tmp, tmpb)) != hipSuccess) { fail(e); goto done; }
    if ((e = hipMemsetAsync(cnt + n, 0, sizeof(int64_t), s)) != hipSuccess) { fail(e); goto done; }
    hipcub::DeviceScan::ExclusiveSum(tmp, tmpb, cnt, colptr, n + 1, s);
    {
        int64_t nnz = 0;
        if ((e = hipMemcpyAsync(&nnz, colptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s)) != hipSuccess) { fail(e); goto done; }
        if ((e = hipStreamSynchronize(s)) != hipSuccess) { fail(e); goto done; }
        *nnz_out = nnz;
        if (!colptr_h) goto done;   // size query only
        if ((e = hipMalloc(&rowval, sizeof(int64_t) * (nnz ? nnz : 1))) != hipSuccess) { fail(e); goto done; }
        if ((e = hipMalloc(&nzval, sizeof(c64) * (nnz ? nnz : 1))) != hipSuccess) { fail(e); goto done; }
        fill(colptr, rowval, nzval);
        k_colptr1<<<nblk(n + 1, 256), 256, 0, s>>>(colptr, n);
        if ((e = hipGetLastError()) != hipSuccess) { fail(e); goto done; }
        if ((e = hipMemcpyAsync(colptr_h, colptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s)) != hipSuccess) { fail(e); goto done; }
        if (nnz) {
            if ((e = hipMemcpyAsync(rowval_h, rowval, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost, s)) != hipSuccess) { fail(e); goto done; }
            if ((e = hipMemcpyAsync(nzval_h, nzval, sizeof(c64) * nnz, hipMemcpyDeviceToHost, s)) != hipSuccess) { fail(e); goto done; }
        }
        if ((e = hipStreamSynchronize(s)) != hipSuccess) { fail(e); goto done; }
    }
done:
    hipStreamSynchronize(s);
    if (cnt) hipFree(cnt);
    if (colptr) hipFree(colptr);
    if (rowval) hipFree(rowval);
    if (nzval) hipFree(nzval);
    if (tmp) hipFree(tmp);
    return rc;
}

int curl_csc(mfd_ctx *ctx, const DevGrid &G, int which, int64_t *nnz, int64_t *colptr, int64_t *rowval,
             mfd_c64 *nzval) {
    const DGrid<double> &g = G.d64;
    int64_t n = 3 * g.M;
    hipStream_t s = ctx->stream;
    return csc_drive(s, n, nnz, colptr, rowval, nzval,
                     [&](int64_t *cnt) { k_curl<false><<<nblk(n, 256), 256, 0, s>>>(g, which, n, cnt, nullptr, nullptr, nullptr); },
                     [&](int64_t *cp, int64_t *rv, c64 *nz) { k_curl<true><<<nblk(n, 256), 256, 0, s>>>(g, which, n, nullptr, cp, rv, nz); });
}

int mean_csc(mfd_ctx *ctx, const DevGrid &G, int which, int64_t *nnz, int64_t *colptr, int64_t *rowval,
             mfd_c64 *nzval) {
    const DGrid<double> &g = G.d64;
    int64_t n = 3 * g.M;
    hipStream_t s = ctx->stream;
    return csc_drive(s, n, nnz, colptr, rowval, nzval,
                     [&](int64_t *cnt) { k_mean<false><<<nblk(n, 256), 256, 0, s>>>(g, which, n, cnt, nullptr, nullptr, nullptr); },
                     [&](int64_t *cp, int64_t *rv, c64 *nz) { k_mean<true><<<nblk(n, 256), 256, 0, s>>>(g, which, n, nullptr, cp, rv, nz); });
}

DevArray::~DevArray() { if (p) hipFree(p); }

int upload_param(mfd_ctx *ctx, const DevGrid &G, const mfd_c64 *parr, DevArray &d, int *full_h) {
    int64_t M = G.d64.M;
    size_t bytes = sizeof(c64) * 9 * M + 64;
    MFD_HIP(hipMalloc(&d.p, bytes));
    MFD_HIP(hipMemcpyAsync(d.p, parr, sizeof(c64) * 9 * M, hipMemcpyHostToDevice, ctx->stream));
    int *flag = (int *)((char *)d.p + sizeof(c64) * 9 * M);
    MFD_HIP(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
    k_offdiag<<<nblk(9 * M, 256), 256, 0, ctx->stream>>>((const c64 *)d.p, M, flag);
    MFD_HIP(hipGetLastError());
    if (full_h) {
        MFD_HIP(hipMemcpyAsync(full_h, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        MFD_HIP(hipStreamSynchronize(ctx->stream));
    }
    return MFD_OK;
}

int param_csc(mfd_ctx *ctx, const DevGrid &G, int which, const mfd_c64 *parr, int64_t *nnz, int64_t *colptr,
              int64_t *rowval, mfd_c64 *nzval) {
    const DGrid<double> &g = G.d64;
    int64_t n = 3 * g.M;
    hipStream_t s = ctx->stream;
    DevArray P;
    MFD_TRY(upload_param(ctx, G, parr, P, nullptr));
    const c64 *pa = (const c64 *)P.p;
    const int *full = (const int *)((const char *)P.p + sizeof(c64) * 9 * g.M);
    return csc_drive(s, n, nnz, colptr, rowval, nzval,
                     [&](int64_t *cnt) { k_param<false><<<nblk(n, 256), 256, 0, s>>>(g, which, pa, full, n, cnt, nullptr, nullptr, nullptr); },
                     [&](int64_t *cp, int64_t *rv, c64 *nz) { k_param<true><<<nblk(n, 256), 256, 0, s>>>(g, which, pa, full, n, nullptr, cp, rv, nz); });
}

int A_csc(mfd_ctx *ctx, const DevGrid &G, int ft, mfd_c64 omega, const mfd_c64 *eps, const mfd_c64 *mu,
          int64_t *nnz, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval) {
    const DGrid<double> &g = G.d64;
    int64_t n = 3 * g.M;
    hipStream_t s = ctx->stream;
    DevArray Dv, Pw;   // divisor (diagonal) and ω²-term parameter arrays
    int dfull = 0;
    MFD_TRY(upload_param(ctx, G, ft == EE ? mu : eps, Dv, &dfull));
    if (dfull) {
        set_error(ft == EE ? "create_A: Pμ must be diagonal (Pμ \\ Cₑ), model.jl:236"
                           : "create_A: Pε must be diagonal (Pε \\ Cₘ), model.jl:239");
        return MFD_ERR_UNSUPPORTED;
    }
    MFD_TRY(upload_param(ctx, G, ft == EE ? eps : mu, Pw, nullptr));
    const c64 *pa = (const c64 *)Pw.p, *da = (const c64 *)Dv.p;
    const int *full = (const int *)((const char *)Pw.p + sizeof(c64) * 9 * g.M);
    c64 om = mk<double>(omega.re, omega.im);
    c64 om2 = om * om;
    int omz = (omega.re == 0.0 && omega.im == 0.0);
    return csc_drive(s, n, nnz, colptr, rowval, nzval,
                     [&](int64_t *cnt) { k_A<false><<<nblk(n, 128), 128, 0, s>>>(g, ft, pa, da, full, om2, omz, n, cnt, nullptr, nullptr, nullptr); },
                     [&](int64_t *cp, int64_t *rv, c64 *nz) { k_A<true><<<nblk(n, 128), 128, 0, s>>>(g, ft, pa, da, full, om2, omz, n, nullptr, cp, rv, nz); });
}

}  // namespace mfd
