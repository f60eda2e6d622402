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


// ------------------------------------------ products of caller-given CSC matrices
//
// create_A / create_b / h_from_e / e_from_h of the reference (model.jl:222-284)
// take the assembled matrices (Pε, Pμ, Cₑ, Cₘ).  These entry points take them
// as Julia hands them over (SparseMatrixCSC{ComplexF64,Int64}: 1-based colptr
// and rowval, rows ascending within a column) and form the same results on the
// device with Julia SparseArrays' structure rules: the divided matrix must be
// diagonal (`Pμ \ Cₑ` scales the rows of Cₑ), the product keeps every
// structural entry (Gustavson's column-by-column product, contributions to an
// entry added in ascending inner index), and `A - ω² Pε` keeps only nonzero
// results, which Julia's sparse map does; with ω = 0 the subtraction is
// skipped.

DevCsc::~DevCsc() { if (buf) hipFree(buf); }

// per column: colptr nondecreasing, rows in [1, m] and strictly ascending
__global__ void k_csc_check(int64_t m, int64_t n, const int64_t *cp, const int64_t *rv, int *flag) {
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t a = cp[j] - 1, b = cp[j + 1] - 1;
    if (b < a) { atomicOr(flag, 1); return; }
    int64_t prev = 0;
    for (int64_t k = a; k < b; ++k) {
        const int64_t r = rv[k];
        if (r < 1 || r > m || r <= prev) { atomicOr(flag, 1); return; }
        prev = r;
    }
}

int csc_upload(mfd_ctx *ctx, const mfd_csc *h, const char *name, DevCsc &d) {
    if (!h || !h->colptr || (h->nnz > 0 && (!h->rowval || !h->nzval)) || h->m < 0 || h->n < 0 || h->nnz < 0) {
        set_error(std::string(name) + ": null or negative CSC fields");
        return MFD_ERR_ARG;
    }
    if (h->colptr[0] != 1 || h->colptr[h->n] != h->nnz + 1) {
        set_error(std::string(name) + ": colptr must start at 1 and end at nnz + 1 (1-based, Julia SparseMatrixCSC)");
        return MFD_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    const size_t bcp = sizeof(int64_t) * (h->n + 1), brv = sizeof(int64_t) * h->nnz, bnz = sizeof(c64) * h->nnz;
    const size_t off_rv = (bcp + 255) / 256 * 256, off_nz = off_rv + (brv + 255) / 256 * 256;
    const size_t off_fl = off_nz + (bnz + 255) / 256 * 256;
    MFD_HIP(hipMalloc(&d.buf, off_fl + 256));
    d.m = h->m;
    d.n = h->n;
    d.nnz = h->nnz;
    d.colptr = (int64_t *)d.buf;
    d.rowval = (int64_t *)((char *)d.buf + off_rv);
    d.nzval = (c64 *)((char *)d.buf + off_nz);
    int *flag = (int *)((char *)d.buf + off_fl);
    MFD_HIP(hipMemcpyAsync(d.colptr, h->colptr, bcp, hipMemcpyHostToDevice, s));
    if (h->nnz) {
        MFD_HIP(hipMemcpyAsync(d.rowval, h->rowval, brv, hipMemcpyHostToDevice, s));
        MFD_HIP(hipMemcpyAsync(d.nzval, h->nzval, bnz, hipMemcpyHostToDevice, s));
    }
    MFD_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
    if (h->n) k_csc_check<<<nblk(h->n, 256), 256, 0, s>>>(h->m, h->n, d.colptr, d.rowval, flag);
    MFD_HIP(hipGetLastError());
    int bad = 0;
    MFD_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    MFD_HIP(hipStreamSynchronize(s));
    if (bad) {
        set_error(std::string(name) + ": colptr must not decrease and rows must be in 1..m, ascending within a column");
        return MFD_ERR_ARG;
    }
    return MFD_OK;
}

// diagonal of a CSC matrix that must be diagonal (the divisor of `P \ C`):
// flag bit 0 = an entry off the diagonal, bit 1 = a zero or missing diagonal entry
__global__ void k_csc_diag(DevCscV P, c64 *d, int *flag) {
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= P.n) return;
    c64 v = mk<double>(0, 0);
    for (int64_t k = P.colptr[j] - 1; k < P.colptr[j + 1] - 1; ++k) {
        if (P.rowval[k] - 1 != j) atomicOr(flag, 1);
        else v = P.nzval[k];
    }
    if (iszero(v)) atomicOr(flag, 2);
    d[j] = v;
}

constexpr int SPGEMM_MAXC = 64;   // distinct rows per product column handled by one thread

// Column j of R = X·(D⁻¹Y), then R − s2·S when sub (numerical zeros dropped);
// count pass (FILL false) or fill pass.  flag bit 2: a column needed more than
// SPGEMM_MAXC distinct rows.
template <bool FILL>
__global__ __launch_bounds__(128) void k_spgemm(DevCscV X, DevCscV Y, const c64 *__restrict__ dinv_div,
                                                DevCscV S, c64 s2, int sub, int64_t n, int64_t *cnt,
                                                const int64_t *colptr, int64_t *rowval, c64 *nzval, int *flag) {
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    int64_t rows[SPGEMM_MAXC];
    c64 acc[SPGEMM_MAXC];
    int nc = 0;
    for (int64_t kb = Y.colptr[j] - 1; kb < Y.colptr[j + 1] - 1; ++kb) {
        const int64_t k = Y.rowval[kb] - 1;
        const c64 bv = cdiv(Y.nzval[kb], dinv_div[k]);   // (P \ C)[k, j]: row k of C divided by P[k, k]
        for (int64_t ka = X.colptr[k] - 1; ka < X.colptr[k + 1] - 1; ++ka) {
            const int64_t i = X.rowval[ka] - 1;
            const c64 v = X.nzval[ka] * bv;
            int x = 0;
            while (x < nc && rows[x] != i) ++x;
            if (x < nc) acc[x] = acc[x] + v;
            else if (nc < SPGEMM_MAXC) { rows[nc] = i; acc[nc] = v; ++nc; }
            else {   // refused by the caller: the column contributes no entries
                atomicOr(flag, 4);
                if (!FILL) cnt[j] = 0;
                return;
            }
        }
    }
    for (int x = 1; x < nc; ++x) {   // rows ascending
        const int64_t r = rows[x];
        const c64 a = acc[x];
        int y = x - 1;
        while (y >= 0 && rows[y] > r) { rows[y + 1] = rows[y]; acc[y + 1] = acc[y]; --y; }
        rows[y + 1] = r;
        acc[y + 1] = a;
    }
    int64_t o = FILL ? colptr[j] : 0, m = 0;
    if (!sub) {
        for (int x = 0; x < nc; ++x) {
            if (FILL) { rowval[o + m] = rows[x] + 1; nzval[o + m] = acc[x]; }
            ++m;
        }
    } else {
        int x = 0;
        int64_t b = S.colptr[j] - 1;
        const int64_t be = S.colptr[j + 1] - 1;
        while (x < nc || b < be) {
            const int64_t ra = x < nc ? rows[x] : INT64_MAX, rb = b < be ? S.rowval[b] - 1 : INT64_MAX;
            c64 v;
            int64_t r;
            if (ra == rb) { v = acc[x] - s2 * S.nzval[b]; r = ra; ++x; ++b; }
            else if (ra < rb) { v = acc[x]; r = ra; ++x; }
            else { v = mk<double>(0, 0) - s2 * S.nzval[b]; r = rb; ++b; }
            if (!iszero(v)) {
                if (FILL) { rowval[o + m] = r + 1; nzval[o + m] = v; }
                ++m;
            }
        }
    }
    if (!FILL) cnt[j] = m;
}

// ---- transpose to row-major (CSR) with rows' columns ascending

__global__ void k_rowcount(DevCscV A, int64_t *rc) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < A.nnz) atomicAdd((unsigned long long *)&rc[A.rowval[e] - 1], 1ull);
}
__global__ void k_csr_scatter(DevCscV A, const int64_t *rowptr, int64_t *fillc, int32_t *colind, c64 *val) {
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    for (int64_t k = A.colptr[j] - 1; k < A.colptr[j + 1] - 1; ++k) {
        const int64_t i = A.rowval[k] - 1;
        const int64_t pos = rowptr[i] + (int64_t)atomicAdd((unsigned long long *)&fillc[i], 1ull);
        colind[pos] = (int32_t)j;
        val[pos] = A.nzval[k];
    }
}
__global__ void k_csr_sortrows(int64_t m, const int64_t *rowptr, int32_t *colind, c64 *val) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t a = rowptr[i], b = rowptr[i + 1];
    for (int64_t x = a + 1; x < b; ++x) {
        const int32_t c = colind[x];
        const c64 v = val[x];
        int64_t y = x - 1;
        while (y >= a && colind[y] > c) { colind[y + 1] = colind[y]; val[y + 1] = val[y]; --y; }
        colind[y + 1] = c;
        val[y + 1] = v;
    }
}

DevCsr::~DevCsr() { if (buf) hipFree(buf); }

int csr_of(mfd_ctx *ctx, const DevCsc &A, DevCsr &R) {
    if (A.n > INT32_MAX) { set_error("matrix too wide for 32-bit column indices"); return MFD_ERR_ARG; }
    hipStream_t s = ctx->stream;
    const int64_t m = A.m, nnz = A.nnz;
    const size_t brp = sizeof(int64_t) * (m + 1), bci = sizeof(int32_t) * nnz, bv = sizeof(c64) * nnz;
    const size_t off_ci = (brp + 255) / 256 * 256, off_v = off_ci + (bci + 255) / 256 * 256;
    const size_t off_c = off_v + (bv + 255) / 256 * 256;
    MFD_HIP(hipMalloc(&R.buf, off_c + brp + 256));
    R.m = m;
    R.n = A.n;
    R.nnz = nnz;
    R.rowptr = (int64_t *)R.buf;
    R.colind = (int32_t *)((char *)R.buf + off_ci);
    R.val = (c64 *)((char *)R.buf + off_v);
    int64_t *rc = (int64_t *)((char *)R.buf + off_c);   // row counts, then fill counters
    MFD_HIP(hipMemsetAsync(rc, 0, brp, s));
    if (nnz) k_rowcount<<<nblk(nnz, 256), 256, 0, s>>>(A.view(), rc);
    size_t tmpb = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, tmpb, rc, R.rowptr, m + 1, s);
    void *tmp = nullptr;
    MFD_HIP(hipMalloc(&tmp, tmpb ? tmpb : 1));
    hipcub::DeviceScan::ExclusiveSum(tmp, tmpb, rc, R.rowptr, m + 1, s);
    hipError_t e = hipStreamSynchronize(s);
    hipFree(tmp);
    MFD_HIP(e);
    MFD_HIP(hipMemsetAsync(rc, 0, brp, s));
    if (A.n) k_csr_scatter<<<nblk(A.n, 256), 256, 0, s>>>(A.view(), R.rowptr, rc, R.colind, R.val);
    if (m) k_csr_sortrows<<<nblk(m, 256), 256, 0, s>>>(m, R.rowptr, R.colind, R.val);
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipStreamSynchronize(s));
    return MFD_OK;
}

// y = A v, one thread per row, columns ascending (the order of a column-by-
// column CSC product into y); then y = fy·(y + add) / div elementwise, each
// step only where its operand is given (fy applied last)
__global__ void k_csr_mv(int64_t m, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
                         const c64 *__restrict__ val, const c64 *__restrict__ v, const c64 *__restrict__ add,
                         c64 addsign, const c64 *__restrict__ div, c64 fy, int scale, c64 *__restrict__ y) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    c64 s = mk<double>(0, 0);
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s = s + val[k] * v[colind[k]];
    if (add) s = s + addsign * add[i];
    if (div) s = cdiv(s, div[i]);
    if (scale) s = fy * s;
    y[i] = s;
}

__global__ void k_divdiag(int64_t n, const c64 *a, const c64 *d, c64 *out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = cdiv(a[i], d[i]);
}

// b .*= -1 (when neg); b .-= iω .* j unless ω = 0
__global__ void k_b_finish(int64_t n, c64 *b, int neg, c64 iw, int omz, const c64 *j) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    c64 v = b[i];
    if (neg) v = v * mk<double>(-1.0, 0.0);
    if (!omz) v = v - iw * j[i];
    b[i] = v;
}

static int square_n(const mfd_csc *const *ms, const char *const *names, int cnt, int64_t *n) {
    *n = -1;
    for (int k = 0; k < cnt; ++k) {
        if (!ms[k]) { set_error(std::string(names[k]) + ": null matrix"); return MFD_ERR_ARG; }
        if (ms[k]->m != ms[k]->n || (*n >= 0 && ms[k]->n != *n)) {
            set_error("the matrices must be square and of one size n (3 unknowns per cell)");
            return MFD_ERR_ARG;
        }
        *n = ms[k]->n;
    }
    return MFD_OK;
}

// diagonal of the divisor `P` of `P \ C`, refused unless diagonal and nonsingular
static int divisor_diag(mfd_ctx *ctx, const DevCsc &P, const char *what, DevArray &d) {
    hipStream_t s = ctx->stream;
    MFD_HIP(hipMalloc(&d.p, sizeof(c64) * P.n + 256));
    int *flag = (int *)((char *)d.p + sizeof(c64) * P.n);
    MFD_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
    if (P.n) k_csc_diag<<<nblk(P.n, 256), 256, 0, s>>>(P.view(), (c64 *)d.p, flag);
    int f = 0;
    MFD_HIP(hipMemcpyAsync(&f, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    MFD_HIP(hipStreamSynchronize(s));
    if (f & 1) {
        set_error(std::string(what) + " must be diagonal (division unsupported otherwise, model.jl:236, 239)");
        return MFD_ERR_UNSUPPORTED;
    }
    if (f & 2) { set_error(std::string(what) + " has a zero or missing diagonal entry (singular)"); return MFD_ERR_ARG; }
    return MFD_OK;
}

int A_from_csc(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
               const mfd_csc *Cm, int64_t *nnz, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval) {
    if (ft != EE && ft != HH) { set_error("create_A: ft must be EE or HH (model.jl:241-245)"); return MFD_ERR_UNSUPPORTED; }
    const mfd_csc *ms[4] = {Pe, Pm, Ce, Cm};
    const char *names[4] = {"Pε", "Pμ", "Cₑ", "Cₘ"};
    int64_t n = 0;
    MFD_TRY(square_n(ms, names, 4, &n));
    // EE: A = Cₘ (Pμ \ Cₑ) − ω² Pε;  HH: A = Cₑ (Pε \ Cₘ) − ω² Pμ
    const mfd_csc *X = ft == EE ? Cm : Ce, *Y = ft == EE ? Ce : Cm, *D = ft == EE ? Pm : Pe, *S = ft == EE ? Pe : Pm;
    DevCsc dX, dY, dD, dS;
    MFD_TRY(csc_upload(ctx, X, ft == EE ? "Cₘ" : "Cₑ", dX));
    MFD_TRY(csc_upload(ctx, Y, ft == EE ? "Cₑ" : "Cₘ", dY));
    MFD_TRY(csc_upload(ctx, D, ft == EE ? "Pμ" : "Pε", dD));
    MFD_TRY(csc_upload(ctx, S, ft == EE ? "Pε" : "Pμ", dS));
    DevArray dd;
    MFD_TRY(divisor_diag(ctx, dD, ft == EE ? "Pμ" : "Pε", dd));
    hipStream_t s = ctx->stream;
    const c64 om = mk<double>(omega.re, omega.im), s2 = om * om;
    const int sub = !(omega.re == 0.0 && omega.im == 0.0);
    int *flag = nullptr;
    MFD_HIP(hipMalloc(&flag, sizeof(int)));
    DevArray fguard;
    fguard.p = flag;
    MFD_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
    const c64 *dv = (const c64 *)dd.p;
    int rc = csc_drive(s, n, nnz, colptr, rowval, nzval,
                       [&](int64_t *cnt) { k_spgemm<false><<<nblk(n, 128), 128, 0, s>>>(dX.view(), dY.view(), dv, dS.view(), s2, sub, n, cnt, nullptr, nullptr, nullptr, flag); },
                       [&](int64_t *cp, int64_t *rv, c64 *nz) { k_spgemm<true><<<nblk(n, 128), 128, 0, s>>>(dX.view(), dY.view(), dv, dS.view(), s2, sub, n, nullptr, cp, rv, nz, flag); });
    int f = 0;
    MFD_HIP(hipMemcpyAsync(&f, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    MFD_HIP(hipStreamSynchronize(s));
    if (f & 4) {
        set_error("create_A: a column of the product has more than 64 distinct rows (not a curl-curl operator)");
        return MFD_ERR_UNSUPPORTED;
    }
    return rc;
}

int b_from_csc(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
               const mfd_csc *Cm, const mfd_c64 *je, const mfd_c64 *jm, mfd_c64 *b) {
    if (ft != EE && ft != HH) { set_error("create_b: ft must be EE or HH (model.jl:269-271)"); return MFD_ERR_UNSUPPORTED; }
    if (!je || !jm || !b) { set_error("create_b: null jₑ, jₘ or b"); return MFD_ERR_ARG; }
    const mfd_csc *ms[4] = {Pe, Pm, Ce, Cm};
    const char *names[4] = {"Pε", "Pμ", "Cₑ", "Cₘ"};
    int64_t n = 0;
    MFD_TRY(square_n(ms, names, 4, &n));
    // EE: b = −Cₘ (Pμ \ jₘ) − iω jₑ;  HH: b = Cₑ (Pε \ jₑ) − iω jₘ
    const mfd_csc *X = ft == EE ? Cm : Ce, *D = ft == EE ? Pm : Pe;
    const mfd_c64 *jd = ft == EE ? jm : je, *jw = ft == EE ? je : jm;
    DevCsc dX, dD;
    MFD_TRY(csc_upload(ctx, X, ft == EE ? "Cₘ" : "Cₑ", dX));
    MFD_TRY(csc_upload(ctx, D, ft == EE ? "Pμ" : "Pε", dD));
    DevArray dd;
    MFD_TRY(divisor_diag(ctx, dD, ft == EE ? "Pμ" : "Pε", dd));
    DevCsr R;
    MFD_TRY(csr_of(ctx, dX, R));
    hipStream_t s = ctx->stream;
    DevArray V;
    MFD_HIP(hipMalloc(&V.p, sizeof(c64) * 3 * n + 64));
    c64 *y = (c64 *)V.p, *jdd = y + n, *jwd = y + 2 * n;
    MFD_HIP(hipMemcpyAsync(jdd, jd, sizeof(c64) * n, hipMemcpyHostToDevice, s));
    MFD_HIP(hipMemcpyAsync(jwd, jw, sizeof(c64) * n, hipMemcpyHostToDevice, s));
    const c64 om = mk<double>(omega.re, omega.im), iw = om * mk<double>(0.0, 1.0);
    const int omz = omega.re == 0.0 && omega.im == 0.0;
    if (n) {
        k_divdiag<<<nblk(n, 256), 256, 0, s>>>(n, jdd, (const c64 *)dd.p, jdd);
        k_csr_mv<<<nblk(n, 256), 256, 0, s>>>(n, R.rowptr, R.colind, R.val, jdd, nullptr, mk<double>(1, 0), nullptr,
                                               mk<double>(1, 0), 0, y);
        k_b_finish<<<nblk(n, 256), 256, 0, s>>>(n, y, ft == EE, iw, omz, jwd);
    }
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipMemcpyAsync(b, y, sizeof(c64) * n, hipMemcpyDeviceToHost, s));
    MFD_HIP(hipStreamSynchronize(s));
    return MFD_OK;
}

int field_from_csc(mfd_ctx *ctx, int which, mfd_c64 omega, const mfd_csc *P, const mfd_csc *Cx,
                   const mfd_c64 *in, const mfd_c64 *src, mfd_c64 *out) {
    const char *pn = which == 0 ? "Pμ" : "Pε", *cn = which == 0 ? "Cₑ" : "Cₘ";
    if (!in || !out) { set_error("null field vector"); return MFD_ERR_ARG; }
    if (omega.re == 0.0 && omega.im == 0.0) { set_error("h_from_e/e_from_h: ω must be nonzero"); return MFD_ERR_ARG; }
    const mfd_csc *ms[2] = {P, Cx};
    const char *names[2] = {pn, cn};
    int64_t n = 0;
    MFD_TRY(square_n(ms, names, 2, &n));
    DevCsc dP, dC;
    MFD_TRY(csc_upload(ctx, P, pn, dP));
    MFD_TRY(csc_upload(ctx, Cx, cn, dC));
    DevArray dd;
    MFD_TRY(divisor_diag(ctx, dP, pn, dd));
    DevCsr R;
    MFD_TRY(csr_of(ctx, dC, R));
    hipStream_t s = ctx->stream;
    DevArray V;
    MFD_HIP(hipMalloc(&V.p, sizeof(c64) * 3 * n + 64));
    c64 *vin = (c64 *)V.p, *vsrc = vin + n, *vout = vin + 2 * n;
    MFD_HIP(hipMemcpyAsync(vin, in, sizeof(c64) * n, hipMemcpyHostToDevice, s));
    if (src) MFD_HIP(hipMemcpyAsync(vsrc, src, sizeof(c64) * n, hipMemcpyHostToDevice, s));
    // h = (i/ω) (Pμ \ (Cₑ e + jₘ));  e = (−i/ω) (Pε \ (Cₘ h − jₑ))
    const c64 om = mk<double>(omega.re, omega.im);
    const c64 f = cdiv(mk<double>(0.0, which == 0 ? 1.0 : -1.0), om);
    if (n)
        k_csr_mv<<<nblk(n, 256), 256, 0, s>>>(n, R.rowptr, R.colind, R.val, vin, src ? vsrc : nullptr,
                                               mk<double>(which == 0 ? 1.0 : -1.0, 0), (const c64 *)dd.p, f, 1, vout);
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipMemcpyAsync(out, vout, sizeof(c64) * n, hipMemcpyDeviceToHost, s));
    MFD_HIP(hipStreamSynchronize(s));
    return MFD_OK;
}

}  // namespace mfd
