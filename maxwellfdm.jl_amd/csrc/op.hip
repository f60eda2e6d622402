// Matrix-free curl-curl operator and device-resident symmetric QMR solver.
//
// A = Cₘ μ⁻¹ Cₑ − ω² Pε (create_A, reference src/model/model.jl:235-237) is
// applied without forming the CSC.  The solver (the user's `A \ b`, which the
// reference leaves to Julia) works on the complex-symmetric
//     P = Q T A T⁻¹ Q,   T = S_E^{1/2} (separable), Q = diag(A)^{-1/2},
// where S_E is the stretched E-cell volume; Aᵀ = S_E A S_E⁻¹ makes P = Pᵀ.
// P needs only Q per unknown: its diagonal mass term −ω² ε q² equals
// 1 − c q² with c the (separable) curl-curl diagonal, so the operator reads
// the iterate, Q and (when ε is anisotropic) the corner ε — never ε itself.
//
// Boundary rules (all boundft = EE): Cₑ forward differences with the ghost at
// the positive end, Cₘ backward differences with the ghost at the negative
// end; a Bloch ghost is the wrapped value times e^{∓ikL}, a non-Bloch ghost
// is zero.
#include "op.hpp"
#include "assembly.hpp"
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>

namespace mfd {

constexpr int NPART = 16384;  // max blocks per reduction (per RHS)
constexpr int NGRID_MAX = 2048;   // grid cap of the grid-stride vector kernels
constexpr int BS = 256;
enum { MODE_A = 0, MODE_P = 1 };
enum { EPI_STORE = 0, EPI_RESID = 1, EPI_DOT = 2 };

// --------------------------------------------------------------- helpers

template <typename R>
__device__ inline cx<R> czero() { return mk<R>((R)0, (R)0); }
template <typename R>
__device__ inline cx<R> cone() { return mk<R>((R)1, (R)0); }

template <typename R>
__device__ inline void cell3(const DGrid<R> &g, int64_t p, int &i, int &j, int &k) {
    i = (int)(p % g.N[0]);
    int64_t t = p / g.N[0];
    j = (int)(t % g.N[1]);
    k = (int)(t / g.N[1]);
}
template <typename R>
__device__ inline void dof2cp(const DGrid<R> &g, int64_t e, int &c, int64_t &p) {
    if (g.cmpfirst) { c = (int)(e % 3); p = e / 3; }
    else { const int k = (int)(e / g.M); c = g.c0 + k; p = e - (int64_t)k * g.M; }
}
// T (or T⁻¹) for E_c at (i,j,k): product over axes of sqrt(s∆l) at E_c's grid
// type (which each axis's boundft sets)
template <typename R>
__device__ inline cx<R> tfac(const DGrid<R> &g, int c, int i, int j, int k, bool inv) {
    if (g.tdof) {
        const int64_t p = i + (int64_t)g.N[0] * (j + (int64_t)g.N[1] * k);
        return inv ? g.tdofinv[3 * p + c] : g.tdof[3 * p + c];
    }
    const cx<R> *const(*t)[3] = inv ? g.ssqinv : g.ssq;
    cx<R> a = t[gtype(g.bft, EE, c, 0)][0][i];
    cx<R> b = t[gtype(g.bft, EE, c, 1)][1][j];
    cx<R> d = t[gtype(g.bft, EE, c, 2)][2][k];
    return (a * b) * d;
}

// scaled iterate value at E_c(i,j,k); indices may be -1 or N on one or more
// axes (ghosts): Bloch wraps with the phase, non-Bloch gives zero
template <typename R, int MODE>
__device__ inline cx<R> ldX(const DGrid<R> &g, const OpData<R> &od, const cx<R> *__restrict__ x, int c, int i,
                            int j, int k) {
    cx<R> ph = cone<R>();
    bool wr = false;
    int id[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (id[a] >= g.N[a]) {
            if (!g.bloch[a]) return czero<R>();
            id[a] -= g.N[a]; ph = ph * g.ph[a]; wr = true;
        } else if (id[a] < 0) {
            if (!g.bloch[a]) return czero<R>();
            id[a] += g.N[a]; ph = ph * g.phinv[a]; wr = true;
        }
    }
    int64_t p = id[0] + (int64_t)g.N[0] * (id[1] + (int64_t)g.N[1] * id[2]);
    cx<R> v = x[dofidx(g, c, p)];
    if (MODE == MODE_P)   // the stored q·T⁻¹ where the operator has it (FP64 data)
        v = v * (od.qs ? od.qs[3 * p + c] : od.q[3 * p + c] * tfac(g, c, id[0], id[1], id[2], true));
    return wr ? v * ph : v;
}

// H_u at (i,j,k) = ∂_v X_w − ∂_w X_v (v = u+1, w = u+2), times 1/μ_u.  Cₑ
// differences run forward along an axis whose boundft is EE and backward where
// it is HH; Cₘ the other way (so one index may be -1 or N: the outer curl's
// ghost, wrapped with the Bloch phase or zero).
template <typename R, int MODE, bool MU>
__device__ inline cx<R> Hval(const DGrid<R> &g, const OpData<R> &od, const cx<R> *__restrict__ x, int u, int i,
                             int j, int k) {
    cx<R> ph = cone<R>();
    bool wr = false;
    int id[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (id[a] < 0) {
            if (!g.bloch[a]) return czero<R>();
            id[a] += g.N[a]; ph = ph * g.phinv[a]; wr = true;
        } else if (id[a] >= g.N[a]) {
            if (!g.bloch[a]) return czero<R>();
            id[a] -= g.N[a]; ph = ph * g.ph[a]; wr = true;
        }
    }
    int v = u == 0 ? 1 : (u == 1 ? 2 : 0), w = u == 0 ? 2 : (u == 1 ? 0 : 1);
    // ∂_a X_c at id without the ∆l factor, in Cₑ's direction along a
    auto dif = [&](int a, int c) -> cx<R> {
        int i1[3] = {id[0], id[1], id[2]};
        if (g.bft[a] == EE) {
            i1[a] += 1;
            return ldX<R, MODE>(g, od, x, c, i1[0], i1[1], i1[2]) - ldX<R, MODE>(g, od, x, c, id[0], id[1], id[2]);
        }
        i1[a] -= 1;
        return ldX<R, MODE>(g, od, x, c, id[0], id[1], id[2]) - ldX<R, MODE>(g, od, x, c, i1[0], i1[1], i1[2]);
    };
    cx<R> h = g.sdlinv[g0type(g.bft, HH, v)][v][id[v]] * dif(v, w) - g.sdlinv[g0type(g.bft, HH, w)][w][id[w]] * dif(w, v);
    if (MU) {
        int64_t p = id[0] + (int64_t)g.N[0] * (id[1] + (int64_t)g.N[1] * id[2]);
        h = h * od.muinv[3 * p + u];
    }
    return wr ? h * ph : h;
}

// curl-curl diagonal for E_u at (i,j,k): Σ over the two derivative axes (Cₘ
// backward along an EE axis, taking H at p and p−1; forward along an HH axis,
// taking H at p and p+1; a non-Bloch ghost H is zero)
template <typename R, bool MU>
__device__ inline cx<R> ccdiag(const DGrid<R> &g, const OpData<R> &od, int u, int i, int j, int k) {
    int id[3] = {i, j, k};
    int64_t p = i + (int64_t)g.N[0] * (j + (int64_t)g.N[1] * k);
    cx<R> s = czero<R>();
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        int a = t == 0 ? (u == 0 ? 1 : (u == 1 ? 2 : 0)) : (u == 0 ? 2 : (u == 1 ? 0 : 1));
        int hc = t == 0 ? (u == 0 ? 2 : (u == 1 ? 0 : 1)) : (u == 0 ? 1 : (u == 1 ? 2 : 0));  // H component
        int ia = id[a], Na = g.N[a];
        const bool fwd = g.bft[a] == HH;   // Cₘ direction along a
        const int gh = g0type(g.bft, HH, a);
        int io = fwd ? (ia < Na - 1 ? ia + 1 : 0) : (ia > 0 ? ia - 1 : Na - 1);   // the other H index
        cx<R> m0 = cone<R>(), m1 = cone<R>();
        int64_t sta = a == 0 ? 1 : (a == 1 ? (int64_t)g.N[0] : (int64_t)g.N[0] * g.N[1]);
        if (MU) {
            m0 = od.muinv[3 * p + hc];
            m1 = od.muinv[3 * (p + (int64_t)(io - ia) * sta) + hc];
        }
        cx<R> ie = g.sdlinv[g0type(g.bft, EE, a)][a][ia];
        cx<R> term;
        if (Na == 1 && g.bloch[a]) {
            cx<R> f = mk<R>((R)2, (R)0) - g.ph[a] - g.phinv[a];
            term = (ie * (g.sdlinv[gh][a][0] * m0)) * f;
        } else if (!g.bloch[a] && (fwd ? ia == Na - 1 : ia == 0)) {
            term = ie * (g.sdlinv[gh][a][ia] * m0);
        } else {
            term = ie * (g.sdlinv[gh][a][ia] * m0 + g.sdlinv[gh][a][io] * m1);
        }
        s = s + term;
    }
    return s;
}

// block reduction of NV doubles; result valid in thread 0
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
    __shared__ double sh[8][NV];
#pragma unroll
    for (int a = 0; a < NV; ++a)
        for (int o = 32; o > 0; o >>= 1) v[a] += __shfl_down(v[a], o, 64);
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < NV; ++a) sh[wv][a] = v[a];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int a = 0; a < NV; ++a) {
            double s = 0;
            for (int b = 0; b < (int)(blockDim.x >> 6); ++b) s += sh[b][a];
            v[a] = s;
        }
    __syncthreads();
}

// write this block's partial (block bid of nb); returns true in every thread of
// the last block to arrive, which then holds the fixed-order total in tot
// (thread 0) and re-arms the counter
template <int NV>
__device__ inline bool reduce_last_n(double (&v)[NV], double *part, unsigned *cnt, double (&tot)[NV], int bid,
                                     int nb) {
    __shared__ int s_last;
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
        for (int a = 0; a < NV; ++a) part[4 * bid + a] = v[a];
        __threadfence();
        unsigned prev = atomicAdd(cnt, 1u);
        s_last = (prev == (unsigned)nb - 1);
    }
    __syncthreads();
    if (!s_last) return false;
    __threadfence();
    double acc[NV];
    for (int a = 0; a < NV; ++a) acc[a] = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int a = 0; a < NV; ++a) acc[a] += __hip_atomic_load(&part[4 * b + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    block_sum<NV>(acc);
    if (threadIdx.x == 0) {
        for (int a = 0; a < NV; ++a) tot[a] = acc[a];
        *cnt = 0;
    }
    return true;
}
template <int NV>
__device__ inline bool reduce_last(double (&v)[NV], double *part, unsigned *cnt, double (&tot)[NV]) {
    return reduce_last_n<NV>(v, part, cnt, tot, (int)blockIdx.x, (int)gridDim.x);
}

// Grid reductions on the solver's hot path: every block writes its partial
// sums with plain stores and a one-block finish kernel, launched next on the
// same stream, adds them in a fixed order (deterministic) and updates the
// solver scalars.  The kernel boundary orders the two, so no block needs a
// device-scope fence or a shared counter.
template <int NV>
__device__ inline void block_partial(double (&v)[NV], double *part, int bid) {
    block_sum<NV>(v);
    if (threadIdx.x == 0)
        for (int a = 0; a < NV; ++a) part[4 * bid + a] = v[a];
}
// fixed-order sum of nb partials by one block; result in thread 0
template <int NV>
__device__ inline void sum_partials(const double *part, int nb, double (&tot)[NV]) {
    for (int a = 0; a < NV; ++a) tot[a] = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int a = 0; a < NV; ++a) tot[a] += part[4 * b + a];
    block_sum<NV>(tot);
}

// after the stencil with EPI_DOT: σ = qᵀPq, α = ρ/σ (one block per RHS)
__global__ __launch_bounds__(BS) void k_fin_dot(const double *part, int nb, Scal *scal) {
    const int rhs = blockIdx.x;
    Scal &S = scal[rhs];
    if (S.done) return;
    double tot[2];
    sum_partials<2>(part + (size_t)rhs * NPART * 4, nb, tot);
    if (threadIdx.x != 0) return;
    if (S.conv) { S.done = 1; return; }
    S.sig_re = tot[0];
    S.sig_im = tot[1];
    if (tot[0] == 0.0 && tot[1] == 0.0) {   // breakdown: no step (the stencil already applied the pending update)
        S.breakdown = 1;
        S.conv = 1;
        S.alpha_re = S.alpha_im = 0;
        S.apend_re = S.apend_im = 0;
        return;
    }
    c64 a = cdiv(mk<double>(S.rho_re, S.rho_im), mk<double>(tot[0], tot[1]));
    S.alpha_re = a.re;
    S.alpha_im = a.im;
    S.apend_re = a.re;
    S.apend_im = a.im;
}

// after the stencil with EPI_RESID: ||b − A x||² per (unmasked) RHS
__global__ __launch_bounds__(BS) void k_fin_resid(const double *part, int nb, double *rn, const int *mask) {
    const int rhs = blockIdx.x;
    if (mask && !mask[rhs]) return;
    double tot[1];
    sum_partials<1>(part + (size_t)rhs * NPART * 4, nb, tot);
    if (threadIdx.x == 0) rn[rhs] = tot[0];
}

// ------------------------------------------------------------ the stencil

// y = A x (MODE_A) or y = P x (MODE_P) for the RHS blockIdx.y.
// EPI_STORE: y = op(x).  EPI_RESID: y = b − op(x), reduce ||y||² into rn.
// EPI_DOT: y = op(x), reduce xᵀy into the solver scalars (σ, α).
// H value from the scratch of k_hpass at (i,j,k) with one index possibly −1
// or N: the Bloch image times the phase, or zero past a non-Bloch end
template <typename R>
__device__ inline cx<R> ldHb(const DGrid<R> &g, const cx<R> *__restrict__ hb, int c, int i, int j, int k) {
    cx<R> ph = cone<R>();
    bool wr = false;
    int id[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (id[a] < 0) {
            if (!g.bloch[a]) return czero<R>();
            id[a] += g.N[a]; ph = ph * g.phinv[a]; wr = true;
        } else if (id[a] >= g.N[a]) {
            if (!g.bloch[a]) return czero<R>();
            id[a] -= g.N[a]; ph = ph * g.ph[a]; wr = true;
        }
    }
    cx<R> v = hb[dofidx(g, c, id[0] + (int64_t)g.N[0] * (id[1] + (int64_t)g.N[1] * id[2]))];
    return wr ? v * ph : v;
}

// first pass of the two-pass direct kernel: H = μ⁻¹ Cₑ X at every cell, once
template <typename R, int MODE, bool MU>
__global__ __launch_bounds__(BS) void k_hpass(DGrid<R> g, OpData<R> od, const cx<R> *__restrict__ xall,
                                              cx<R> *__restrict__ hall, int64_t vs, const Scal *scal, int epi_dot) {
    const int rhs = blockIdx.y;
    if (epi_dot && (scal[rhs].done || scal[rhs].conv)) return;
    const cx<R> *x = xall + rhs * vs;
    cx<R> *hb = hall + rhs * vs;
    for (int64_t p = blockIdx.x * (int64_t)BS + threadIdx.x; p < g.M; p += (int64_t)gridDim.x * BS) {
        int i, j, k;
        cell3(g, p, i, j, k);
#pragma unroll
        for (int c = 0; c < 3; ++c) hb[dofidx(g, c, p)] = Hval<R, MODE, MU>(g, od, x, c, i, j, k);
    }
}

// HB: H comes from k_hpass's scratch (hall) instead of being recomputed at p
// and its six neighbours
template <typename R, int MODE, bool OFFD, bool MU, int EPI, bool HB = false>
__global__ __launch_bounds__(BS) void k_apply(DGrid<R> g, OpData<R> od, const cx<R> *__restrict__ xall,
                                              cx<R> *__restrict__ yall, const c64 *__restrict__ ball, int64_t vs,
                                              double *part, const Scal *scal,
                                              const int *__restrict__ mask = nullptr,
                                              const cx<R> *__restrict__ hall = nullptr) {
    const int rhs = blockIdx.y;
    const cx<R> *x = xall + rhs * vs;
    cx<R> *y = yall + rhs * vs;
    const cx<R> *hb = HB ? hall + rhs * vs : nullptr;
    bool skip = false;
    if (EPI == EPI_RESID && mask && !mask[rhs]) return;
    if (EPI == EPI_DOT) {
        if (scal[rhs].done) return;
        skip = scal[rhs].conv != 0;
    }
    double acc[2] = {0.0, 0.0};
    if (!skip) {
        for (int64_t p = blockIdx.x * (int64_t)BS + threadIdx.x; p < g.M; p += (int64_t)gridDim.x * BS) {
            int i, j, k;
            cell3(g, p, i, j, k);
            // H_c(p) and Hm[c][a] = H_c(p ∓ â), the neighbour Cₘ's difference
            // along a takes (p − â along an EE axis, p + â along an HH axis)
            cx<R> H[3], Hm[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) H[c] = HB ? hb[dofidx(g, c, p)] : Hval<R, MODE, MU>(g, od, x, c, i, j, k);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (a == c) continue;
                    int id[3] = {i, j, k};
                    id[a] += g.bft[a] == EE ? -1 : 1;
                    Hm[c][a] = HB ? ldHb(g, hb, c, id[0], id[1], id[2])
                                  : Hval<R, MODE, MU>(g, od, x, c, id[0], id[1], id[2]);
                }
            int id[3] = {i, j, k};
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                int v = u == 0 ? 1 : (u == 1 ? 2 : 0), w = u == 0 ? 2 : (u == 1 ? 0 : 1);
                // E_u = ∂_v H_w − ∂_w H_v (backward along EE axes, forward along HH axes)
                cx<R> dv = g.bft[v] == EE ? H[w] - Hm[w][v] : Hm[w][v] - H[w];
                cx<R> dw = g.bft[w] == EE ? H[v] - Hm[v][w] : Hm[v][w] - H[v];
                cx<R> cc = g.sdlinv[g0type(g.bft, EE, v)][v][id[v]] * dv - g.sdlinv[g0type(g.bft, EE, w)][w][id[w]] * dw;
                int64_t e = dofidx(g, u, p);
                cx<R> xu = x[e];
                cx<R> out;
                cx<R> lf;
                if (MODE == MODE_A) {
                    out = cc - od.om2 * (od.epsd[3 * p + u] * xu);
                } else {
                    cx<R> qu = od.q[3 * p + u];
                    lf = qu * tfac(g, u, i, j, k, false);
                    cx<R> cd = ccdiag<R, MU>(g, od, u, i, j, k);
                    out = lf * cc + (cone<R>() - cd * (qu * qu)) * xu;
                }
                if (OFFD) {
                    cx<R> term = czero<R>();
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        int ww = t == 0 ? v : w;   // input component
                        int eix = (u + ww == 1) ? 0 : ((u + ww == 2) ? 1 : 2);   // (xy, xz, yz)
#pragma unroll
                        for (int oc = 0; oc < 2; ++oc) {
                            int cid[3] = {i, j, k};
                            cx<R> fo = mk<R>((R)0.5, (R)0);
                            if (oc == 1) {
                                cid[u] += 1;
                                if (cid[u] >= g.N[u]) {
                                    if (!g.bloch[u]) continue;
                                    cid[u] -= g.N[u];
                                    fo = (R)0.5 * g.ph[u];
                                }
                            }
                            int64_t cp = cid[0] + (int64_t)g.N[0] * (cid[1] + (int64_t)g.N[1] * cid[2]);
                            cx<R> ec = od.epso[3 * cp + eix];
                            int cw = cid[ww];
                            cx<R> ia = g.sdlinv[PRIM][ww][cw];
                            cx<R> wi0 = ((R)0.5 * g.sdl[DUAL][ww][cw]) * ia;
                            cx<R> s = wi0 * ldX<R, MODE>(g, od, x, ww, cid[0], cid[1], cid[2]);
                            int fid[3] = {cid[0], cid[1], cid[2]};
                            fid[ww] -= 1;
                            if (fid[ww] >= 0 || g.bloch[ww]) {
                                int fi = fid[ww] >= 0 ? fid[ww] : g.N[ww] - 1;
                                cx<R> wi1 = ((R)0.5 * g.sdl[DUAL][ww][fi]) * ia;
                                s = s + wi1 * ldX<R, MODE>(g, od, x, ww, fid[0], fid[1], fid[2]);
                            }
                            term = term + (fo * ec) * s;
                        }
                    }
                    if (MODE == MODE_A) out = out - od.om2 * term;
                    else out = out - od.om2 * (lf * term);
                }
                if (EPI == EPI_RESID) {
                    c64 bb = ball[rhs * vs + e];
                    cx<R> rr = mk<R>((R)bb.re - out.re, (R)bb.im - out.im);
                    if (yall) y[e] = rr;   // null output: the residual norm only
                    acc[0] += (double)rr.re * rr.re + (double)rr.im * rr.im;
                } else {
                    y[e] = out;
                    if (EPI == EPI_DOT) {
                        acc[0] += (double)xu.re * out.re - (double)xu.im * out.im;
                        acc[1] += (double)xu.re * out.im + (double)xu.im * out.re;
                    }
                }
            }
        }
    }
    if (EPI == EPI_STORE) return;
    block_partial<2>(acc, part + (size_t)rhs * NPART * 4, (int)blockIdx.x);
}

// Sparse product for an operator built from a caller's CSC matrix A (stored
// row-major): MODE_A y = A x, MODE_P y = Q T A T⁻¹ Q x with qin = Q T⁻¹ and
// qout = Q T per DOF (Q = diag(A)^-1/2, T the symmetrizing scaling); the
// epilogues are those of k_apply.
// LPR lanes share a row (its entries are contiguous), summed by shuffles.
template <typename R, int MODE, int EPI>
__global__ __launch_bounds__(BS) void k_spmv(int64_t m, const int64_t *__restrict__ rowptr,
                                             const int32_t *__restrict__ colind, const c64 *__restrict__ val,
                                             const c64 *__restrict__ qin, const c64 *__restrict__ qout,
                                             const cx<R> *__restrict__ xall,
                                             cx<R> *__restrict__ yall, const c64 *__restrict__ ball, int64_t vs,
                                             double *part, const Scal *scal, const int *__restrict__ mask) {
    constexpr int LPR = 4;
    const int rhs = blockIdx.y;
    const cx<R> *x = xall + rhs * vs;
    cx<R> *y = yall + rhs * vs;
    bool skip = false;
    if (EPI == EPI_RESID && mask && !mask[rhs]) return;
    if (EPI == EPI_DOT) {
        if (scal[rhs].done) return;
        skip = scal[rhs].conv != 0;
    }
    double acc[2] = {0.0, 0.0};
    const int lane = threadIdx.x % LPR;
    if (!skip) {
        for (int64_t i = (blockIdx.x * (int64_t)BS + threadIdx.x) / LPR; i < m; i += (int64_t)gridDim.x * (BS / LPR)) {
            c64 s = mk<double>(0, 0);
            for (int64_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += LPR) {
                const int32_t j = colind[k];
                c64 xv = cvt<double>(x[j]);
                if (MODE == MODE_P) xv = xv * qin[j];
                s = s + val[k] * xv;
            }
#pragma unroll
            for (int o = 1; o < LPR; o <<= 1) {
                s.re += __shfl_xor(s.re, o, LPR);
                s.im += __shfl_xor(s.im, o, LPR);
            }
            if (lane) continue;
            const c64 xi = cvt<double>(x[i]);
            const c64 out = MODE == MODE_P ? qout[i] * s : s;
            if (EPI == EPI_RESID) {
                const c64 bb = ball[rhs * vs + i];
                const c64 rr = mk<double>(bb.re - out.re, bb.im - out.im);
                if (yall) y[i] = cvt<R>(rr);   // null output: the residual norm only
                acc[0] += rr.re * rr.re + rr.im * rr.im;
            } else {
                y[i] = cvt<R>(out);
                if (EPI == EPI_DOT) {
                    acc[0] += xi.re * out.re - xi.im * out.im;
                    acc[1] += xi.re * out.im + xi.im * out.re;
                }
            }
        }
    }
    if (EPI == EPI_STORE) return;
    block_partial<2>(acc, part + (size_t)rhs * NPART * 4, (int)blockIdx.x);
}

// Jacobi factor q = diag(A)^-1/2 of a row-major A; flag set for a zero or
// missing diagonal entry
__global__ void k_csr_jacobi(int64_t m, const int64_t *rowptr, const int32_t *colind, const c64 *val, c64 *q,
                             int *flag) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    c64 d = mk<double>(0, 0);
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
        if (colind[k] == i) d = val[k];
    if (iszero(d)) { atomicOr(flag, 1); d = mk<double>(1, 0); }
    q[i] = cinv(csqrt_(d));
}

// Symmetrizing scaling of an assembled A: a diagonal S with S A symmetric,
// s_i A_ij = s_j A_ji, exists for the A of create_A (S = the E-cell volumes,
// since Cₘ = S⁻¹ Cₑᵀ S_H) but A alone does not carry it.  It is spread
// breadth-first from one DOF (s = 1): each sweep gives every unscaled DOF i
// the scale of its first scaled neighbour j (ascending j, A_ij and A_ji both
// nonzero), s_i = s_j A_ji / A_ij, reading only the previous sweep's scales
// (s = 0: not yet scaled).
__global__ void k_sym_sweep(int64_t n, const int64_t *rowptr, const int32_t *colind, const c64 *val,
                            const int64_t *colptr, const int64_t *rowval, const c64 *nzval, const c64 *s_old,
                            c64 *s_new, int *changed) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const c64 si = s_old[i];
    if (!iszero(si)) { s_new[i] = si; return; }
    c64 v = mk<double>(0, 0);
    // row i (A_ij, columns ascending) against column i (A_ji, rows ascending)
    int64_t a = rowptr[i], ae = rowptr[i + 1], b = colptr[i] - 1, be = colptr[i + 1] - 1;
    while (a < ae && b < be) {
        const int64_t ja = colind[a], jb = rowval[b] - 1;
        if (ja == jb) {
            const c64 sj = s_old[ja];
            if (!iszero(sj) && !iszero(val[a]) && !iszero(nzval[b])) { v = sj * cdiv(nzval[b], val[a]); break; }
            ++a;
            ++b;
        } else if (ja < jb) ++a;
        else ++b;
    }
    s_new[i] = v;
    if (!iszero(v)) *changed = 1;
}
// first DOF still unscaled (a new component) gets s = 1
__global__ void k_sym_seed(int64_t n, c64 *s, unsigned long long *first) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && iszero(s[i])) atomicMin(first, (unsigned long long)i);
}
__global__ void k_sym_set1(c64 *s, const unsigned long long *first, int64_t n) {
    if (threadIdx.x == 0 && *first < (unsigned long long)n) s[*first] = mk<double>(1, 0);
}
// flag if some stored pair misses |s_i A_ij − s_j A_ji| ≤ 1e-10 (|s_i A_ij| + |s_j A_ji|)
__global__ void k_sym_check(int64_t n, const int64_t *rowptr, const int32_t *colind, const c64 *val,
                            const int64_t *colptr, const int64_t *rowval, const c64 *nzval, const c64 *sv, int *flag) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t a = rowptr[i], ae = rowptr[i + 1], b = colptr[i] - 1, be = colptr[i + 1] - 1;
    while (a < ae || b < be) {
        const int64_t ja = a < ae ? colind[a] : INT64_MAX, jb = b < be ? rowval[b] - 1 : INT64_MAX;
        const int64_t j = ja < jb ? ja : jb;
        const c64 aij = ja == j ? val[a] : mk<double>(0, 0), aji = jb == j ? nzval[b] : mk<double>(0, 0);
        if (ja == j) ++a;
        if (jb == j) ++b;
        const c64 l = sv[i] * aij, r = sv[j] * aji;
        const double d = hypot(l.re - r.re, l.im - r.im), m = hypot(l.re, l.im) + hypot(r.re, r.im);
        if (!(d <= 1e-10 * m)) { atomicOr(flag, 1); return; }
    }
}
// T = √s and the solver's scalings: qs = q T⁻¹ (input), qo = q T (output)
__global__ void k_sym_scale(int64_t n, const c64 *sv, const c64 *q, c64 *t, c64 *tinv, c64 *qs, c64 *qo) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const c64 ti = csqrt_(sv[i]), tv = cinv(ti);
    t[i] = ti;
    tinv[i] = tv;
    qs[i] = q[i] * tv;
    qo[i] = q[i] * ti;
}

#include "stencil_tiled.hpp"

// ------------------------------------------------------- vector kernels

// Non-temporal (streaming) vector access: loads and stores that do not
// allocate in the caches.  When an iteration's vectors and operator data
// exceed the 256 MB Infinity Cache, the SQMR vector updates stream through
// them this way (d and z are reused only an iteration later), so the cache
// keeps the stencil's operands instead.
typedef double ntd2 __attribute__((ext_vector_type(2)));
typedef float ntf2 __attribute__((ext_vector_type(2)));
__device__ inline c64 ntld(const c64 *p) {
    const ntd2 v = __builtin_nontemporal_load((const ntd2 *)p);
    return mk<double>(v.x, v.y);
}
__device__ inline c32 ntld(const c32 *p) {
    const ntf2 v = __builtin_nontemporal_load((const ntf2 *)p);
    return mk<float>(v.x, v.y);
}
__device__ inline void ntst(c64 *p, c64 a) { __builtin_nontemporal_store(ntd2{a.re, a.im}, (ntd2 *)p); }
__device__ inline void ntst(c32 *p, c32 a) { __builtin_nontemporal_store(ntf2{a.re, a.im}, (ntf2 *)p); }
template <bool NT, typename T>
__device__ inline T vld(const T *p) {
    if (NT) return ntld(p);
    return *p;
}
template <bool NT, typename T>
__device__ inline void vst(T *p, T a) {
    if (NT) ntst(p, a);
    else *p = a;
}

// The SQMR step's r update with the stencil's finish folded in: every block
// sums the stencil's qᵀPq partials in the fixed order k_fin_dot uses (so σ and
// α = ρ/σ are bitwise those of the separate finish), block 0 records them in
// the solver scalars, and r -= α t follows with the partials of ||r||², rᵀr.
// The scalars' flags are read once per block (thread 0), so a block never
// splits on a flag that block 0 updates meanwhile.
template <typename R, bool NT>
__global__ __launch_bounds__(BS) void k_upd_r_dot(cx<R> *__restrict__ rall, const cx<R> *__restrict__ tall,
                                                  int64_t n, int64_t vs, const double *dpart, int nbd, double *part,
                                                  Scal *scal) {
    const int rhs = blockIdx.y;
    Scal &S = scal[rhs];
    __shared__ int fl[2];
    __shared__ double bc[4];
    if (threadIdx.x == 0) {
        fl[0] = S.done;
        fl[1] = S.conv;
        bc[2] = S.rho_re;
        bc[3] = S.rho_im;
    }
    __syncthreads();
    if (fl[0]) return;
    double tot[2];
    sum_partials<2>(dpart + (size_t)rhs * NPART * 4, nbd, tot);
    if (threadIdx.x == 0) {
        bc[0] = tot[0];
        bc[1] = tot[1];
    }
    __syncthreads();
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (fl[1]) {   // converged last step: this step does nothing
        if (lead) S.done = 1;
        return;
    }
    const double sr = bc[0], si = bc[1];
    if (sr == 0.0 && si == 0.0) {   // breakdown: no step
        if (lead) {
            S.sig_re = S.sig_im = 0;
            S.breakdown = 1;
            S.conv = 1;
            S.alpha_re = S.alpha_im = 0;
            S.apend_re = S.apend_im = 0;
            // k_fin_r skips a converged RHS, so k_upd_dzq would otherwise
            // apply the previous step's coefficients once more: d = 0, z kept
            S.c2t2 = 0;
            S.c2a_re = S.c2a_im = 0;
        }
        return;
    }
    const c64 al = cdiv(mk<double>(bc[2], bc[3]), mk<double>(sr, si));
    if (lead) {
        S.sig_re = sr;
        S.sig_im = si;
        S.alpha_re = S.apend_re = al.re;
        S.alpha_im = S.apend_im = al.im;
    }
    cx<R> *r = rall + rhs * vs;
    const cx<R> *t = tall + rhs * vs;
    double acc[3] = {0, 0, 0};
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        c64 v = cvt<double>(vld<NT>(&r[e])) - al * cvt<double>(vld<NT>(&t[e]));
        vst<NT>(&r[e], cvt<R>(v));
        acc[0] += (double)v.re * v.re + (double)v.im * v.im;
        acc[1] += (double)v.re * v.re - (double)v.im * v.im;
        acc[2] += 2.0 * (double)v.re * v.im;
    }
    block_partial<3>(acc, part + (size_t)rhs * NPART * 4, (int)blockIdx.x);
}

// after k_upd_r_dot: ||r||, θ, c, τ, β, ρ of the SQMR step (one block per RHS)
__global__ __launch_bounds__(BS) void k_fin_r(const double *part, int nb, Scal *scal) {
    const int rhs = blockIdx.x;
    Scal &S = scal[rhs];
    if (S.done || S.conv) return;
    double tot[3];
    sum_partials<3>(part + (size_t)rhs * NPART * 4, nb, tot);
    if (threadIdx.x != 0) return;
    double rn = sqrt(tot[0]);
    double th1 = rn / S.tau, c = 1.0 / sqrt(1.0 + th1 * th1);
    S.c2t2 = c * c * S.theta * S.theta;
    S.c2a_re = c * c * S.alpha_re;
    S.c2a_im = c * c * S.alpha_im;
    S.tau = S.tau * th1 * c;
    c64 rho1 = mk<double>(tot[1], tot[2]);
    if (S.rho_re == 0.0 && S.rho_im == 0.0) { S.beta_re = S.beta_im = 0; S.breakdown = 1; S.conv = 1; }
    else {
        c64 b = cdiv(rho1, mk<double>(S.rho_re, S.rho_im));
        S.beta_re = b.re;
        S.beta_im = b.im;
    }
    S.rho_re = rho1.re;
    S.rho_im = rho1.im;
    S.theta = th1;
    S.rnorm2 = tot[0];
    S.it += 1;
    if (S.tau <= S.target || S.it >= S.itmax || rn == 0.0) S.conv = 1;
}

// d = c²θ² d + c²α q;  z += d;  q = r + β q
// QNT: the new q is streamed too (the q vectors of all RHS exceed the cache,
// so keeping them there only evicts other data)
template <typename R, bool NT, bool QNT = false>
__global__ __launch_bounds__(BS) void k_upd_dzq(cx<R> *__restrict__ dall, cx<R> *__restrict__ zall,
                                                cx<R> *__restrict__ qall, const cx<R> *__restrict__ rall, int64_t n,
                                                int64_t vs, const Scal *scal) {
    const int rhs = blockIdx.y;
    const Scal &S = scal[rhs];
    if (S.done) return;
    cx<R> *d = dall + rhs * vs, *z = zall + rhs * vs, *q = qall + rhs * vs;
    const cx<R> *r = rall + rhs * vs;
    double c2t2 = S.c2t2;
    c64 c2a = mk<double>(S.c2a_re, S.c2a_im), be = mk<double>(S.beta_re, S.beta_im);
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        c64 qe = cvt<double>(vld<NT>(&q[e]));
        c64 de = c2t2 * cvt<double>(vld<NT>(&d[e])) + c2a * qe;
        vst<NT>(&d[e], cvt<R>(de));
        vst<NT>(&z[e], cvt<R>(cvt<double>(vld<NT>(&z[e])) + de));
        const cx<R> qn = cvt<R>(cvt<double>(vld<NT>(&r[e])) + be * qe);   // read next by the stencil
        if (QNT) vst<true>(&q[e], qn);
        else q[e] = qn;
    }
}

// cycle start: c = (Q T) r64 in precision R; r = q = c; d = z = 0
template <typename R>
__global__ __launch_bounds__(BS) void k_init(DGrid<double> g, OpData<double> od, const c64 *__restrict__ r64all,
                                             cx<R> *rall, cx<R> *qall, cx<R> *dall, cx<R> *zall, int64_t n,
                                             int64_t vs, double *part, unsigned *cnt, Scal *scal, double rel_target,
                                             const int *active, long long itmax) {
    const int rhs = blockIdx.y;
    const c64 *r64 = r64all + rhs * vs;
    cx<R> *r = rall + rhs * vs, *q = qall + rhs * vs, *d = dall + rhs * vs, *z = zall + rhs * vs;
    bool act = active[rhs] != 0;
    double acc[3] = {0, 0, 0};
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        int c;
        int64_t p;
        dof2cp(g, e, c, p);
        int i, j, k;
        cell3(g, p, i, j, k);
        cx<R> v = czero<R>();
        if (act) {
            c64 l = cvt<double>(od.q[3 * p + c]) * tfac(g, c, i, j, k, false);
            c64 s = l * r64[e];
            v = mk<R>((R)s.re, (R)s.im);
        }
        r[e] = v;
        q[e] = v;
        d[e] = czero<R>();
        z[e] = czero<R>();
        acc[0] += (double)v.re * v.re + (double)v.im * v.im;
        acc[1] += (double)v.re * v.re - (double)v.im * v.im;
        acc[2] += 2.0 * (double)v.re * v.im;
    }
    double tot[3];
    if (!reduce_last<3>(acc, part + (size_t)rhs * NPART * 4, cnt + rhs, tot)) return;
    if (threadIdx.x != 0) return;
    Scal &S = scal[rhs];
    memset(&S, 0, sizeof(Scal));
    S.tau = sqrt(tot[0]);
    S.cn = S.tau;
    S.rho_re = tot[1];
    S.rho_im = tot[2];
    S.target = rel_target * S.cn;
    S.itmax = itmax;
    S.tau_ref = S.tau;
    S.it_ref = 0;
    if (!act || S.cn == 0.0) { S.conv = 1; S.done = 1; }
}

// e64 += (Q T⁻¹) z
template <typename R>
__global__ __launch_bounds__(BS) void k_upd_e(DGrid<double> g, OpData<double> od, c64 *__restrict__ eall,
                                              const cx<R> *__restrict__ zall, int64_t n, int64_t vs) {
    const int rhs = blockIdx.y;
    c64 *e64 = eall + rhs * vs;
    const cx<R> *z = zall + rhs * vs;
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        cx<R> ze = z[e];
        if (ze.re == 0 && ze.im == 0) continue;
        int c;
        int64_t p;
        dof2cp(g, e, c, p);
        int i, j, k;
        cell3(g, p, i, j, k);
        c64 rf = cvt<double>(od.q[3 * p + c]) * tfac(g, c, i, j, k, true);
        e64[e] = e64[e] + rf * cvt<double>(ze);
    }
}


// probe of the true residual inside a cycle: et = e + (Q T⁻¹) z (FP64)
template <typename R>
__global__ __launch_bounds__(BS) void k_probe_x(DGrid<double> g, OpData<double> od, const c64 *__restrict__ eall,
                                                const cx<R> *__restrict__ zall, c64 *__restrict__ etall, int64_t n,
                                                int64_t vs) {
    const int rhs = blockIdx.y;
    const c64 *e64 = eall + rhs * vs;
    const cx<R> *z = zall + rhs * vs;
    c64 *et = etall + rhs * vs;
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        int c;
        int64_t p;
        dof2cp(g, e, c, p);
        int i, j, k;
        cell3(g, p, i, j, k);
        c64 rf = cvt<double>(od.q[3 * p + c]) * tfac(g, c, i, j, k, true);
        et[e] = e64[e] + rf * cvt<double>(z[e]);
    }
}

// ||b||² per RHS
__global__ __launch_bounds__(BS) void k_norm2(const c64 *__restrict__ ball, int64_t n, int64_t vs, double *part,
                                              unsigned *cnt, double *out) {
    const int rhs = blockIdx.y;
    const c64 *b = ball + rhs * vs;
    double acc[1] = {0};
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS)
        acc[0] += b[e].re * b[e].re + b[e].im * b[e].im;
    double tot[1];
    if (!reduce_last<1>(acc, part + (size_t)rhs * NPART * 4, cnt + rhs, tot)) return;
    if (threadIdx.x == 0) out[rhs] = tot[0];
}

// y = s ∘ x with s = T² (S_E, inv = false) or T⁻² (inv = true), in place
__global__ __launch_bounds__(BS) void k_scale_S(DGrid<double> g, c64 *x, int64_t n, int inv) {
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        int c;
        int64_t p;
        dof2cp(g, e, c, p);
        int i, j, k;
        cell3(g, p, i, j, k);
        c64 t = tfac(g, c, i, j, k, inv != 0);
        x[e] = (t * t) * x[e];
    }
}

// grad = 2 Re(ω² λ e)
__global__ __launch_bounds__(BS) void k_grad(const c64 *e, const c64 *lam, c64 om2, double *grad, int64_t n) {
    for (int64_t k = blockIdx.x * (int64_t)BS + threadIdx.x; k < n; k += (int64_t)gridDim.x * BS) {
        c64 v = om2 * (lam[k] * e[k]);
        grad[k] = 2.0 * v.re;
    }
}

// ----------------------------------------------------- operator set-up

// pack (Nx,Ny,Nz,3,3) arrays into per-cell triples
__global__ void k_pack(const c64 *eps, const c64 *mu, int64_t M, c64 *epsd, c64 *epso, c64 *muinv, int *asym) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= M) return;
    for (int c = 0; c < 3; ++c) {
        epsd[3 * p + c] = eps[p + M * (c + 3 * c)];
        if (muinv) muinv[3 * p + c] = cinv(mu[p + M * (c + 3 * c)]);
    }
    if (epso) {
        const int vv[3] = {0, 0, 1}, ww[3] = {1, 2, 2};
        for (int x = 0; x < 3; ++x) {
            // ε_vw and ε_wv share one stored value: they may differ by rounding
            // only, measured against the size of the tensor (its diagonal too)
            c64 a = eps[p + M * (vv[x] + 3 * ww[x])], b = eps[p + M * (ww[x] + 3 * vv[x])];
            c64 dv = eps[p + M * (vv[x] + 3 * vv[x])], dw = eps[p + M * (ww[x] + 3 * ww[x])];
            double sc = fabs(a.re) + fabs(a.im) + fabs(b.re) + fabs(b.im) + fabs(dv.re) + fabs(dv.im) +
                        fabs(dw.re) + fabs(dw.im);
            if (fabs(a.re - b.re) + fabs(a.im - b.im) > 1e-12 * sc) atomicOr(asym, 1);
            epso[3 * p + x] = mk<double>(0.5 * (a.re + b.re), 0.5 * (a.im + b.im));
        }
    }
}

// Jacobi scaling q = (c − ω² ε_cc)^{-1/2}
template <bool MU>
__global__ void k_jacobi(DGrid<double> g, OpData<double> od, c64 *q, int use) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= g.M) return;
    int i, j, k;
    cell3(g, p, i, j, k);
    for (int c = 0; c < 3; ++c) {
        if (!use) { q[3 * p + c] = mk<double>(1.0, 0.0); continue; }
        c64 dg = ccdiag<double, MU>(g, od, c, i, j, k) - od.om2 * od.epsd[3 * p + c];
        q[3 * p + c] = cinv(csqrt_(dg));
    }
}

// qs = q·T⁻¹ per E DOF: the tiled stencil scales its staged input with this one
// factor and recovers q = qs·T for the owner's output
__global__ void k_qscale(DGrid<double> g, const c64 *q, c64 *qs) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= g.M) return;
    int i, j, k;
    cell3(g, p, i, j, k);
    for (int c = 0; c < 3; ++c) qs[3 * p + c] = q[3 * p + c] * tfac(g, c, i, j, k, true);
}

// grad-div weight W = 1/ε̄² at each primal node, ε̄ the mean of the ε_ww of
// the (up to 6) E edges meeting at the node
__global__ void k_wnode(DGrid<double> g, const c64 *epsd, c64 *w, c32 *w32) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= g.M) return;
    int id[3];
    cell3(g, p, id[0], id[1], id[2]);
    c64 s = mk<double>(0, 0);
    int n = 0;
    for (int a = 0; a < 3; ++a) {
        s = s + epsd[3 * p + a];
        ++n;
        int64_t st = a == 0 ? 1 : (a == 1 ? (int64_t)g.N[0] : (int64_t)g.N[0] * g.N[1]);
        if (id[a] > 0) { s = s + epsd[3 * (p - st) + a]; ++n; }
        else if (g.bloch[a]) { s = s + epsd[3 * (p + (int64_t)(g.N[a] - 1) * st) + a]; ++n; }
    }
    c64 m = mk<double>(s.re / n, s.im / n);
    c64 v = cinv(m * m);   // W = 1/ε̄² at the node
    w[p] = v;
    w32[p] = mk<float>((float)v.re, (float)v.im);
}

// Jacobi scaling of the grad-div form: q = (c − ω² ε + ε² (G W Div)-diagonal)^{-1/2}
__global__ void k_jacobi_gd(DGrid<double> g, OpData<double> od, c64 *q) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= g.M) return;
    int id[3];
    cell3(g, p, id[0], id[1], id[2]);
    for (int u = 0; u < 3; ++u) {
        c64 e = od.epsd[3 * p + u];
        c64 dg = ccdiag<double, false>(g, od, u, id[0], id[1], id[2]) - od.om2 * e;
        int Nu = g.N[u], iu = id[u];
        int64_t st = u == 0 ? 1 : (u == 1 ? (int64_t)g.N[0] : (int64_t)g.N[0] * g.N[1]);
        c64 invD = g.sdlinv[DUAL][u][iu];
        c64 gd;
        if (Nu == 1 && g.bloch[u]) {
            c64 f = (g.ph[u] - mk<double>(1, 0)) * (mk<double>(1, 0) - g.phinv[u]);
            gd = -((invD * g.sdlinv[PRIM][u][0]) * (od.wnode[p] * f));
        } else {
            gd = od.wnode[p] * g.sdlinv[PRIM][u][iu];
            if (iu < Nu - 1) gd = gd + od.wnode[p + st] * g.sdlinv[PRIM][u][iu + 1];
            else if (g.bloch[u]) gd = gd + od.wnode[p - (int64_t)(Nu - 1) * st] * g.sdlinv[PRIM][u][0];
            gd = invD * gd;
        }
        dg = dg + (e * e) * gd;
        q[3 * p + u] = cinv(csqrt_(dg));
    }
}

// ψ = W Div(r) at the primal nodes (backward differences), per RHS
__global__ __launch_bounds__(BS) void k_gd_psi(DGrid<double> g, const c64 *wnode, const c64 *rall, c64 *psiall,
                                               int64_t vs) {
    const int rhs = blockIdx.y;
    const c64 *r = rall + rhs * vs;
    c64 *psi = psiall + rhs * g.M;
    for (int64_t p = blockIdx.x * (int64_t)BS + threadIdx.x; p < g.M; p += (int64_t)gridDim.x * BS) {
        int id[3];
        cell3(g, p, id[0], id[1], id[2]);
        c64 dv = mk<double>(0, 0);
        for (int w = 0; w < 3; ++w) {
            int64_t st = w == 0 ? 1 : (w == 1 ? (int64_t)g.N[0] : (int64_t)g.N[0] * g.N[1]);
            c64 a = r[dofidx(g, w, p)], b = mk<double>(0, 0);
            if (id[w] > 0) b = r[dofidx(g, w, p - st)];
            else if (g.bloch[w]) b = r[dofidx(g, w, p + (int64_t)(g.N[w] - 1) * st)] * g.phinv[w];
            dv = dv + g.sdlinv[PRIM][w][id[w]] * (a - b);
        }
        psi[p] = wnode[p] * dv;
    }
}

// r' = r + (1/ω²) ε G ψ (forward differences of ψ to the E edges), in place
__global__ __launch_bounds__(BS) void k_gd_rhs(DGrid<double> g, OpData<double> od, const c64 *psiall, c64 *rall,
                                               int64_t vs, c64 inv_om2) {
    const int rhs = blockIdx.y;
    c64 *r = rall + rhs * vs;
    const c64 *psi = psiall + rhs * g.M;
    for (int64_t p = blockIdx.x * (int64_t)BS + threadIdx.x; p < g.M; p += (int64_t)gridDim.x * BS) {
        int id[3];
        cell3(g, p, id[0], id[1], id[2]);
        for (int u = 0; u < 3; ++u) {
            int64_t st = u == 0 ? 1 : (u == 1 ? (int64_t)g.N[0] : (int64_t)g.N[0] * g.N[1]);
            c64 pn = mk<double>(0, 0);
            if (id[u] < g.N[u] - 1) pn = psi[p + st];
            else if (g.bloch[u]) pn = psi[p - (int64_t)(g.N[u] - 1) * st] * g.ph[u];
            c64 gv = g.sdlinv[DUAL][u][id[u]] * (pn - psi[p]);
            int64_t e = dofidx(g, u, p);
            r[e] = r[e] + inv_om2 * (od.epsd[3 * p + u] * gv);
        }
    }
}

__global__ void k_down(const c64 *a, c32 *b, int64_t n) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < n) b[e] = mk<float>((float)a[e].re, (float)a[e].im);
}
__global__ void k_up(const c32 *a, c64 *b, int64_t n) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < n) b[e] = mk<double>((double)a[e].re, (double)a[e].im);
}

static inline unsigned nb(int64_t n, int bs = BS) { return (unsigned)((n + bs - 1) / bs); }
static inline unsigned ngrid(int64_t n) {
    int64_t b = (n + BS - 1) / BS;
    return (unsigned)(b < NGRID_MAX ? (b < 1 ? 1 : b) : NGRID_MAX);
}

void Workspace::release() {
    if (fixed) {   // caller-provided memory: forget it, free nothing
        *this = Workspace();
        return;
    }
    if (graph) hipGraphExecDestroy(graph);
    graph = nullptr;
    if (buf) hipFree(buf);
    buf = nullptr;
    if (scal_h) hipHostFree(scal_h);
    scal_h = nullptr;
    if (rn_h) hipHostFree(rn_h);
    rn_h = nullptr;
    nrhs = 0;
}

// scratch of the two-pass direct kernel for nrhs right-hand sides (operators
// the tiled stencils do not serve); allocated outside any graph capture
static int hbuf_ensure(mfd_op *op, int64_t nrhs) {
    if (op->csr) return MFD_OK;
    const bool tiled_ok = !(op->has_mu && op->full) && (op->ee_all || op->rev);
    if (tiled_ok && op->tile64[0] > 0 && op->tile32[0] > 0) return MFD_OK;
    const size_t need = sizeof(c64) * 3 * (size_t)op->G.h.M * (size_t)nrhs;
    if (op->hbuf_bytes >= need) return MFD_OK;
    if (op->hbuf) {
        MFD_HIP(hipStreamSynchronize(op->stream));
        hipFree(op->hbuf);
        op->hbuf = nullptr;
        op->hbuf_bytes = 0;
    }
    if (hipMalloc(&op->hbuf, need) != hipSuccess) {
        (void)hipGetLastError();   // clear it: the single-pass kernel serves without the scratch
        op->hbuf = nullptr;
        return MFD_OK;
    }
    op->hbuf_bytes = need;
    return MFD_OK;
}

static inline size_t ws_up(size_t x) { return (x + 255) / 256 * 256; }

// Bytes of a solver workspace for nrhs right-hand sides of n unknowns.  The
// Krylov vectors are sized for FP64 in both modes: a mixed solve whose FP32
// cycles stop paying off continues with FP64 cycles in place.
size_t ws_bytes(int64_t n, int64_t nrhs) {
    const size_t vb = ws_up(sizeof(c64) * n * nrhs);
    return 9 * vb + ws_up(sizeof(c64) * (n / 3) * nrhs) + ws_up(sizeof(double) * 3 * nrhs * NPART * 4) +
           ws_up(sizeof(unsigned) * 3 * nrhs) + ws_up(sizeof(Scal) * nrhs) + 2 * ws_up(sizeof(double) * nrhs) +
           ws_up(sizeof(int) * 3 * nrhs);
}

// Lays out the device part of a workspace of ws_bytes(n, nrhs) bytes at base
// (256-byte aligned); the host mirrors scal_h and rn_h are the caller's.
void ws_carve(Workspace &w, void *base, int64_t n, int64_t nrhs) {
    const size_t vb = ws_up(sizeof(c64) * n * nrhs);
    char *p = (char *)base;
    w.r = p; p += vb;
    w.q = p; p += vb;
    w.t = p; p += vb;
    w.d = p; p += vb;
    w.z = p; p += vb;
    w.r64 = p; p += vb;
    w.et = p; p += vb;
    w.et2 = p; p += vb;
    w.ebest = p; p += vb;
    w.psi = p; p += ws_up(sizeof(c64) * (n / 3) * nrhs);
    w.part = (double *)p; p += ws_up(sizeof(double) * 3 * nrhs * NPART * 4);
    w.cnt = (unsigned *)p; p += ws_up(sizeof(unsigned) * 3 * nrhs);
    w.scal = (Scal *)p; p += ws_up(sizeof(Scal) * nrhs);
    w.bn = (double *)p; p += ws_up(sizeof(double) * nrhs);
    w.rn = (double *)p; p += ws_up(sizeof(double) * nrhs);
    w.active = (int *)p;
    w.nrhs = nrhs;
    w.n = n;
    w.bytes = ws_bytes(n, nrhs);
    w.graph_iters = 0;
}

static int ws_ensure(mfd_op *op, int64_t nrhs, int prec) {
    Workspace &w = op->ws;
    int64_t n = op->n;
    if (w.fixed) {   // a caller-provided workspace: never reallocated here
        if (w.nrhs != nrhs || w.n != n) { set_error("fixed workspace does not match nrhs"); return MFD_ERR_ARG; }
        w.prec = prec;
        return MFD_OK;
    }
    MFD_TRY(hbuf_ensure(op, nrhs));
    if (w.buf && w.nrhs == nrhs && w.prec == prec) return MFD_OK;
    w.release();
    const size_t bytes = ws_bytes(n, nrhs);
    MFD_HIP(hipMalloc(&w.buf, bytes));
    ws_carve(w, w.buf, n, nrhs);
    MFD_HIP(hipMemsetAsync(w.cnt, 0, sizeof(unsigned) * 3 * nrhs, op->stream));
    MFD_HIP(hipHostMalloc((void **)&w.scal_h, sizeof(Scal) * nrhs, hipHostMallocDefault));
    MFD_HIP(hipHostMalloc((void **)&w.rn_h, sizeof(double) * nrhs, hipHostMallocDefault));
    w.prec = prec;
    return MFD_OK;
}

template <typename R> inline const OpData<R> &opd(const mfd_op *op);
template <> inline const OpData<double> &opd<double>(const mfd_op *op) { return op->d64; }
template <> inline const OpData<float> &opd<float>(const mfd_op *op) { return op->d32; }

// finish kernel of a stencil epilogue with a reduction (nb partials per RHS)
template <int EPI>
static void launch_finish(hipStream_t s, int nb, int64_t nrhs, double *part, Scal *scal, double *rn, const int *mask) {
    if (EPI == EPI_DOT) k_fin_dot<<<(unsigned)nrhs, BS, 0, s>>>(part, nb, scal);
    else if (EPI == EPI_RESID) k_fin_resid<<<(unsigned)nrhs, BS, 0, s>>>(part, nb, rn, mask);
}

// launch one operator apply; returns the number of partial sums per RHS its
// reduction wrote (finish = false leaves them to the next kernel)
template <typename R, int MODE, int EPI>
static int launch_apply(mfd_op *op, const cx<R> *x, cx<R> *y, const c64 *b, int64_t nrhs, double *part,
                        Scal *scal, double *rn, const int *mask = nullptr, hipEvent_t after_main = nullptr,
                        bool finish = true) {
    const DGrid<R> &g = op->G.view<R>();
    const OpData<R> &od = opd<R>(op);
    int64_t vs = op->n;
    hipStream_t s = op->stream;
    int ty = std::is_same<R, double>::value ? op->tile64[0] : op->tile32[0];
    int kz = std::is_same<R, double>::value ? op->tile64[1] : op->tile32[1];
    int xr = std::is_same<R, double>::value ? op->tile64[2] : op->tile32[2];
    if (op->csr) {   // an operator built from a caller's CSC matrix
        dim3 grid(ngrid(vs), (unsigned)nrhs);
        k_spmv<R, MODE, EPI><<<grid, BS, 0, s>>>(vs, op->csr_rowptr, op->csr_colind, op->csr_val, op->d64.qs,
                                                 op->csr_qo, x, y, b, vs, part, scal, mask);
        if (after_main) hipEventRecord(after_main, s);
        if (finish) launch_finish<EPI>(s, (int)grid.x, nrhs, part, scal, rn, mask);
        return (int)grid.x;
    }
    if (MODE == MODE_P && op->gd && ty == 8) ty = 4;   // the grad-div tile keeps two more LDS planes
    // μ ≠ 1 has a tiled form with diagonal ε (the grad-div form is never on with μ)
    if ((op->ee_all || op->rev) && ty > 0 && !(op->has_mu && op->full)) {
        // order bit 1: 32-wide tiles of 2·TY rows (same threads per block);
        // the off-diagonal, grad-div, μ and reflected stencils are built 64-wide only
        const bool narrow = (xr & 2) && !op->full && !(MODE == MODE_P && op->gd) && !op->has_mu && !op->rev && !op->win;
        xr &= 1;
        const int tw = narrow ? 32 : 64, rows = narrow ? 2 * ty : ty;
        int nx = (g.N[0] + tw - 1) / tw, ny = (g.N[1] + rows - 1) / rows;
        int nch = (g.N[2] + kz - 1) / kz;
        while ((int64_t)nx * ny * nch > NPART) { kz *= 2; nch = (g.N[2] + kz - 1) / kz; }
        dim3 grid(nx, ny, (unsigned)(nch * nrhs));
        const DGrid<double> &gr = op->G.r64;
        if (op->win && op->full) {   // a component window (2D and 1D models), μ = 1, boundft EE
            if (ty == 8) k_apply_tiled<R, MODE, EPI, 8, false, false, true, 64, false, false, true><<<grid, 512, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 2) k_apply_tiled<R, MODE, EPI, 2, false, false, true, 64, false, false, true><<<grid, 128, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 4, false, false, true, 64, false, false, true><<<grid, 256, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 1, false, false, true, 64, false, false, true><<<grid, 64, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        } else if (op->win) {
            if (ty == 8) k_apply_tiled<R, MODE, EPI, 8, false, false, false, 64, false, false, true><<<grid, 512, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 2) k_apply_tiled<R, MODE, EPI, 2, false, false, false, 64, false, false, true><<<grid, 128, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 4, false, false, false, 64, false, false, true><<<grid, 256, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 1, false, false, false, 64, false, false, true><<<grid, 64, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        } else if (op->rev && op->has_mu) {   // boundft HH on some axis: the grid reflected along those axes
            if (ty == 8) k_apply_tiled<R, MODE, EPI, 8, false, false, false, 64, true, true><<<grid, 512, 0, s>>>(gr, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 2) k_apply_tiled<R, MODE, EPI, 2, false, false, false, 64, true, true><<<grid, 128, 0, s>>>(gr, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 4, false, false, false, 64, true, true><<<grid, 256, 0, s>>>(gr, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 1, false, false, false, 64, true, true><<<grid, 64, 0, s>>>(gr, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        } else if (op->rev) {
            if (ty == 8) k_apply_tiled<R, MODE, EPI, 8, false, false, false, 64, false, true><<<grid, 512, 0, s>>>(gr, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 2) k_apply_tiled<R, MODE, EPI, 2, false, false, false, 64, false, true><<<grid, 128, 0, s>>>(gr, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 4, false, false, false, 64, false, true><<<grid, 256, 0, s>>>(gr, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 1, false, false, false, 64, false, true><<<grid, 64, 0, s>>>(gr, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        } else if (op->full) {   // off-diagonal ε (grad-div is never enabled with it)
            if (ty == 8) k_apply_tiled<R, MODE, EPI, 8, false, false, true><<<grid, 512, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 2) k_apply_tiled<R, MODE, EPI, 2, false, false, true><<<grid, 128, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 4, false, false, true><<<grid, 256, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 1, false, false, true><<<grid, 64, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        } else if (MODE == MODE_P && op->gd) {
            if (ty == 2) k_apply_tiled<R, MODE, EPI, 2, true><<<grid, 128, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 4, true><<<grid, 256, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 1, true><<<grid, 64, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        } else if (op->has_mu) {
            if (ty == 8) k_apply_tiled<R, MODE, EPI, 8, false, false, false, 64, true><<<grid, 512, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 2) k_apply_tiled<R, MODE, EPI, 2, false, false, false, 64, true><<<grid, 128, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 4, false, false, false, 64, true><<<grid, 256, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 1, false, false, false, 64, true><<<grid, 64, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        } else if (narrow) {
            if (ty == 8) k_apply_tiled<R, MODE, EPI, 16, false, false, false, 32><<<grid, 512, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 2) k_apply_tiled<R, MODE, EPI, 4, false, false, false, 32><<<grid, 128, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 8, false, false, false, 32><<<grid, 256, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 2, false, false, false, 32><<<grid, 64, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        } else {
            if (ty == 8) k_apply_tiled<R, MODE, EPI, 8, false><<<grid, 512, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 2) k_apply_tiled<R, MODE, EPI, 2, false><<<grid, 128, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else if (ty == 4) k_apply_tiled<R, MODE, EPI, 4, false><<<grid, 256, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
            else k_apply_tiled<R, MODE, EPI, 1, false><<<grid, 64, 0, s>>>(op->G.d64, op->d64, op->d64.om2, x, y, b, vs, kz, nch, part, scal, mask, xr);
        }
        if (after_main) hipEventRecord(after_main, s);
        if (finish) launch_finish<EPI>(s, nx * ny * nch, nrhs, part, scal, rn, mask);
        return nx * ny * nch;
    }
    dim3 grid(ngrid(g.M), (unsigned)nrhs);
    // two passes when the scratch is there: H = μ⁻¹ Cₑ X once per cell, then
    // E from H (instead of recomputing H at each cell and its six neighbours)
    const bool two = op->hbuf && op->hbuf_bytes >= sizeof(cx<R>) * (size_t)vs * (size_t)nrhs;
    cx<R> *hb = (cx<R> *)op->hbuf;
    if (two) {
        if (op->has_mu) k_hpass<R, MODE, true><<<grid, BS, 0, s>>>(g, od, x, hb, vs, scal, EPI == EPI_DOT);
        else k_hpass<R, MODE, false><<<grid, BS, 0, s>>>(g, od, x, hb, vs, scal, EPI == EPI_DOT);
    }
#define MFD_DIRECT(OFF, MU_)                                                                                   \
    do {                                                                                                       \
        if (two) k_apply<R, MODE, OFF, MU_, EPI, true><<<grid, BS, 0, s>>>(g, od, x, y, b, vs, part, scal, mask, hb); \
        else k_apply<R, MODE, OFF, MU_, EPI><<<grid, BS, 0, s>>>(g, od, x, y, b, vs, part, scal, mask);          \
    } while (0)
    if (op->full) {
        if (op->has_mu) MFD_DIRECT(true, true);
        else MFD_DIRECT(true, false);
    } else {
        if (op->has_mu) MFD_DIRECT(false, true);
        else MFD_DIRECT(false, false);
    }
#undef MFD_DIRECT
    if (after_main) hipEventRecord(after_main, s);
    if (finish) launch_finish<EPI>(s, (int)grid.x, nrhs, part, scal, rn, mask);
    return (int)grid.x;
}

template <typename R>
static void launch_upd_dzq(dim3 gv, hipStream_t s, cx<R> *d, cx<R> *z, cx<R> *q, const cx<R> *r, int64_t n,
                           const Scal *scal, bool nt) {
    // the new q of every RHS fits the 256 MB Infinity Cache: keep it there for
    // the next stencil; otherwise stream it as well
    const bool qnt = nt && (double)n * gv.y * sizeof(cx<R>) > 256.0 * 1024 * 1024;
    if (qnt) k_upd_dzq<R, true, true><<<gv, BS, 0, s>>>(d, z, q, r, n, n, scal);
    else if (nt) k_upd_dzq<R, true><<<gv, BS, 0, s>>>(d, z, q, r, n, n, scal);
    else k_upd_dzq<R, false><<<gv, BS, 0, s>>>(d, z, q, r, n, n, scal);
}

// Whether one SQMR step's vectors (r, q, t, d, z of every RHS) and the
// stencil's operator data (qs) exceed the Infinity Cache (256 MB on MI355X):
// then the vector updates stream non-temporally (measured on 128³ FP64: step
// 0.273 → 0.242 ms; on 64³, where everything stays cached, 0.074 → 0.076 ms,
// so small problems keep cached accesses).
static bool stream_nt(int64_t n, int64_t nrhs, size_t vbytes) {
    const double foot = (double)n * (double)(5 * nrhs) * (double)vbytes + (double)n * sizeof(c64);
    return foot > 256.0 * 1024 * 1024;
}

// One SQMR step: stencil t = P q with qᵀt partials, r update (with the α
// finish folded in), finish of ||r||, rᵀr, d/z/q update.  E (optional): six
// events around the main kernels, for timing in context.
template <typename R>
static void launch_iter(mfd_op *op, int64_t nrhs, hipEvent_t *E = nullptr) {
    Workspace &w = op->ws;
    int64_t n = w.n;
    hipStream_t s = op->stream;
    // about one resident round of blocks in total: every RHS streams at once
    // (with ngrid(n) blocks per RHS the launch ran as nrhs successive rounds,
    // measured 24% slower per RHS for the FP64 d/z/q update at 4 RHS)
    const unsigned gx = std::max<unsigned>(ngrid(n) / (unsigned)nrhs, std::min<unsigned>(ngrid(n), 64u));
    dim3 gv(gx, (unsigned)nrhs);
    size_t pslot = (size_t)nrhs * NPART * 4;
    const bool nt = stream_nt(n, nrhs, sizeof(cx<R>));
    if (E) hipEventRecord(E[0], s);
    const int nbd = launch_apply<R, MODE_P, EPI_DOT>(op, (const cx<R> *)w.q, (cx<R> *)w.t, nullptr, nrhs, w.part,
                                                     w.scal, nullptr, nullptr, E ? E[1] : nullptr, false);
    if (E) hipEventRecord(E[2], s);
    if (nt) k_upd_r_dot<R, true><<<gv, BS, 0, s>>>((cx<R> *)w.r, (const cx<R> *)w.t, n, n, w.part, nbd, w.part + pslot, w.scal);
    else k_upd_r_dot<R, false><<<gv, BS, 0, s>>>((cx<R> *)w.r, (const cx<R> *)w.t, n, n, w.part, nbd, w.part + pslot, w.scal);
    if (E) hipEventRecord(E[3], s);
    k_fin_r<<<(unsigned)nrhs, BS, 0, s>>>(w.part + pslot, (int)gv.x, w.scal);
    if (E) hipEventRecord(E[4], s);
    launch_upd_dzq<R>(gv, s, (cx<R> *)w.d, (cx<R> *)w.z, (cx<R> *)w.q, (const cx<R> *)w.r, n, w.scal, nt);
    if (E) hipEventRecord(E[5], s);
}

}  // namespace mfd

using namespace mfd;

// ------------------------------------------------------------- C-ABI glue

int op_create_impl(mfd_ctx *ctx, const mfd_grid *g, int ft, mfd_c64 omega, const mfd_c64 *eps_arr,
                   const mfd_c64 *mu_arr, mfd_op **out) {
    if (ft != MFD_EE && ft != MFD_HH) { set_error("ft must be EE or HH"); return MFD_ERR_ARG; }
    // E↔H duality: the HH operator Cₑ (Pε \ Cₘ) − ω² Pμ of a model is, entry
    // for entry, the EE operator of the same grid with every boundft flipped
    // and ε, μ swapped, so an HH operator is built as that EE operator
    mfd_grid gd = *g;
    if (ft == MFD_HH) {
        for (int a = 0; a < 3; ++a) gd.boundft[a] = gd.boundft[a] == MFD_EE ? MFD_HH : MFD_EE;
        std::swap(eps_arr, mu_arr);
    }
    g = &gd;
    mfd_op *op = new mfd_op();
    op->ctx = ctx;
    op->stream = ctx->stream;
    op->ft = ft;
    for (int a = 0; a < 3; ++a)
        if (g->boundft[a] != MFD_EE) {
            op->ee_all = 0;
            op->rev |= 1 << a;
        }
    int rc = op->G.create(g, ctx->stream);
    if (!rc && op->rev) rc = op->G.reflect(op->rev, ctx->stream);
    if (rc) { delete op; return rc; }
    hipStream_t s = ctx->stream;
    int64_t M = op->G.h.M;
    op->n = 3 * M;
    op->omega = mk<double>(omega.re, omega.im);
    c64 om2 = op->omega * op->omega;
    DevArray E, U;
    int full = 0, mufull = 0;
    auto bail = [&](int code) { delete op; return code; };
    if ((rc = upload_param(ctx, op->G, eps_arr, E, &full))) return bail(rc);
    if ((rc = upload_param(ctx, op->G, mu_arr, U, &mufull))) return bail(rc);
    if (mufull) {
        set_error(ft == MFD_EE ? "create_A: Pμ must be diagonal (Pμ \\ Cₑ), model.jl:236"
                               : "create_A: Pε must be diagonal (Pε \\ Cₘ), model.jl:239");
        return bail(MFD_ERR_UNSUPPORTED);
    }
    if (full && !op->ee_all) {
        set_error(ft == MFD_EE ? "matrix-free operator: off-diagonal ε needs boundft = EE on every axis"
                               : "matrix-free HH operator: off-diagonal μ needs boundft = HH on every axis");
        return bail(MFD_ERR_UNSUPPORTED);
    }
    // is μ identically 1?
    {
        std::vector<int> dummy;
        bool ones = true;
        for (int c = 0; c < 3 && ones; ++c)
            for (int64_t p = 0; p < M; ++p) {
                const mfd_c64 &m = mu_arr[p + M * (c + 3 * c)];
                if (m.re != 1.0 || m.im != 0.0) { ones = false; break; }
            }
        op->has_mu = !ones;
    }
    op->full = full;
    // device arrays: epsd, epso, muinv, q (FP64) and their FP32 copies, + asym flag
    size_t n3 = 3 * (size_t)M;
    size_t b64 = sizeof(c64) * n3, b32 = sizeof(c32) * n3;
    size_t bytes = 4 * (b64 + b32) + b64 + (b64 + b32) / 3 + 512;
    if (hipMalloc(&op->buf, bytes) != hipSuccess) { set_error("hipMalloc failed for operator arrays"); return bail(MFD_ERR_ALLOC); }
    char *p = (char *)op->buf;
    c64 *epsd = (c64 *)p; p += b64;
    c64 *epso = (c64 *)p; p += b64;
    c64 *muinv = (c64 *)p; p += b64;
    c64 *q = (c64 *)p; p += b64;
    c64 *qs = (c64 *)p; p += b64;
    c32 *epsd32 = (c32 *)p; p += b32;
    c32 *epso32 = (c32 *)p; p += b32;
    c32 *muinv32 = (c32 *)p; p += b32;
    c32 *q32 = (c32 *)p; p += b32;
    c64 *wn = (c64 *)p; p += b64 / 3;
    c32 *wn32 = (c32 *)p; p += b32 / 3;
    p = (char *)(((uintptr_t)p + 63) & ~(uintptr_t)63);
    int *asym = (int *)p;
    if (hipMemsetAsync(asym, 0, sizeof(int), s) != hipSuccess) return bail(MFD_ERR_HIP);
    k_pack<<<nb(M), BS, 0, s>>>((const c64 *)E.p, (const c64 *)U.p, M, epsd, full ? epso : nullptr,
                                op->has_mu ? muinv : nullptr, asym);
    int asym_h = 0;
    if (hipMemcpyAsync(&asym_h, asym, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { set_error("operator pack failed"); return bail(MFD_ERR_HIP); }
    if (asym_h) { set_error("matrix-free operator: off-diagonal ε must be symmetric"); return bail(MFD_ERR_UNSUPPORTED); }
    op->d64.epsd = epsd;
    op->d64.epso = full ? epso : nullptr;
    op->d64.muinv = op->has_mu ? muinv : nullptr;
    op->d64.q = q;
    op->d64.qs = qs;
    op->d64.om2 = om2;
    op->d64.wnode = wn;
    // with a complex Bloch phase e^{-ikL} the operator is not complex-symmetric
    // (the wrap entries of Cₑ and Cₘ are conjugate, not transposed, images)
    for (int a = 0; a < 3; ++a)
        if (op->G.d64.bloch[a] && fabs(op->G.d64.ph[a].im) > 1e-14) op->sym = 0;
    op->gd_ok = !full && !op->has_mu && op->ee_all && !(om2.re == 0.0 && om2.im == 0.0);
    if (!op->ee_all && !op->rev) op->tile64[0] = op->tile32[0] = 0;   // the direct kernel serves mixed boundaries
    else {
        // default tile (until mfd_op_autotune or mfd_op_set_tile): 4 rows x 16
        // planes, with fewer planes and then fewer rows per block on grids too
        // small for that to give one tile per CU (256); 32-wide tiles (of
        // twice the rows) where they leave fewer x lanes idle than 64-wide ones
        const int64_t N0 = op->G.h.N[0];
        const bool narrow = !full && !op->has_mu && !op->rev && 64 * ((N0 + 63) / 64) - N0 > 32 * ((N0 + 31) / 32) - N0;
        const int64_t nx = narrow ? (N0 + 31) / 32 : (N0 + 63) / 64;
        auto tiles = [&](int ty, int kz) {
            const int rows = narrow ? 2 * ty : ty;
            return nx * ((op->G.h.N[1] + rows - 1) / rows) * ((op->G.h.N[2] + kz - 1) / kz);
        };
        int ty = 4, kz = 16;
        while (kz > 4 && tiles(ty, kz) < 256) kz /= 2;
        while (ty > 1 && tiles(ty, kz) < 256) ty /= 2;
        // a one-cell z axis with full ε (the smoothed 2D models): each tile
        // marches a single plane, and the 8-row tile is the faster one
        // wherever it still gives a tile per CU (2D TE 1024²: 8.67 → 8.31 s)
        if (full && op->G.h.N[2] == 1 && ty == 4 && tiles(8, kz) >= 256) ty = 8;
        op->tile64[0] = op->tile32[0] = ty;
        op->tile64[1] = op->tile32[1] = kz;
        op->tile64[2] = op->tile32[2] = narrow ? 3 : 1;
    }
    op->gd = 0;   // opt-in: it helps small PML-dominated grids and slows larger ones
    k_wnode<<<nb(M), BS, 0, s>>>(op->G.d64, epsd, wn, wn32);
    if (op->gd) k_jacobi_gd<<<nb(M), BS, 0, s>>>(op->G.d64, op->d64, q);
    else if (op->has_mu) k_jacobi<true><<<nb(M), BS, 0, s>>>(op->G.d64, op->d64, q, 1);
    else k_jacobi<false><<<nb(M), BS, 0, s>>>(op->G.d64, op->d64, q, 1);
    k_qscale<<<nb(M), BS, 0, s>>>(op->G.d64, q, qs);
    k_down<<<nb(n3), BS, 0, s>>>(epsd, epsd32, n3);
    if (full) k_down<<<nb(n3), BS, 0, s>>>(epso, epso32, n3);
    if (op->has_mu) k_down<<<nb(n3), BS, 0, s>>>(muinv, muinv32, n3);
    k_down<<<nb(n3), BS, 0, s>>>(q, q32, n3);
    op->d32.epsd = epsd32;
    op->d32.epso = full ? epso32 : nullptr;
    op->d32.muinv = op->has_mu ? muinv32 : nullptr;
    op->d32.q = q32;
    op->d32.qs = nullptr;   // the tiled stencil reads the FP64 operator data only
    op->d32.wnode = wn32;
    op->d32.om2 = mk<float>((float)om2.re, (float)om2.im);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { set_error("operator set-up failed"); return bail(MFD_ERR_HIP); }
    hipEventCreate(&op->ev0);
    hipEventCreate(&op->ev1);
    *out = op;
    return MFD_OK;
}

int op_destroy_impl(mfd_op *op) {
    if (!op) return MFD_OK;
    hipStreamSynchronize(op->stream);
    op->ws.release();
    if (op->buf) hipFree(op->buf);
    if (op->hbuf) hipFree(op->hbuf);
    if (op->csr_buf) hipFree(op->csr_buf);
    if (op->ev0) hipEventDestroy(op->ev0);
    if (op->ev1) hipEventDestroy(op->ev1);
    delete op;
    return MFD_OK;
}

// Operator from the caller's assembled A (the `A` of create_A/create_linsys,
// model.jl:209-246): a row-major device copy applied by k_spmv, Jacobi
// scaling from its diagonal, and a 1D unit grid of n/3 cells for the solver's
// DOF numbering (so T = 1 and P = Q A Q).  A that equals its transpose is
// complex-symmetric (SQMR, COCG, adjoint); otherwise BiCGSTAB serves it.
int op_create_csc_impl(mfd_ctx *ctx, const mfd_csc *A, mfd_op **out) {
    if (!A || A->m != A->n || A->n < 3 || A->n % 3) {
        set_error("mfd_op_create_csc: A must be square with n = 3·(cells) rows");
        return MFD_ERR_ARG;
    }
    const int64_t n = A->n, M = n / 3;
    std::vector<double> l0(M + 1), l1 = {0.0, 1.0};
    for (int64_t i = 0; i <= M; ++i) l0[i] = (double)i;
    mfd_grid g;
    memset(&g, 0, sizeof g);
    g.N[0] = M;
    g.N[1] = g.N[2] = 1;
    g.lprim[0] = l0.data();
    g.lprim[1] = g.lprim[2] = l1.data();
    g.omega_pml.re = 1.0;
    g.pml_m = 4.0;
    g.pml_lnR = -16.0;
    g.order_cmpfirst = 1;
    DevCsc dA;
    MFD_TRY(csc_upload(ctx, A, "A", dA));
    DevCsr R;
    MFD_TRY(csr_of(ctx, dA, R));
    mfd_op *op = new mfd_op();
    op->ctx = ctx;
    op->stream = ctx->stream;
    op->ft = MFD_EE;
    auto bail = [&](int code) { op_destroy_impl(op); return code; };
    int rc = op->G.create(&g, ctx->stream);
    if (rc) return bail(rc);
    hipStream_t s = ctx->stream;
    op->n = n;
    op->csr = 1;
    op->csr_buf = R.buf;
    op->csr_rowptr = R.rowptr;
    op->csr_colind = R.colind;
    op->csr_val = R.val;
    R.buf = nullptr;   // the operator owns it now
    op->tile64[0] = op->tile32[0] = 0;
    op->omega = mk<double>(0, 0);
    const size_t b64 = sizeof(c64) * n, b32 = sizeof(c32) * n;
    // q, qs, qo, T, T⁻¹ (FP64), q (FP32), flags
    if (hipMalloc(&op->buf, 5 * b64 + b32 + 256) != hipSuccess) { set_error("hipMalloc failed for operator arrays"); return bail(MFD_ERR_ALLOC); }
    c64 *q = (c64 *)op->buf, *qs = q + n, *qo = q + 2 * n, *t = q + 3 * n, *tinv = q + 4 * n;
    c32 *q32 = (c32 *)(q + 5 * n);
    int *flag = (int *)((char *)op->buf + 5 * b64 + b32 + 64);
    int fl[2] = {0, 0};
    if (hipMemsetAsync(flag, 0, 2 * sizeof(int), s) != hipSuccess) return bail(MFD_ERR_HIP);
    k_csr_jacobi<<<nb(n), BS, 0, s>>>(n, op->csr_rowptr, op->csr_colind, op->csr_val, q, flag);
    // symmetrizing scaling: breadth-first sweeps (checked every 16), a new
    // seed for each component, then the check
    void *tmp = nullptr;
    if (hipMalloc(&tmp, 2 * sizeof(c64) * n + 256) != hipSuccess) { set_error("hipMalloc failed"); return bail(MFD_ERR_ALLOC); }
    c64 *w = (c64 *)tmp, *w2 = w + n;
    int *chg = (int *)(w2 + n);
    unsigned long long *first = (unsigned long long *)(chg + 32);
    hipError_t se = hipMemsetAsync(w, 0, sizeof(c64) * n, s);
    int64_t sweeps = 0;
    while (se == hipSuccess) {
        // seed the first unscaled DOF (DOF 0 at the start)
        const unsigned long long none = ~0ull;
        hipMemcpyAsync(first, &none, sizeof none, hipMemcpyHostToDevice, s);
        k_sym_seed<<<nb(n), BS, 0, s>>>(n, w, first);
        unsigned long long f = none;
        hipMemcpyAsync(&f, first, sizeof f, hipMemcpyDeviceToHost, s);
        se = hipStreamSynchronize(s);
        if (se != hipSuccess || f == none) break;   // every DOF scaled
        k_sym_set1<<<1, 64, 0, s>>>(w, first, n);
        while (sweeps <= n) {
            hipMemsetAsync(chg, 0, 16 * sizeof(int), s);
            for (int k = 0; k < 16; ++k) {
                k_sym_sweep<<<nb(n), BS, 0, s>>>(n, op->csr_rowptr, op->csr_colind, op->csr_val, dA.colptr, dA.rowval,
                                                  dA.nzval, w, w2, chg + k);
                std::swap(w, w2);
            }
            sweeps += 16;
            int last = 1;
            hipMemcpyAsync(&last, chg + 15, sizeof(int), hipMemcpyDeviceToHost, s);
            se = hipStreamSynchronize(s);
            if (se != hipSuccess || last == 0) break;   // the component is covered
        }
    }
    k_sym_check<<<nb(n), BS, 0, s>>>(n, op->csr_rowptr, op->csr_colind, op->csr_val, dA.colptr, dA.rowval, dA.nzval, w,
                                     flag + 1);
    k_sym_scale<<<nb(n), BS, 0, s>>>(n, w, q, t, tinv, qs, qo);
    k_down<<<nb(n), BS, 0, s>>>(q, q32, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(fl, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(tmp);
    if (e != hipSuccess) { set_error("operator set-up failed"); return bail(MFD_ERR_HIP); }
    if (fl[0]) { set_error("mfd_op_create_csc: A has a zero or missing diagonal entry (Jacobi scaling undefined)"); return bail(MFD_ERR_ARG); }
    op->sym = fl[1] == 0;
    if (op->sym) {   // P = Q T A T⁻¹ Q is complex-symmetric: SQMR, COCG and the adjoint apply
        op->G.d64.tdof = t;
        op->G.d64.tdofinv = tinv;
        op->csr_qo = qo;
        op->d64.qs = qs;
    } else {         // T = 1 (BiCGSTAB)
        op->csr_qo = q;
        op->d64.qs = q;
    }
    op->d64.q = q;
    op->d32.q = q32;
    hipEventCreate(&op->ev0);
    hipEventCreate(&op->ev1);
    *out = op;
    return MFD_OK;
}

// Component window (the 2D and 1D models: te.jl:4-13 E = (Ex, Ey), tm.jl:4-13
// E = Ez, tem.jl:4-13 E = Ex): the unknowns are the E components c0 ..
// c0+nc−1 of the single-cell 3D operator, whose restriction A[w,w] to them is
// the model's create_A.  Vectors hold nc·M values, component-last (p + M·k
// for component c0 + k).  The tiled stencil reads the window's components
// (the others are zero) and writes only them; the per-DOF solver kernels
// find each unknown's component through the window.
int op_set_components_impl(mfd_op *op, int c0, int nc) {
    if (op->csr) { set_error("components: an operator built from a CSC matrix has no component layout"); return MFD_ERR_UNSUPPORTED; }
    if (c0 < 0 || nc < 1 || c0 + nc > 3) { set_error("components: need 0 <= c0, 1 <= nc, c0 + nc <= 3"); return MFD_ERR_ARG; }
    if (op->G.h.cmpfirst && nc < 3) { set_error("components: a window needs the component-last order (order_cmpfirst = 0)"); return MFD_ERR_ARG; }
    if ((op->has_mu || op->rev) && nc < 3) {
        set_error("components: a window is built for μ = 1 and boundft = EE on every axis");
        return MFD_ERR_UNSUPPORTED;
    }
    if (op->gd && nc < 3) { set_error("components: switch the grad-div form off first"); return MFD_ERR_UNSUPPORTED; }
    MFD_HIP(hipStreamSynchronize(op->stream));
    for (DGrid<double> *g : {&op->G.d64, &op->G.r64}) { g->c0 = c0; g->nc = nc; }
    op->G.d32.c0 = c0;
    op->G.d32.nc = nc;
    op->n = (int64_t)nc * op->G.h.M;
    op->win = nc < 3;
    // the grad-div form's node scratch and transforms take all components
    op->gd_ok = !op->win && !op->full && !op->has_mu && op->ee_all &&
                !(op->d64.om2.re == 0.0 && op->d64.om2.im == 0.0);
    if (op->win)
        for (int *t : {op->tile64, op->tile32})
            if (t[0] == 0) { t[0] = 1; t[1] = 4; t[2] = 1; }
    op->ws.release();   // sized for the old vector length
    return MFD_OK;
}

int op_set_graddiv_impl(mfd_op *op, int enable) {
    if (enable && !op->gd_ok) {
        set_error("grad-div form needs diagonal ε, μ = 1 and ω ≠ 0");
        return MFD_ERR_UNSUPPORTED;
    }
    op->gd = enable ? 1 : 0;
    hipStream_t s = op->stream;
    int64_t M = op->G.h.M;
    c64 *q = (c64 *)op->d64.q;
    if (op->gd) k_jacobi_gd<<<nb(M), BS, 0, s>>>(op->G.d64, op->d64, q);
    else if (op->has_mu) k_jacobi<true><<<nb(M), BS, 0, s>>>(op->G.d64, op->d64, q, 1);
    else k_jacobi<false><<<nb(M), BS, 0, s>>>(op->G.d64, op->d64, q, 1);
    k_qscale<<<nb(M), BS, 0, s>>>(op->G.d64, q, (c64 *)op->d64.qs);
    k_down<<<nb(3 * M), BS, 0, s>>>(q, (c32 *)op->d32.q, 3 * M);
    op->ws.release();   // captured graphs hold the old kernels
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipStreamSynchronize(s));
    return MFD_OK;
}

int op_apply_dev_impl(mfd_op *op, const void *x, void *y, int64_t nrhs, int prec) {
    if (nrhs < 1 || nrhs > 65535) { set_error("nrhs out of range"); return MFD_ERR_ARG; }
    MFD_TRY(hbuf_ensure(op, nrhs));
    if (prec == MFD_FP64) launch_apply<double, MODE_A, EPI_STORE>(op, (const c64 *)x, (c64 *)y, nullptr, nrhs, nullptr, nullptr, nullptr);
    else launch_apply<float, MODE_A, EPI_STORE>(op, (const c32 *)x, (c32 *)y, nullptr, nrhs, nullptr, nullptr, nullptr);
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipStreamSynchronize(op->stream));
    return MFD_OK;
}

int op_apply_impl(mfd_op *op, const mfd_c64 *x, mfd_c64 *y, int64_t nrhs) {
    int64_t n = op->n;
    size_t b = sizeof(c64) * n * nrhs;
    void *dx = nullptr, *dy = nullptr;
    MFD_HIP(hipMalloc(&dx, b));
    if (hipMalloc(&dy, b) != hipSuccess) { hipFree(dx); set_error("hipMalloc failed"); return MFD_ERR_ALLOC; }
    hipStream_t s = op->stream;
    int rc = MFD_OK;
    if (hipMemcpyAsync(dx, x, b, hipMemcpyHostToDevice, s) != hipSuccess) rc = MFD_ERR_HIP;
    if (!rc) rc = op_apply_dev_impl(op, dx, dy, nrhs, MFD_FP64);
    if (!rc && (hipMemcpyAsync(y, dy, b, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess))
        rc = MFD_ERR_HIP;
    hipFree(dx);
    hipFree(dy);
    return rc;
}

// average ms of one stencil launch (EPI_STORE): MODE_A or MODE_P, FP64 or FP32 vectors
static int time_apply(mfd_op *op, int64_t nrhs, int prec, int mode, int iters, double *ms) {
    MFD_TRY(hbuf_ensure(op, nrhs));
    int64_t n = op->n;
    size_t es = prec == MFD_FP64 ? sizeof(c64) : sizeof(c32);
    void *dx = nullptr, *dy = nullptr;
    MFD_HIP(hipMalloc(&dx, es * n * nrhs));
    if (hipMalloc(&dy, es * n * nrhs) != hipSuccess) { hipFree(dx); return MFD_ERR_ALLOC; }
    hipStream_t s = op->stream;
    hipMemsetAsync(dx, 0, es * n * nrhs, s);
    auto one = [&]() {
        if (prec == MFD_FP64) {
            if (mode == MODE_A) launch_apply<double, MODE_A, EPI_STORE>(op, (const c64 *)dx, (c64 *)dy, nullptr, nrhs, nullptr, nullptr, nullptr);
            else launch_apply<double, MODE_P, EPI_STORE>(op, (const c64 *)dx, (c64 *)dy, nullptr, nrhs, nullptr, nullptr, nullptr);
        } else {
            if (mode == MODE_A) launch_apply<float, MODE_A, EPI_STORE>(op, (const c32 *)dx, (c32 *)dy, nullptr, nrhs, nullptr, nullptr, nullptr);
            else launch_apply<float, MODE_P, EPI_STORE>(op, (const c32 *)dx, (c32 *)dy, nullptr, nrhs, nullptr, nullptr, nullptr);
        }
    };
    for (int w = 0; w < 3; ++w) one();
    hipEventRecord(op->ev0, s);
    for (int it = 0; it < iters; ++it) one();
    hipEventRecord(op->ev1, s);
    hipError_t e = hipEventSynchronize(op->ev1);
    float t = 0;
    hipEventElapsedTime(&t, op->ev0, op->ev1);
    *ms = t / iters;
    hipFree(dx);
    hipFree(dy);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return MFD_ERR_HIP; }
    return MFD_OK;
}

int op_bench_apply_impl(mfd_op *op, int64_t nrhs, int prec, int iters, double *ms) {
    return time_apply(op, nrhs, prec, prec == MFD_FP64 ? MODE_A : MODE_P, iters, ms);
}

// Scalars of an active iteration that never converges (ρ = 1, τ = 1, no
// iteration cap), for timing complete solver iterations back to back
__global__ void k_bench_scal(Scal *scal, int64_t nrhs) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= nrhs) return;
    Scal S = {};
    S.rho_re = 1.0;
    S.tau = S.cn = 1.0;
    S.itmax = 1LL << 62;
    scal[r] = S;
}

// Average duration of each kernel of one SQMR iteration (stencil with the
// qᵀPq epilogue, r update, d/z/q update), timed in context: `iters` complete
// iterations (finish kernels included) run back to back on the context stream
// with the solver's workspace and launch shapes, with events around the three
// main kernels only.
int op_bench_iteration_impl(mfd_op *op, int64_t nrhs, int prec, int iters, double *ms) {
    if (nrhs < 1 || iters < 1 || iters > 1000) { set_error("nrhs must be positive and iters in [1, 1000]"); return MFD_ERR_ARG; }
    prec = prec == MFD_MIXED ? MFD_MIXED : MFD_FP64;
    MFD_TRY(ws_ensure(op, nrhs, prec));
    Workspace &w = op->ws;
    hipStream_t s = op->stream;
    int64_t n = w.n;
    size_t es = prec == MFD_FP64 ? sizeof(c64) : sizeof(c32);
    // bounded nonzero contents (every byte 0x3c: ~1e-2 in FP32, ~2e-19 in FP64)
    for (void *v : {w.r, w.q, w.t, w.d, w.z}) MFD_HIP(hipMemsetAsync(v, 0x3c, es * n * nrhs, s));
    dim3 gv(ngrid(n), (unsigned)nrhs);
    const size_t ps = (size_t)nrhs * NPART * 4;
    std::vector<hipEvent_t> ev(6 * (size_t)iters, nullptr);
    int rc = MFD_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) { rc = MFD_ERR_HIP; break; }
    auto iteration = [&](hipEvent_t *E) {
        if (prec == MFD_FP64) launch_iter<double>(op, nrhs, E);
        else launch_iter<float>(op, nrhs, E);
    };
    if (!rc) {
        k_bench_scal<<<(unsigned)((nrhs + 63) / 64), 64, 0, s>>>(w.scal, nrhs);
        for (int k = 0; k < 3; ++k) iteration(nullptr);
        for (int it = 0; it < iters; ++it) iteration(&ev[6 * (size_t)it]);
        if (hipStreamSynchronize(s) != hipSuccess) rc = MFD_ERR_HIP;
    }
    if (!rc) {
        double acc[3] = {0, 0, 0};
        for (int it = 0; it < iters; ++it)
            for (int k = 0; k < 3; ++k) {
                float t = 0;
                hipEventElapsedTime(&t, ev[6 * (size_t)it + 2 * k], ev[6 * (size_t)it + 2 * k + 1]);
                acc[k] += t;
            }
        for (int k = 0; k < 3; ++k) ms[k] = acc[k] / iters;
    }
    for (auto &e : ev)
        if (e) hipEventDestroy(e);
    if (rc) { set_error("kernel timing failed"); return rc; }
    MFD_HIP(hipGetLastError());
    return MFD_OK;
}

// Tile-shape autotuner: times the Krylov iteration's stencil (the Jacobi-scaled
// operator P on FP64 or complex-FP32 vectors) for each candidate (TY, KZ, tile
// order) and keeps the fastest per precision.
int op_autotune_impl(mfd_op *op, int64_t nrhs, int iters, int *best64, int *best32, double *ms_out) {
    if (op->csr) { set_error("autotune: an operator built from a CSC matrix has no stencil tiles"); return MFD_ERR_UNSUPPORTED; }
    if ((op->has_mu && op->full) || !(op->ee_all || op->rev)) {   // the direct kernel serves these operators
        op->tile64[0] = op->tile32[0] = 0;
        if (best64) best64[0] = best64[1] = best64[2] = 0;
        if (best32) best32[0] = best32[1] = best32[2] = 0;
        return MFD_OK;
    }
    // Each candidate is timed in context: complete SQMR steps (stencil, r
    // update, finish, d/z/q update) on the solver workspace with nonzero data,
    // scored by the whole step, so the caches hold what they hold inside a
    // solve (a tile order whose halo rows miss L2 costs the stencil and, through
    // the caches, the vector kernels around it).
    // 32-wide tiles (orders 2, 3) are candidates for the diagonal-ε stencil
    // where 64-wide tiles would leave lanes of the last x tile idle (with an x
    // size that is a multiple of 64 they only add halo columns: at 128³ they
    // tie on the Krylov step within noise and lose 19% on the A apply)
    const int tys[4] = {1, 2, 4, 8}, kzs[4] = {4, 8, 16, 32};
    const int norders = (op->G.h.N[0] % 64 != 0 && !op->full && !op->gd && !op->has_mu && !op->rev && !op->win) ? 4 : 2;
    double bestms[2] = {1e30, 1e30};
    int bt[2][3] = {{4, 16, 1}, {4, 16, 1}};
    const int its = std::max(2, std::min(iters, 50));
    for (int prec = 0; prec < 2; ++prec)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                for (int xo = 0; xo < norders; ++xo) {
                    int *tile = prec == 0 ? op->tile64 : op->tile32;
                    tile[0] = tys[a];
                    tile[1] = kzs[b];
                    tile[2] = xo;
                    double part[3] = {0, 0, 0};
                    MFD_TRY(op_bench_iteration_impl(op, nrhs, prec == 0 ? MFD_FP64 : MFD_MIXED, its, part));
                    const double ms = part[0] + part[1] + part[2];
                    if (ms < bestms[prec]) {
                        bestms[prec] = ms;
                        bt[prec][0] = tys[a];
                        bt[prec][1] = kzs[b];
                        bt[prec][2] = xo;
                    }
                }
    for (int k = 0; k < 3; ++k) {
        op->tile64[k] = bt[0][k];
        op->tile32[k] = bt[1][k];
        if (best64) best64[k] = bt[0][k];
        if (best32) best32[k] = bt[1][k];
    }
    op->ws.release();   // captured graphs hold the old launch shapes
    if (ms_out) { ms_out[0] = bestms[0]; ms_out[1] = bestms[1]; }
    return MFD_OK;
}

// ------------------------------------------------ COCG with reliable updates
//
// Conjugate orthogonal CG (van der Vorst & Melissen 1990) on the complex
// symmetric P, iterating on vectors of precision R with FP64 reductions.
// Reliable updates (Sleijpen & van der Vorst 1996; as in mixed-precision CG,
// Clark et al. 2010): when the recursive residual has fallen by δ below its
// running maximum, the iterate x is flushed into the FP64 solution e, the
// true residual b − A e is recomputed with the FP64 operator and replaces r;
// the search direction p is kept.  Convergence is decided on that FP64
// residual only.

// r −= α t;  partial sums of ||r||², rᵀr (k_fin_cocg finishes).  The iterate
// update x += α p of this step is deferred to the next step's stencil, which
// reads p anyway.
template <typename R>
__global__ __launch_bounds__(BS) void k_cocg_r(cx<R> *__restrict__ rall, const cx<R> *__restrict__ tall, int64_t n,
                                               int64_t vs, double *part, const Scal *scal) {
    const int rhs = blockIdx.y;
    const Scal &S = scal[rhs];
    if (S.done || S.breakdown) return;
    cx<R> *r = rall + rhs * vs;
    const cx<R> *t = tall + rhs * vs;
    c64 al = mk<double>(S.alpha_re, S.alpha_im);
    double acc[3] = {0, 0, 0};
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        c64 v = cvt<double>(r[e]) - al * cvt<double>(t[e]);
        r[e] = cvt<R>(v);
        acc[0] += (double)v.re * v.re + (double)v.im * v.im;
        acc[1] += (double)v.re * v.re - (double)v.im * v.im;
        acc[2] += 2.0 * (double)v.re * v.im;
    }
    block_partial<3>(acc, part + (size_t)rhs * NPART * 4, (int)blockIdx.x);
}

// after k_cocg_r: β, ρ, running max of ||r|| (one block per RHS)
__global__ __launch_bounds__(BS) void k_fin_cocg(const double *part, int nb, Scal *scal) {
    const int rhs = blockIdx.x;
    Scal &S = scal[rhs];
    if (S.done || S.breakdown) return;
    double tot[3];
    sum_partials<3>(part + (size_t)rhs * NPART * 4, nb, tot);
    if (threadIdx.x != 0) return;
    c64 rho1 = mk<double>(tot[1], tot[2]);
    if (S.rho_re == 0.0 && S.rho_im == 0.0) { S.breakdown = 1; S.beta_re = S.beta_im = 0; }
    else {
        c64 b = cdiv(rho1, mk<double>(S.rho_re, S.rho_im));
        S.beta_re = b.re;
        S.beta_im = b.im;
    }
    S.rho_re = rho1.re;
    S.rho_im = rho1.im;
    S.rnorm2 = tot[0];
    double rn = sqrt(tot[0]);
    if (rn > S.tau_ref) S.tau_ref = rn;   // running maximum since the last update
    S.it += 1;
}

// p = r + β p_old and the deferred x += α_prev p_old, for operators served by
// the direct stencil (the tiled stencil does this while staging)
template <typename R>
__global__ __launch_bounds__(BS) void k_cocg_p(cx<R> *__restrict__ pnall, const cx<R> *__restrict__ poall,
                                               const cx<R> *__restrict__ rall, cx<R> *__restrict__ xall, int64_t n,
                                               int64_t vs, const Scal *scal) {
    const int rhs = blockIdx.y;
    const Scal &S = scal[rhs];
    if (S.done || S.breakdown) return;
    cx<R> *pn = pnall + rhs * vs, *x = xall + rhs * vs;
    const cx<R> *po = poall + rhs * vs, *r = rall + rhs * vs;
    c64 be = mk<double>(S.beta_re, S.beta_im), ap = mk<double>(S.apend_re, S.apend_im);
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        c64 pv = cvt<double>(po[e]);
        pn[e] = cvt<R>(cvt<double>(r[e]) + be * pv);
        x[e] = cvt<R>(cvt<double>(x[e]) + ap * pv);
    }
}

// e64 += (Q T⁻¹)(x + α_pend p);  x = 0   (masked RHS; k_ru_replace then clears α_pend)
template <typename R>
__global__ __launch_bounds__(BS) void k_ru_flush(DGrid<double> g, OpData<double> od, c64 *__restrict__ eall,
                                                 cx<R> *__restrict__ xall, const cx<R> *__restrict__ pall, int64_t n,
                                                 int64_t vs, const int *mask, const Scal *scal) {
    const int rhs = blockIdx.y;
    if (!mask[rhs]) return;
    c64 *e64 = eall + rhs * vs;
    cx<R> *x = xall + rhs * vs;
    const cx<R> *pv = pall + rhs * vs;
    const c64 ap = mk<double>(scal[rhs].apend_re, scal[rhs].apend_im);
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        c64 xe = cvt<double>(x[e]) + ap * cvt<double>(pv[e]);
        if (xe.re == 0 && xe.im == 0) continue;
        int c;
        int64_t p;
        dof2cp(g, e, c, p);
        int i, j, k;
        cell3(g, p, i, j, k);
        c64 rf = cvt<double>(od.q[3 * p + c]) * tfac(g, c, i, j, k, true);
        e64[e] = e64[e] + rf * xe;
        x[e] = czero<R>();
    }
}

// r = (Q T) r64 in precision R (p = r as well on a restart); ρ = rᵀr, running
// max and scale reset (masked RHS)
template <typename R>
__global__ __launch_bounds__(BS) void k_ru_replace(DGrid<double> g, OpData<double> od, const c64 *__restrict__ r64all,
                                                   cx<R> *__restrict__ rall, cx<R> *__restrict__ pall, int64_t n,
                                                   int64_t vs, double *part, unsigned *cnt, Scal *scal,
                                                   const int *mask, const int *restart, const double *rn_true) {
    const int rhs = blockIdx.y;
    if (!mask[rhs]) return;
    const c64 *r64 = r64all + rhs * vs;
    cx<R> *r = rall + rhs * vs, *p = pall + rhs * vs;
    const bool rs = restart[rhs] != 0;
    double acc[3] = {0, 0, 0};
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        int c;
        int64_t pc;
        dof2cp(g, e, c, pc);
        int i, j, k;
        cell3(g, pc, i, j, k);
        c64 l = cvt<double>(od.q[3 * pc + c]) * tfac(g, c, i, j, k, false);
        c64 sv = l * r64[e];
        cx<R> v = mk<R>((R)sv.re, (R)sv.im);
        r[e] = v;
        if (rs) p[e] = v;
        acc[0] += (double)v.re * v.re + (double)v.im * v.im;
        acc[1] += (double)v.re * v.re - (double)v.im * v.im;
        acc[2] += 2.0 * (double)v.re * v.im;
    }
    double tot[3];
    if (!reduce_last<3>(acc, part + (size_t)rhs * NPART * 4, cnt + rhs, tot)) return;
    if (threadIdx.x != 0) return;
    Scal &S = scal[rhs];
    S.rho_re = tot[1];
    S.rho_im = tot[2];
    S.rnorm2 = tot[0];
    S.cn = sqrt(tot[0]);       // scaled norm of the replaced residual
    S.tau_ref = S.cn;          // running maximum restarts here
    S.tau = sqrt(rn_true[rhs]);   // its true (unscaled) norm
    S.breakdown = 0;
    S.apend_re = S.apend_im = 0;  // the flush took the pending update
    if (restart[rhs]) S.beta_re = S.beta_im = 0;   // next p = r
    if (S.cn == 0.0) S.done = 1;
}

__global__ void k_set_done(Scal *scal, const int *done, int64_t nrhs) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrhs && done[r]) scal[r].done = 1;
}

__global__ void k_scal_reset(Scal *scal, int64_t nrhs) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrhs) memset(&scal[r], 0, sizeof(Scal));
}

// COCG p buffers: step parity k reads p_old from buffer k and writes p to k^1
template <typename R>
static inline cx<R> *cocg_p(Workspace &w, int k) { return (cx<R> *)(k ? w.d : w.q); }

// one COCG step of parity par: p = r + β p_old (formed in the stencil), t = P p,
// σ = pᵀt, α; the deferred x += α_prev p_old; r −= α t; β, ρ
template <typename R>
static void launch_cocg_iter(mfd_op *op, int64_t nrhs, int par) {
    Workspace &w = op->ws;
    int64_t n = w.n;
    hipStream_t s = op->stream;
    dim3 gv(ngrid(n), (unsigned)nrhs);
    size_t pslot = (size_t)nrhs * NPART * 4;
    cx<R> *po = cocg_p<R>(w, par), *pn = cocg_p<R>(w, par ^ 1);
    const DGrid<R> &g = op->G.view<R>();
    int ty = std::is_same<R, double>::value ? op->tile64[0] : op->tile32[0];
    if (!op->csr && !op->full && !op->has_mu && !op->gd && op->ee_all && !op->win && ty > 0) {
        int kz = std::is_same<R, double>::value ? op->tile64[1] : op->tile32[1];
        // the fused COCG stencil is built 64-wide only: the narrow-tile bit is ignored
        int xr = (std::is_same<R, double>::value ? op->tile64[2] : op->tile32[2]) & 1;
        int nx = (g.N[0] + 63) / 64, ny = (g.N[1] + ty - 1) / ty, nch = (g.N[2] + kz - 1) / kz;
        while ((int64_t)nx * ny * nch > NPART) { kz *= 2; nch = (g.N[2] + kz - 1) / kz; }
        dim3 grid(nx, ny, (unsigned)(nch * nrhs));
        int64_t vs = op->n;
        const cx<R> *rr = (const cx<R> *)w.r;
        cx<R> *xi = (cx<R> *)w.z, *t = (cx<R> *)w.t;
        if (ty == 8)
            k_apply_tiled<R, MODE_P, EPI_DOT, 8, false, true><<<grid, 512, 0, s>>>(op->G.d64, op->d64, op->d64.om2, po, t, nullptr, vs, kz, nch, w.part, w.scal, nullptr, xr, rr, pn, xi);
        else if (ty == 2)
            k_apply_tiled<R, MODE_P, EPI_DOT, 2, false, true><<<grid, 128, 0, s>>>(op->G.d64, op->d64, op->d64.om2, po, t, nullptr, vs, kz, nch, w.part, w.scal, nullptr, xr, rr, pn, xi);
        else if (ty == 4)
            k_apply_tiled<R, MODE_P, EPI_DOT, 4, false, true><<<grid, 256, 0, s>>>(op->G.d64, op->d64, op->d64.om2, po, t, nullptr, vs, kz, nch, w.part, w.scal, nullptr, xr, rr, pn, xi);
        else
            k_apply_tiled<R, MODE_P, EPI_DOT, 1, false, true><<<grid, 64, 0, s>>>(op->G.d64, op->d64, op->d64.om2, po, t, nullptr, vs, kz, nch, w.part, w.scal, nullptr, xr, rr, pn, xi);
        launch_finish<EPI_DOT>(s, nx * ny * nch, nrhs, w.part, w.scal, nullptr, nullptr);
    } else {
        k_cocg_p<R><<<gv, BS, 0, s>>>(pn, po, (const cx<R> *)w.r, (cx<R> *)w.z, n, n, w.scal);
        launch_apply<R, MODE_P, EPI_DOT>(op, pn, (cx<R> *)w.t, nullptr, nrhs, w.part, w.scal, nullptr);
    }
    k_cocg_r<R><<<gv, BS, 0, s>>>((cx<R> *)w.r, (const cx<R> *)w.t, n, n, w.part + pslot, w.scal);
    k_fin_cocg<<<(unsigned)nrhs, BS, 0, s>>>(w.part + pslot, (int)gv.x, w.scal);
}

// ‖b_r‖² of every RHS into bn (host); a zero right-hand side has the exact
// solution 0, which is stored into its column of e here (whatever the initial
// guess), so every solve path then treats that RHS as converged
static int rhs_norms(mfd_op *op, const void *b, void *e, int64_t nrhs, double *bn) {
    Workspace &w = op->ws;
    hipStream_t s = op->stream;
    const int64_t n = w.n;
    const size_t pslot = (size_t)nrhs * NPART * 4;
    k_norm2<<<dim3(ngrid(n), (unsigned)nrhs), BS, 0, s>>>((const c64 *)b, n, n, w.part + 2 * pslot, w.cnt + 2 * nrhs, w.bn);
    MFD_HIP(hipGetLastError());
    MFD_HIP(hipMemcpyAsync(bn, w.bn, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
    MFD_HIP(hipStreamSynchronize(s));
    for (int64_t r = 0; r < nrhs; ++r)
        if (bn[r] == 0.0) MFD_HIP(hipMemsetAsync((c64 *)e + r * n, 0, sizeof(c64) * n, s));
    return MFD_OK;
}

// relative residual from ‖r‖² and ‖b‖² (0 for a zero right-hand side, whose
// column rhs_norms set to the exact solution; NaN stays NaN)
static inline double rel_residual(double rn2, double bn2) { return bn2 == 0.0 ? 0.0 : sqrt(rn2 / bn2); }

// largest relative residual, NaN if any is NaN
static inline double max_rel(const std::vector<double> &rel) {
    double m = 0.0;
    for (double v : rel) {
        if (std::isnan(v)) return v;
        m = std::max(m, v);
    }
    return m;
}

// COCG with reliable updates; b, e device FP64; returns iterations and updates
template <typename R>
static int run_cocg(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts &o, double delta,
                    long long *iters_out, long long *updates_out, double *worst_out, bool *conv_out,
                    bool bail_on_stall) {
    Workspace &w = op->ws;
    hipStream_t s = op->stream;
    int64_t n = w.n;
    dim3 gv(ngrid(n), (unsigned)nrhs);
    size_t pslot = (size_t)nrhs * NPART * 4;
    const DGrid<R> &g = op->G.view<R>();
    const OpData<R> &od = opd<R>(op);
    std::vector<double> bn(nrhs);
    MFD_TRY(rhs_norms(op, b, e, nrhs, bn.data()));
    k_scal_reset<<<(unsigned)((nrhs + 63) / 64), 64, 0, s>>>(w.scal, nrhs);
    MFD_HIP(hipMemsetAsync(w.z, 0, sizeof(cx<R>) * n * nrhs, s));
    MFD_HIP(hipMemsetAsync(w.q, 0, sizeof(cx<R>) * n * nrhs, s));
    MFD_HIP(hipMemsetAsync(w.d, 0, sizeof(cx<R>) * n * nrhs, s));
    int par = 0;   // parity of the next step (which p buffer holds the current p)
    std::vector<int> mask(nrhs, 1), restart(nrhs, 1), done(nrhs, 0), failed(nrhs, 0);
    std::vector<double> relres(nrhs, 1.0);
    std::vector<long long> last_ru(nrhs, 0);
    int *mask_d = w.active, *aux_d = w.active + nrhs;   // mask and restart/done flags on the device
    int *flush_d = w.active + 2 * nrhs;
    auto up_mask = [&]() -> const int * {
        hipMemcpyAsync(flush_d, mask.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, s);
        return flush_d;
    };
    auto residual = [&](double *worst) -> int {   // r64 = b − A e for masked RHS; marks converged ones
        MFD_HIP(hipMemcpyAsync(mask_d, mask.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, s));
        launch_apply<double, MODE_A, EPI_RESID>(op, (const c64 *)e, (c64 *)w.r64, (const c64 *)b, nrhs,
                                                w.part + 2 * pslot, nullptr, w.rn, mask_d);
        MFD_HIP(hipGetLastError());
        MFD_HIP(hipMemcpyAsync(w.rn_h, w.rn, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
        MFD_HIP(hipStreamSynchronize(s));
        std::vector<double> open_rel;
        for (int64_t r = 0; r < nrhs; ++r) {
            if (mask[r]) {
                double rel = rel_residual(w.rn_h[r], bn[r]);
                relres[r] = rel;
                if (rel <= o.tol) done[r] = 1;
                // a NaN residual stops that RHS without converging it
                if (std::isnan(rel)) done[r] = failed[r] = 1;
            }
            if (!done[r] || failed[r]) open_rel.push_back(relres[r]);
        }
        *worst = max_rel(open_rel);
        return MFD_OK;
    };
    std::vector<int> m2(nrhs);
    auto replace = [&]() -> int {   // r (and p on restart) from r64 for masked, unconverged RHS
        for (int64_t r = 0; r < nrhs; ++r) m2[r] = mask[r] && !done[r];
        MFD_HIP(hipMemcpyAsync(mask_d, m2.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, s));
        MFD_HIP(hipMemcpyAsync(aux_d, restart.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, s));
        k_ru_replace<R><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)w.r64, (cx<R> *)w.r, cocg_p<R>(w, par), n, n,
                                          w.part + 2 * pslot, w.cnt + 2 * nrhs, w.scal, mask_d, aux_d, w.rn);
        MFD_HIP(hipMemcpyAsync(aux_d, done.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, s));
        k_set_done<<<(unsigned)((nrhs + 63) / 64), 64, 0, s>>>(w.scal, aux_d, nrhs);
        MFD_HIP(hipGetLastError());
        return MFD_OK;
    };
    double worst = 0;
    MFD_TRY(residual(&worst));
    bool all = true;
    for (int64_t r = 0; r < nrhs; ++r) all = all && done[r];
    long long updates = 0, mx = 0, best_upd = 0;
    double best_worst = worst;
    bool stalled = false;
    if (!all) {
        MFD_TRY(replace());
        std::fill(restart.begin(), restart.end(), 0);
        int chunk = o.check_every > 0 ? o.check_every : 32;
        bool use_graph = o.use_graph != 0;
        if (use_graph && (chunk & 1)) ++chunk;   // a replayed graph must end on the parity it starts with
        const int kind = sizeof(R) == 4 ? 2 : 5;
        if (use_graph && (w.graph_iters != chunk || w.graph_kind != kind)) {
            if (w.graph) { hipGraphExecDestroy(w.graph); w.graph = nullptr; }
            hipGraph_t gr;
            MFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int it = 0; it < chunk; ++it) launch_cocg_iter<R>(op, nrhs, (par + it) & 1);
            MFD_HIP(hipStreamEndCapture(s, &gr));
            MFD_HIP(hipGraphInstantiate(&w.graph, gr, nullptr, nullptr, 0));
            hipGraphDestroy(gr);
            w.graph_iters = chunk;
            w.graph_kind = kind;
        }
        while (true) {
            if (use_graph) MFD_HIP(hipGraphLaunch(w.graph, s));
            else
                for (int it = 0; it < chunk; ++it) launch_cocg_iter<R>(op, nrhs, (par + it) & 1);
            par = (par + chunk) & 1;
            MFD_HIP(hipGetLastError());
            MFD_HIP(hipMemcpyAsync(w.scal_h, w.scal, sizeof(Scal) * nrhs, hipMemcpyDeviceToHost, s));
            MFD_HIP(hipStreamSynchronize(s));
            bool any = false;
            mx = 0;
            for (int64_t r = 0; r < nrhs; ++r) {
                const Scal &S = w.scal_h[r];
                if (S.it > mx) mx = S.it;
                mask[r] = 0;
                restart[r] = 0;
                if (done[r]) continue;
                double rn = sqrt(S.rnorm2);
                double est = S.cn > 0 ? rn / S.cn * S.tau : 0.0;   // estimated true residual norm
                // refresh when the recursive residual fell by δ below its running
                // maximum, when convergence looks near, on breakdown, and as a
                // guard when it has grown 1e4 past the last true residual or
                // 2000 iterations passed without a refresh
                bool trig = S.breakdown || rn < delta * S.tau_ref || est <= 0.5 * o.tol * sqrt(bn[r]) ||
                            S.it >= o.maxit || (delta > 0 && (rn > 1e4 * S.cn || S.it - last_ru[r] >= 2000));
                if (trig) { mask[r] = 1; restart[r] = S.breakdown || rn > 1e4 * S.cn; any = true; last_ru[r] = S.it; }
            }
            if (any) {
                k_ru_flush<R><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (c64 *)e, (cx<R> *)w.z, cocg_p<R>(w, par), n, n, up_mask(), w.scal);
                MFD_TRY(residual(&worst));
                // a stalling low-precision run hands over (the flush above put
                // everything it produced into e): three updates without a 2x
                // gain on the best true residual, or a 10x rise above it
                if (bail_on_stall) {
                    if (worst < 0.5 * best_worst) { best_worst = worst; best_upd = updates; }
                    else if (worst > 10 * best_worst || updates - best_upd >= 3) stalled = true;
                }
                MFD_TRY(replace());
                ++updates;
                if (getenv("MFD_SOLVE_TRACE"))
                    fprintf(stderr, "[mfd cocg] it %lld update %lld worst residual %.3e\n", mx, updates, worst);
                if (stalled) break;
            }
            all = true;
            for (int64_t r = 0; r < nrhs; ++r) all = all && done[r];
            if (all || mx >= o.maxit) break;
        }
        if (!all) {   // flush what the last iterations produced
            std::fill(mask.begin(), mask.end(), 1);
            k_ru_flush<R><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (c64 *)e, (cx<R> *)w.z, cocg_p<R>(w, par), n, n, up_mask(), w.scal);
            MFD_TRY(residual(&worst));
        }
    }
    worst = max_rel(relres);
    bool nan_rhs = false;
    for (int64_t r = 0; r < nrhs; ++r) nan_rhs = nan_rhs || failed[r];
    *iters_out = mx;
    *updates_out = updates;
    *worst_out = worst;
    *conv_out = all && !nan_rhs;
    return MFD_OK;
}

int op_set_tile_impl(mfd_op *op, int prec, int ty, int kz, int xcd) {
    if (op->csr) { set_error("tile: an operator built from a CSC matrix has no stencil tiles"); return MFD_ERR_UNSUPPORTED; }
    if (op->win && ty == 0) { set_error("tile: an operator with a component window runs the tiled stencil only (TY > 0)"); return MFD_ERR_UNSUPPORTED; }
    if (!(ty == 0 || ty == 1 || ty == 2 || ty == 4 || ty == 8) || kz < 1 || xcd < 0 || xcd > 3) {
        set_error("tile: TY in {0,1,2,4,8}, KZ >= 1, order in {0,1,2,3} (bit 0: XCD order, bit 1: 32-wide tiles)");
        return MFD_ERR_ARG;
    }
    int *t = prec == MFD_FP64 ? op->tile64 : op->tile32;
    t[0] = ty;
    t[1] = kz;
    t[2] = xcd;
    op->ws.release();
    return MFD_OK;
}

void solve_opts_default(mfd_solve_opts *o) {
    o->tol = 1e-8;
    o->maxit = 200000;
    o->precision = MFD_FP64;
    o->method = -1;
    o->inner_tol = 1e-5;
    o->check_every = 32;
    o->use_graph = 1;
}

// with MFD_SOLVE_TRACE set, a long cycle reports the largest recurrence
// residual estimate τ/‖c‖ of its running RHS every 512 steps
static void trace_progress(const Scal *scal_h, int64_t nrhs, long long done_it, long long it) {
    if (done_it % 512 != 0 || !getenv("MFD_SOLVE_TRACE")) return;
    double worst = 0;
    for (int64_t r = 0; r < nrhs; ++r)
        if (!scal_h[r].conv && scal_h[r].cn > 0) worst = std::max(worst, scal_h[r].tau / scal_h[r].cn);
    fprintf(stderr, "[mfd step] it %lld recurrence tau/cn %.3e\n", it, worst);
}

// stalled (optional): set when an FP32 cycle ended because its probes stopped
// improving, rather than at its target or budget
template <typename R>
static int run_inner(mfd_op *op, int64_t nrhs, const mfd_solve_opts *o, double rel_target, const int *active_d,
                     long long budget, long long *iters_out, const void *b = nullptr, const void *e = nullptr,
                     const double *bn = nullptr, bool *stalled = nullptr, double tol = 0, double cyc_res = 0) {
    if (stalled) *stalled = false;
    Workspace &w = op->ws;
    hipStream_t s = op->stream;
    int64_t n = w.n;
    dim3 gv(ngrid(n), (unsigned)nrhs);
    size_t pslot = (size_t)nrhs * NPART * 4;
    const DGrid<R> &g = op->G.view<R>();
    const OpData<R> &od = opd<R>(op);
    k_init<R><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)w.r64, (cx<R> *)w.r, (cx<R> *)w.q, (cx<R> *)w.d, (cx<R> *)w.z, n, n,
                                w.part + 2 * pslot, w.cnt + 2 * nrhs, w.scal, rel_target, active_d, budget);
    MFD_HIP(hipGetLastError());
    int chunk = o->check_every > 0 ? o->check_every : 32;
    bool use_graph = o->use_graph != 0;
    const int kind = sizeof(R) == 4 ? 1 : 3;
    if (use_graph && (w.graph_iters != chunk || w.graph_kind != kind)) {
        if (w.graph) { hipGraphExecDestroy(w.graph); w.graph = nullptr; }
        hipGraph_t gr;
        MFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int it = 0; it < chunk; ++it) launch_iter<R>(op, nrhs);
        MFD_HIP(hipStreamEndCapture(s, &gr));
        MFD_HIP(hipGraphInstantiate(&w.graph, gr, nullptr, nullptr, 0));
        hipGraphDestroy(gr);
        w.graph_iters = chunk;
        w.graph_kind = kind;
    }
    long long done_it = 0;
    // FP32 cycles probe the true FP64 residual every `probe` steps and end after
    // two probes in a row that fail to halve it
    const long long probe = std::max<long long>(chunk, 512) / chunk * chunk;
    const int stall_max = 2;
    double best_probe = 1e300;
    int stall = 0;
    // FP64 cycles (given tol and the cycle's starting residual cyc_res) end on
    // the true residual: the recurrence estimate τ runs to the cycle target
    // (0.1 tol / cyc_res, a 10x margin for the gap between τ and the true
    // residual), but once every running RHS has τ/‖c‖ · cyc_res ≤ 3 tol the
    // true FP64 residual is probed every `probe64` steps and the cycle ends as
    // soon as it is ≤ 0.9 tol for every RHS (the outer loop recomputes it the
    // same way).  A probe costs about one step.
    const long long probe64 = std::max<long long>(chunk, 256) / chunk * chunk;
    long long next64 = 0;
    while (true) {
        if (use_graph) MFD_HIP(hipGraphLaunch(w.graph, s));
        else
            for (int it = 0; it < chunk; ++it) launch_iter<R>(op, nrhs);
        MFD_HIP(hipGetLastError());
        MFD_HIP(hipMemcpyAsync(w.scal_h, w.scal, sizeof(Scal) * nrhs, hipMemcpyDeviceToHost, s));
        MFD_HIP(hipStreamSynchronize(s));
        done_it += chunk;
        bool all = true;
        long long mx = 0;
        for (int64_t r = 0; r < nrhs; ++r) {
            if (!w.scal_h[r].conv) all = false;
            if (w.scal_h[r].it > mx) mx = w.scal_h[r].it;
        }
        *iters_out = mx;
        if (all || mx >= budget) break;
        trace_progress(w.scal_h, nrhs, done_it, mx);
        bool probe_now = false;
        if (sizeof(R) == 4) probe_now = b && e && done_it % probe == 0;
        else if (b && e && tol > 0 && cyc_res > 0 && done_it >= next64) {
            bool near = true;
            for (int64_t r = 0; r < nrhs; ++r) {
                const Scal &S = w.scal_h[r];
                if (!S.conv && S.tau > 3.0 * (tol / cyc_res) * S.cn) near = false;
            }
            if (near) {
                probe_now = true;
                next64 = done_it + probe64;
            }
        }
        // FP32 cycles: probe the true FP64 residual every `probe` iterations and
        // end the cycle once it stops improving (2x per probe window); FP64
        // cycles: end once it meets the tolerance
        if (probe_now) {
            dim3 gv2(ngrid(n), (unsigned)nrhs);
            k_probe_x<R><<<gv2, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)e, (const cx<R> *)w.z, (c64 *)w.et, n, n);
            size_t ps = (size_t)nrhs * NPART * 4;
            launch_apply<double, MODE_A, EPI_RESID>(op, (const c64 *)w.et, (c64 *)w.et2, (const c64 *)b, nrhs,
                                                    w.part + 2 * ps, nullptr, w.rn);
            MFD_HIP(hipMemcpyAsync(w.rn_h, w.rn, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
            MFD_HIP(hipStreamSynchronize(s));
            double worst = 0;
            for (int64_t r = 0; r < nrhs; ++r) worst = std::max(worst, bn[r] > 0 ? sqrt(w.rn_h[r] / bn[r]) : 0.0);
            if (getenv("MFD_SOLVE_TRACE")) fprintf(stderr, "[mfd probe] it %lld true residual %.3e\n", mx, worst);
            if (sizeof(R) == 8) {
                if (worst <= 0.9 * tol) break;
                continue;
            }
            if (worst > 0.5 * best_probe) {
                if (++stall >= stall_max) {
                    if (stalled) *stalled = true;
                    break;
                }
            } else stall = 0;
            best_probe = std::min(best_probe, worst);
        }
    }
    return MFD_OK;
}

// ------------------------------------------------ BiCGSTAB (non-symmetric P)
//
// van der Vorst's BiCGSTAB on the Jacobi-scaled P, FP64, for operators that
// are not complex-symmetric (Bloch phases with k ≠ 0), where SQMR and COCG do
// not apply.  Vectors: r (w.r), r̂ (w.et), p (w.q), v (w.t), s (w.d),
// t (w.et2), x (w.z).  Hermitian inner products; the recurrence residual ‖r‖
// is compared with the cycle target as in SQMR.

// cycle start: r = r̂ = p = (Q T) r64, x = v = 0; ρ = r̂ᴴr, τ = ‖r‖
__global__ __launch_bounds__(BS) void k_bi_init(DGrid<double> g, OpData<double> od, const c64 *__restrict__ r64all,
                                                c64 *rall, c64 *rhall, c64 *pall, c64 *vall, c64 *xall, int64_t n,
                                                int64_t vs, double *part, unsigned *cnt, Scal *scal,
                                                double rel_target, const int *active, long long itmax) {
    const int rhs = blockIdx.y;
    const c64 *r64 = r64all + rhs * vs;
    c64 *r = rall + rhs * vs, *rh = rhall + rhs * vs, *pp = pall + rhs * vs, *v = vall + rhs * vs, *x = xall + rhs * vs;
    bool act = active[rhs] != 0;
    double acc[1] = {0};
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        int c;
        int64_t p;
        dof2cp(g, e, c, p);
        int i, j, k;
        cell3(g, p, i, j, k);
        c64 val = czero<double>();
        if (act) val = (od.q[3 * p + c] * tfac(g, c, i, j, k, false)) * r64[e];
        r[e] = val;
        rh[e] = val;
        pp[e] = val;
        v[e] = czero<double>();
        x[e] = czero<double>();
        acc[0] += val.re * val.re + val.im * val.im;
    }
    double tot[1];
    if (!reduce_last<1>(acc, part + (size_t)rhs * NPART * 4, cnt + rhs, tot)) return;
    if (threadIdx.x != 0) return;
    Scal &S = scal[rhs];
    memset(&S, 0, sizeof(Scal));
    S.tau = sqrt(tot[0]);
    S.cn = S.tau;
    S.rho_re = tot[0];
    S.target = rel_target * S.cn;
    S.itmax = itmax;
    if (!act || S.cn == 0.0) { S.conv = 1; S.done = 1; }
}

// partials of aᴴb (and ‖a‖² when NV = 3)
template <int NV>
__global__ __launch_bounds__(BS) void k_bi_dot(const c64 *__restrict__ aall, const c64 *__restrict__ ball, int64_t n,
                                               int64_t vs, double *part, const Scal *scal) {
    const int rhs = blockIdx.y;
    const Scal &S = scal[rhs];
    if (S.done || S.conv) return;
    const c64 *a = aall + rhs * vs, *b = ball + rhs * vs;
    double acc[NV];
    for (int k = 0; k < NV; ++k) acc[k] = 0;
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        c64 x = a[e], y = b[e];
        acc[0] += x.re * y.re + x.im * y.im;
        acc[1] += x.re * y.im - x.im * y.re;
        if (NV == 3) acc[2] += x.re * x.re + x.im * x.im;
    }
    block_partial<NV>(acc, part + (size_t)rhs * NPART * 4, (int)blockIdx.x);
}

// α = ρ / (r̂ᴴv)
__global__ __launch_bounds__(BS) void k_bi_fin1(const double *part, int nb, Scal *scal) {
    const int rhs = blockIdx.x;
    Scal &S = scal[rhs];
    if (S.done || S.conv) return;
    double tot[2];
    sum_partials<2>(part + (size_t)rhs * NPART * 4, nb, tot);
    if (threadIdx.x != 0) return;
    if (tot[0] == 0.0 && tot[1] == 0.0) { S.breakdown = 1; S.conv = 1; S.alpha_re = S.alpha_im = 0; return; }
    c64 a = cdiv(mk<double>(S.rho_re, S.rho_im), mk<double>(tot[0], tot[1]));
    S.alpha_re = a.re;
    S.alpha_im = a.im;
}

// s = r − α v
__global__ __launch_bounds__(BS) void k_bi_s(c64 *__restrict__ sall, const c64 *__restrict__ rall,
                                             const c64 *__restrict__ vall, int64_t n, int64_t vs, const Scal *scal) {
    const int rhs = blockIdx.y;
    const Scal &S = scal[rhs];
    if (S.done || S.conv) return;
    c64 *s = sall + rhs * vs;
    const c64 *r = rall + rhs * vs, *v = vall + rhs * vs;
    c64 al = mk<double>(S.alpha_re, S.alpha_im);
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) s[e] = r[e] - al * v[e];
}

// ω = tᴴs / tᴴt
__global__ __launch_bounds__(BS) void k_bi_fin2(const double *part, int nb, Scal *scal) {
    const int rhs = blockIdx.x;
    Scal &S = scal[rhs];
    if (S.done || S.conv) return;
    double tot[3];
    sum_partials<3>(part + (size_t)rhs * NPART * 4, nb, tot);
    if (threadIdx.x != 0) return;
    if (tot[2] == 0.0) { S.om_re = S.om_im = 0; return; }   // t = 0: s is the residual, ω unused
    S.om_re = tot[0] / tot[2];
    S.om_im = tot[1] / tot[2];
}

// x += α p + ω s;  r = s − ω t;  partials of r̂ᴴr and ‖r‖²
__global__ __launch_bounds__(BS) void k_bi_xr(c64 *__restrict__ xall, c64 *__restrict__ rall,
                                              const c64 *__restrict__ pall, const c64 *__restrict__ sall,
                                              const c64 *__restrict__ tall, const c64 *__restrict__ rhall, int64_t n,
                                              int64_t vs, double *part, const Scal *scal) {
    const int rhs = blockIdx.y;
    const Scal &S = scal[rhs];
    if (S.done || S.conv) return;
    c64 *x = xall + rhs * vs, *r = rall + rhs * vs;
    const c64 *pp = pall + rhs * vs, *s = sall + rhs * vs, *t = tall + rhs * vs, *rh = rhall + rhs * vs;
    c64 al = mk<double>(S.alpha_re, S.alpha_im), om = mk<double>(S.om_re, S.om_im);
    double acc[3] = {0, 0, 0};
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS) {
        c64 se = s[e];
        x[e] = x[e] + al * pp[e] + om * se;
        c64 rv = se - om * t[e];
        r[e] = rv;
        c64 h = rh[e];
        acc[0] += h.re * rv.re + h.im * rv.im;
        acc[1] += h.re * rv.im - h.im * rv.re;
        acc[2] += rv.re * rv.re + rv.im * rv.im;
    }
    block_partial<3>(acc, part + (size_t)rhs * NPART * 4, (int)blockIdx.x);
}

// β = (ρ_new/ρ)(α/ω), ρ, τ = ‖r‖, convergence
__global__ __launch_bounds__(BS) void k_bi_fin3(const double *part, int nb, Scal *scal) {
    const int rhs = blockIdx.x;
    Scal &S = scal[rhs];
    if (S.done || S.conv) return;
    double tot[3];
    sum_partials<3>(part + (size_t)rhs * NPART * 4, nb, tot);
    if (threadIdx.x != 0) return;
    c64 rho1 = mk<double>(tot[0], tot[1]), om = mk<double>(S.om_re, S.om_im);
    S.tau = sqrt(tot[2]);
    S.it += 1;
    if ((S.rho_re == 0.0 && S.rho_im == 0.0) || (om.re == 0.0 && om.im == 0.0)) {
        S.breakdown = 1;
        S.conv = 1;
        S.beta_re = S.beta_im = 0;
    } else {
        c64 b = cdiv(rho1, mk<double>(S.rho_re, S.rho_im)) * cdiv(mk<double>(S.alpha_re, S.alpha_im), om);
        S.beta_re = b.re;
        S.beta_im = b.im;
    }
    S.rho_re = rho1.re;
    S.rho_im = rho1.im;
    if (S.tau <= S.target || S.it >= S.itmax || S.tau == 0.0) S.conv = 1;
}

// p = r + β (p − ω v)
__global__ __launch_bounds__(BS) void k_bi_p(c64 *__restrict__ pall, const c64 *__restrict__ rall,
                                             const c64 *__restrict__ vall, int64_t n, int64_t vs, const Scal *scal) {
    const int rhs = blockIdx.y;
    const Scal &S = scal[rhs];
    if (S.done || S.conv) return;
    c64 *pp = pall + rhs * vs;
    const c64 *r = rall + rhs * vs, *v = vall + rhs * vs;
    c64 be = mk<double>(S.beta_re, S.beta_im), om = mk<double>(S.om_re, S.om_im);
    for (int64_t e = blockIdx.x * (int64_t)BS + threadIdx.x; e < n; e += (int64_t)gridDim.x * BS)
        pp[e] = r[e] + be * (pp[e] - om * v[e]);
}

static void launch_bicgstab_iter(mfd_op *op, int64_t nrhs) {
    Workspace &w = op->ws;
    int64_t n = w.n;
    hipStream_t s = op->stream;
    dim3 gv(ngrid(n), (unsigned)nrhs);
    const size_t ps = (size_t)nrhs * NPART * 4;
    c64 *r = (c64 *)w.r, *rh = (c64 *)w.et, *pp = (c64 *)w.q, *v = (c64 *)w.t, *sv = (c64 *)w.d, *t = (c64 *)w.et2,
        *x = (c64 *)w.z;
    launch_apply<double, MODE_P, EPI_STORE>(op, pp, v, nullptr, nrhs, nullptr, nullptr, nullptr);
    k_bi_dot<2><<<gv, BS, 0, s>>>(rh, v, n, n, w.part, w.scal);
    k_bi_fin1<<<(unsigned)nrhs, BS, 0, s>>>(w.part, (int)gv.x, w.scal);
    k_bi_s<<<gv, BS, 0, s>>>(sv, r, v, n, n, w.scal);
    launch_apply<double, MODE_P, EPI_STORE>(op, sv, t, nullptr, nrhs, nullptr, nullptr, nullptr);
    k_bi_dot<3><<<gv, BS, 0, s>>>(t, sv, n, n, w.part + ps, w.scal);
    k_bi_fin2<<<(unsigned)nrhs, BS, 0, s>>>(w.part + ps, (int)gv.x, w.scal);
    k_bi_xr<<<gv, BS, 0, s>>>(x, r, pp, sv, t, rh, n, n, w.part + 2 * ps, w.scal);
    k_bi_fin3<<<(unsigned)nrhs, BS, 0, s>>>(w.part + 2 * ps, (int)gv.x, w.scal);
    k_bi_p<<<gv, BS, 0, s>>>(pp, r, v, n, n, w.scal);
}

// one BiCGSTAB cycle from the FP64 residual in w.r64; x accumulates in w.z.
// Given b, e and the host per-RHS state (‖b‖², best, active), the cycle
// probes the true FP64 residual of e + x every `probe` steps (w.r64, free
// once the cycle has started, holds the probed iterate; the residual vector
// itself is not stored): a column whose probe beats its best residual so far
// is copied to w.ebest and its best lowered, the cycle ends when every
// running column meets 0.9·tol on the probe, and it ends early — for a restart
// from the best iterate — when the probed residual has risen 10× above the
// cycle's best probe (BiCGSTAB's erratic phases, in which its updated
// residual also drifts from the true one).
static int run_inner_bicgstab(mfd_op *op, int64_t nrhs, const mfd_solve_opts *o, double rel_target,
                              const int *active_d, long long budget, long long *iters_out,
                              const void *b = nullptr, const void *e = nullptr, const double *bn = nullptr,
                              double *best = nullptr, const int *active_h = nullptr) {
    Workspace &w = op->ws;
    hipStream_t s = op->stream;
    int64_t n = w.n;
    dim3 gv(ngrid(n), (unsigned)nrhs);
    size_t pslot = (size_t)nrhs * NPART * 4;
    k_bi_init<<<gv, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)w.r64, (c64 *)w.r, (c64 *)w.et, (c64 *)w.q,
                                (c64 *)w.t, (c64 *)w.z, n, n, w.part + 2 * pslot, w.cnt + 2 * nrhs, w.scal, rel_target,
                                active_d, budget);
    MFD_HIP(hipGetLastError());
    int chunk = o->check_every > 0 ? o->check_every : 32;
    bool use_graph = o->use_graph != 0;
    long long done_it = 0;
    const long long probe = std::max<long long>(chunk, 2048) / chunk * chunk;
    double cyc_best = HUGE_VAL;   // smallest worst-column probe of this cycle
    if (use_graph && (w.graph_iters != chunk || w.graph_kind != 4)) {
        if (w.graph) { hipGraphExecDestroy(w.graph); w.graph = nullptr; }
        hipGraph_t gr;
        MFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int it = 0; it < chunk; ++it) launch_bicgstab_iter(op, nrhs);
        MFD_HIP(hipStreamEndCapture(s, &gr));
        MFD_HIP(hipGraphInstantiate(&w.graph, gr, nullptr, nullptr, 0));
        hipGraphDestroy(gr);
        w.graph_iters = chunk;
        w.graph_kind = 4;
    }
    while (true) {
        if (use_graph) MFD_HIP(hipGraphLaunch(w.graph, s));
        else
            for (int it = 0; it < chunk; ++it) launch_bicgstab_iter(op, nrhs);
        MFD_HIP(hipGetLastError());
        MFD_HIP(hipMemcpyAsync(w.scal_h, w.scal, sizeof(Scal) * nrhs, hipMemcpyDeviceToHost, s));
        MFD_HIP(hipStreamSynchronize(s));
        bool all = true;
        long long mx = 0;
        for (int64_t r = 0; r < nrhs; ++r) {
            if (!w.scal_h[r].conv) all = false;
            if (w.scal_h[r].it > mx) mx = w.scal_h[r].it;
        }
        *iters_out = mx;
        if (all || mx >= budget) break;
        done_it += chunk;
        trace_progress(w.scal_h, nrhs, done_it, mx);
        if (!(b && e && bn && best && active_h) || done_it % probe) continue;
        k_probe_x<double><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)e, (const c64 *)w.z, (c64 *)w.r64, n, n);
        launch_apply<double, MODE_A, EPI_RESID>(op, (const c64 *)w.r64, nullptr, (const c64 *)b, nrhs,
                                                w.part + 2 * pslot, nullptr, w.rn, active_d);
        MFD_HIP(hipGetLastError());
        MFD_HIP(hipMemcpyAsync(w.rn_h, w.rn, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
        MFD_HIP(hipStreamSynchronize(s));
        double worst = 0;
        bool done = true;
        for (int64_t r = 0; r < nrhs; ++r) {
            if (!active_h[r] || w.scal_h[r].conv) continue;
            const double rel = rel_residual(w.rn_h[r], bn[r]);
            if (rel < best[r]) {
                best[r] = rel;
                MFD_HIP(hipMemcpyAsync((c64 *)w.ebest + r * n, (const c64 *)w.r64 + r * n, sizeof(c64) * n,
                                       hipMemcpyDeviceToDevice, s));
            }
            worst = std::isnan(rel) || std::isnan(worst) ? NAN : std::max(worst, rel);
            if (!(rel <= 0.9 * o->tol)) done = false;
        }
        if (getenv("MFD_SOLVE_TRACE")) fprintf(stderr, "[mfd probe] it %lld true residual %.3e\n", mx, worst);
        if (done) break;
        if (std::isnan(worst) || worst > 10.0 * cyc_best) break;
        cyc_best = std::min(cyc_best, worst);
    }
    return MFD_OK;
}

// e (device, FP64, nrhs) solves A e = b (device) to o->tol.
int solve_dev_impl(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *oin, mfd_solve_stats *st) {
    mfd_solve_opts o;
    if (oin) o = *oin; else solve_opts_default(&o);
    if (nrhs < 1 || nrhs > 65535) { set_error("nrhs out of range"); return MFD_ERR_ARG; }
    int prec = o.precision == MFD_MIXED ? MFD_MIXED : MFD_FP64;
    MFD_TRY(ws_ensure(op, nrhs, prec));
    Workspace &w = op->ws;
    hipStream_t s = op->stream;
    int64_t n = w.n;
    size_t pslot = (size_t)nrhs * NPART * 4;
    dim3 gv(ngrid(n), (unsigned)nrhs);
    hipEventRecord(op->ev0, s);
    int method = o.method >= 0 ? o.method : (op->sym ? MFD_SQMR : MFD_BICGSTAB);
    if (method != MFD_SQMR && method != MFD_COCG && method != MFD_BICGSTAB) { set_error("unknown method"); return MFD_ERR_ARG; }
    if (!op->sym && method != MFD_BICGSTAB) {
        set_error("SQMR and COCG need a complex-symmetric operator; with a complex Bloch phase use BiCGSTAB");
        return MFD_ERR_UNSUPPORTED;
    }
    if (method == MFD_COCG) {
        long long it = 0, upd = 0;
        double worst = 0;
        bool conv = false;
        // δ: FP32 recurrences are refreshed once the residual has fallen 100x;
        // FP64 ones only when convergence is near
        double delta = prec == MFD_MIXED ? 1e-2 : 0.0;
        long long it32 = 0;
        int rc;
        if (prec == MFD_MIXED) {
            // FP32 recurrences until the true residual stops improving, then FP64
            // from the current iterate
            rc = run_cocg<float>(op, b, e, nrhs, o, delta, &it, &upd, &worst, &conv, true);
            if (rc) return rc;
            it32 = it;
            if (!conv && it < o.maxit) {
                mfd_solve_opts o2 = o;
                o2.maxit = o.maxit - it;
                long long it2 = 0, upd2 = 0;
                MFD_TRY(ws_ensure(op, nrhs, MFD_MIXED));
                rc = run_cocg<double>(op, b, e, nrhs, o2, 0.0, &it2, &upd2, &worst, &conv, false);
                it += it2;
                upd += upd2;
            }
        } else rc = run_cocg<double>(op, b, e, nrhs, o, delta, &it, &upd, &worst, &conv, false);
        if (rc) return rc;
        hipEventRecord(op->ev1, s);
        hipEventSynchronize(op->ev1);
        float ms = 0;
        hipEventElapsedTime(&ms, op->ev0, op->ev1);
        if (st) {
            st->iterations = it;
            st->outer_cycles = upd;
            st->relres = worst;
            st->seconds = ms * 1e-3;
            st->converged = conv;
            st->iterations_fp32 = it32;
        }
        if (!conv) { set_error("solve did not reach the tolerance within maxit"); return MFD_ERR_NOCONV; }
        return MFD_OK;
    }
    std::vector<double> bn(nrhs);
    MFD_TRY(rhs_norms(op, b, e, nrhs, bn.data()));
    std::vector<int> active(nrhs, 1);
    int *active_d = w.active;
    long long total_it = 0, cycles = 0, it32 = 0;
    double worst = 0, prev_worst = 0;
    // per RHS: the smallest FP64 residual seen at a cycle start (its iterate
    // is kept in ebest), the latest one, and whether it is NaN (that RHS stops)
    std::vector<double> best(nrhs, HUGE_VAL), rel(nrhs, HUGE_VAL);
    std::vector<char> failed(nrhs, 0);
    bool conv = false;
    // FP32 cycles run while each gains at least 10x on the FP64 residual and
    // ends at its target; a cycle that gains less, or whose probes stopped
    // improving (the FP32 recurrence stagnated; a restart from the refined
    // residual stagnates again), switches the rest of the solve to FP64 cycles
    int cyc_prec = method == MFD_BICGSTAB ? MFD_FP64 : prec;
    bool stalled32 = false, restored = false;
    while (true) {
        launch_apply<double, MODE_A, EPI_RESID>(op, (const c64 *)e, (c64 *)w.r64, (const c64 *)b, nrhs,
                                                w.part + 2 * pslot, nullptr, w.rn);
        MFD_HIP(hipGetLastError());
        MFD_HIP(hipMemcpyAsync(w.rn_h, w.rn, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
        MFD_HIP(hipStreamSynchronize(s));
        bool all = true, any_active = false;
        for (int64_t r = 0; r < nrhs; ++r) {
            rel[r] = rel_residual(w.rn_h[r], bn[r]);
            // a NaN residual is neither converged nor worth iterating on
            if (std::isnan(rel[r])) failed[r] = 1;
            const bool ok = rel[r] <= o.tol;
            active[r] = !ok && !failed[r];
            all = all && ok;
            any_active = any_active || active[r];
            // the iterate with the smallest residual at a cycle start is kept
            // per RHS: a solve that ends without converging returns it (a
            // cycle of a stagnating recurrence can raise the residual)
            if (rel[r] < best[r]) {
                best[r] = rel[r];
                MFD_HIP(hipMemcpyAsync((c64 *)w.ebest + r * n, (const c64 *)e + r * n, sizeof(c64) * n,
                                       hipMemcpyDeviceToDevice, s));
            }
        }
        worst = max_rel(rel);
        if (all) { conv = true; break; }
        if (!any_active || total_it >= o.maxit || cycles > 200) break;
        // BiCGSTAB: a column whose cycle ended above twice its best iterate (a
        // probe snapshot or an earlier cycle start) restarts from that iterate;
        // its residual is then recomputed, and equals the best exactly
        if (method == MFD_BICGSTAB && !restored) {
            for (int64_t r = 0; r < nrhs; ++r) {
                if (!active[r] || best[r] == HUGE_VAL || !(rel[r] > 2.0 * best[r])) continue;
                MFD_HIP(hipMemcpyAsync((c64 *)e + r * n, (const c64 *)w.ebest + r * n, sizeof(c64) * n,
                                       hipMemcpyDeviceToDevice, s));
                restored = true;
            }
            if (restored) continue;
        }
        restored = false;
        if (cyc_prec == MFD_MIXED && cycles > 0 && (worst > 0.1 * prev_worst || stalled32)) cyc_prec = MFD_FP64;
        prev_worst = worst;
        MFD_HIP(hipMemcpyAsync(active_d, active.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, s));
        if (op->gd) {
            // the inner system is A' z = r' with r' = r + (1/ω²) ε G W Div r
            k_gd_psi<<<gv, BS, 0, s>>>(op->G.d64, op->d64.wnode, (const c64 *)w.r64, (c64 *)w.psi, n);
            k_gd_rhs<<<gv, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)w.psi, (c64 *)w.r64, n, cinv(op->d64.om2));
        }
        // inner target, relative to this cycle's residual: enough to finish
        // (0.1 tol / worst), never looser than 1e-2 so every restart makes
        // progress, and in mixed precision never tighter than inner_tol
        // (FP32 iterations stagnate below that)
        double need = std::max(0.1 * o.tol / std::max(worst, 1e-300), 1e-15);
        double target = std::min(need, 1e-2);
        if (cyc_prec == MFD_MIXED) target = std::max(target, o.inner_tol);
        long long it = 0;
        int rc;
        if (method == MFD_BICGSTAB)
            rc = run_inner_bicgstab(op, nrhs, &o, target, active_d, o.maxit - total_it, &it, b, e, bn.data(),
                                    best.data(), active.data());
        else if (cyc_prec == MFD_MIXED)
            rc = run_inner<float>(op, nrhs, &o, target, active_d, o.maxit - total_it, &it, b, e, bn.data(), &stalled32);
        else
            rc = run_inner<double>(op, nrhs, &o, target, active_d, o.maxit - total_it, &it, b, e, bn.data(), nullptr,
                                   op->gd ? 0.0 : o.tol, worst);
        if (rc) return rc;
        total_it += it;
        if (cyc_prec == MFD_MIXED) it32 += it;
        ++cycles;
        if (getenv("MFD_SOLVE_TRACE")) {
            MFD_HIP(hipMemcpyAsync(w.scal_h, w.scal, sizeof(Scal) * nrhs, hipMemcpyDeviceToHost, s));
            MFD_HIP(hipStreamSynchronize(s));
            fprintf(stderr, "[mfd solve] cycle %lld prec %s residual %.3e target %.2e iterations %lld tau/cn %.3e stag %d\n",
                    cycles, cyc_prec == MFD_MIXED ? "fp32" : "fp64", worst, target, it,
                    w.scal_h[0].cn > 0 ? w.scal_h[0].tau / w.scal_h[0].cn : 0.0, w.scal_h[0].stag);
        }
        if (cyc_prec == MFD_MIXED) k_upd_e<float><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (c64 *)e, (const c32 *)w.z, n, n);
        else k_upd_e<double><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (c64 *)e, (const c64 *)w.z, n, n);
        MFD_HIP(hipGetLastError());
    }
    if (!conv) {
        // columns whose last residual is above their best (or NaN) go back to
        // the best iterate; converged columns keep theirs
        for (int64_t r = 0; r < nrhs; ++r) {
            if (!(rel[r] > best[r]) && !std::isnan(rel[r])) continue;
            if (best[r] == HUGE_VAL) continue;   // never finite: nothing to restore
            MFD_HIP(hipMemcpyAsync((c64 *)e + r * n, (const c64 *)w.ebest + r * n, sizeof(c64) * n,
                                   hipMemcpyDeviceToDevice, s));
            rel[r] = best[r];
        }
        worst = max_rel(rel);
    }
    hipEventRecord(op->ev1, s);
    hipEventSynchronize(op->ev1);
    float ms = 0;
    hipEventElapsedTime(&ms, op->ev0, op->ev1);
    if (st) {
        st->iterations = total_it;
        st->outer_cycles = cycles;
        st->relres = worst;
        st->seconds = ms * 1e-3;
        st->converged = conv;
        st->iterations_fp32 = it32;
    }
    if (!conv) { set_error("solve did not reach the tolerance within maxit"); return MFD_ERR_NOCONV; }
    return MFD_OK;
}

// copies each problem's solver scalars into one array (one readback per chunk)
__global__ void k_gather_scal(Scal *const *src, Scal *dst, int count) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) dst[p] = *src[p];
}

namespace {
struct BatchProb {
    mfd_op *op;
    const void *b;
    void *e;
    int prec;                  // precision of the current cycle
    double bn = 0, worst = 1, prev = 0, best_probe = 1e300;
    double best = HUGE_VAL;    // smallest FP64 residual at a cycle start (iterate kept in ebest)
    int stall = 0;
    long long it_total = 0, it32 = 0, cycles = 0, next_probe = 0, next64 = 0;
    bool conv = false, done = false, stalled32 = false;   // stalled32: an FP32 cycle stagnated
};
}  // namespace

// Independent problems (own grid, ε, ω; one RHS each) solved in lockstep: each
// chunk of SQMR iterations of every live problem is one hipGraph launch whose
// per-problem branches run on separate streams; between chunks the host reads
// all solver scalars in one copy and runs the FP64 refinement step of each
// problem whose inner cycle ended (same rules as solve_dev_impl's SQMR path).
int solve_batch_dev_impl(int count, mfd_op *const *ops, const void *const *b, void *const *e,
                         const mfd_solve_opts *oin, mfd_solve_stats *st) {
    if (count < 1 || count > 1024 || !ops || !b || !e) { set_error("count must be in [1, 1024] with non-null arrays"); return MFD_ERR_ARG; }
    mfd_solve_opts o;
    if (oin) o = *oin; else solve_opts_default(&o);
    if (o.method == MFD_COCG) { set_error("batched solve: only the SQMR method is implemented"); return MFD_ERR_UNSUPPORTED; }
    mfd_ctx *ctx = ops[0] ? ops[0]->ctx : nullptr;
    for (int p = 0; p < count; ++p) {
        if (!ops[p] || ops[p]->ctx != ctx || !b[p] || !e[p]) { set_error("batched solve: every problem needs an operator of the same context and device b, e"); return MFD_ERR_ARG; }
        for (int q = 0; q < p; ++q)
            if (ops[q] == ops[p]) { set_error("batched solve: an operator appears twice"); return MFD_ERR_ARG; }
        if (!ops[p]->sym) { set_error("batched solve: SQMR needs complex-symmetric operators (no complex Bloch phase)"); return MFD_ERR_UNSUPPORTED; }
    }
    const int prec0 = o.precision == MFD_MIXED ? MFD_MIXED : MFD_FP64;
    hipStream_t s = ctx->stream;
    std::vector<BatchProb> P(count);
    for (int p = 0; p < count; ++p) {
        P[p].op = ops[p];
        P[p].b = b[p];
        P[p].e = e[p];
        P[p].prec = prec0;
        MFD_TRY(ws_ensure(ops[p], 1, prec0));
    }
    const int nstreams = std::min(count, 32);
    std::vector<hipStream_t> streams(nstreams, nullptr);
    std::vector<hipEvent_t> joins(nstreams, nullptr);
    hipEvent_t fork = nullptr, tev0 = nullptr, tev1 = nullptr;
    Scal **src_d = nullptr;
    Scal *all_d = nullptr, *all_h = nullptr;
    hipGraphExec_t gexec = nullptr;
    int rc = MFD_OK;
    auto fail = [&](hipError_t err) { set_error(hipGetErrorString(err)); rc = MFD_ERR_HIP; };
#define BCHK(x) do { hipError_t err_ = (x); if (err_ != hipSuccess) { fail(err_); goto cleanup; } } while (0)
    {
        for (int i = 0; i < nstreams; ++i) {
            BCHK(hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking));
            BCHK(hipEventCreateWithFlags(&joins[i], hipEventDisableTiming));
        }
        BCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        BCHK(hipEventCreate(&tev0));
        BCHK(hipEventCreate(&tev1));
        BCHK(hipMalloc(&src_d, sizeof(Scal *) * count));
        BCHK(hipMalloc(&all_d, sizeof(Scal) * count));
        BCHK(hipHostMalloc((void **)&all_h, sizeof(Scal) * count, hipHostMallocDefault));
        {
            std::vector<Scal *> src(count);
            for (int p = 0; p < count; ++p) src[p] = ops[p]->ws.scal;
            BCHK(hipMemcpyAsync(src_d, src.data(), sizeof(Scal *) * count, hipMemcpyHostToDevice, s));
            BCHK(hipStreamSynchronize(s));
        }
        hipEventRecord(tev0, s);
        for (int p = 0; p < count; ++p) {
            Workspace &w = ops[p]->ws;
            dim3 gv(ngrid(w.n), 1);
            k_norm2<<<gv, BS, 0, s>>>((const c64 *)b[p], w.n, w.n, w.part + 2 * NPART * 4, w.cnt + 2, w.bn);
        }
        for (int p = 0; p < count; ++p) {
            BCHK(hipMemcpyAsync(ops[p]->ws.rn_h, ops[p]->ws.bn, sizeof(double), hipMemcpyDeviceToHost, s));
            BCHK(hipStreamSynchronize(s));
            P[p].bn = ops[p]->ws.rn_h[0];
            // b = 0: the exact solution is 0 whatever the initial guess
            if (P[p].bn == 0.0) BCHK(hipMemsetAsync(e[p], 0, sizeof(c64) * ops[p]->ws.n, s));
        }
        const int chunk = o.check_every > 0 ? o.check_every : 32;
        const long long probe_every = std::max<long long>(chunk, 512) / chunk * chunk;
        const long long probe64 = std::max<long long>(chunk, 256) / chunk * chunk;
        // FP64 residual of problem p; ends it or starts its next inner cycle
        auto start_cycle = [&](BatchProb &B) -> int {
            mfd_op *op = B.op;
            Workspace &w = op->ws;
            int64_t n = w.n;
            dim3 gv(ngrid(n), 1);
            launch_apply<double, MODE_A, EPI_RESID>(op, (const c64 *)B.e, (c64 *)w.r64, (const c64 *)B.b, 1,
                                                    w.part + 2 * NPART * 4, nullptr, w.rn);
            MFD_HIP(hipMemcpyAsync(w.rn_h, w.rn, sizeof(double), hipMemcpyDeviceToHost, s));
            MFD_HIP(hipStreamSynchronize(s));
            const double rel = rel_residual(w.rn_h[0], B.bn);
            B.worst = rel;
            if (rel <= o.tol) {
                B.conv = B.done = true;
                if (getenv("MFD_SOLVE_TRACE"))
                    fprintf(stderr, "[mfd batch] a problem of %lld cells converged after %lld iterations\n",
                            (long long)op->G.h.M, B.it_total);
                return MFD_OK;
            }
            if (rel < B.best) {   // the best iterate at a cycle start, returned if the problem does not converge
                B.best = rel;
                MFD_HIP(hipMemcpyAsync(w.ebest, B.e, sizeof(c64) * n, hipMemcpyDeviceToDevice, s));
            }
            if (std::isnan(rel) || B.it_total >= o.maxit || B.cycles > 200) {
                B.done = true;
                if (B.best < HUGE_VAL && !(rel <= B.best)) {
                    MFD_HIP(hipMemcpyAsync(B.e, w.ebest, sizeof(c64) * n, hipMemcpyDeviceToDevice, s));
                    B.worst = B.best;
                }
                return MFD_OK;
            }
            if (B.prec == MFD_MIXED && B.cycles > 0 && (rel > 0.1 * B.prev || B.stalled32)) B.prec = MFD_FP64;
            B.prev = rel;
            if (op->gd) {
                k_gd_psi<<<gv, BS, 0, s>>>(op->G.d64, op->d64.wnode, (const c64 *)w.r64, (c64 *)w.psi, n);
                k_gd_rhs<<<gv, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)w.psi, (c64 *)w.r64, n, cinv(op->d64.om2));
            }
            double need = std::max(0.1 * o.tol / std::max(rel, 1e-300), 1e-15);
            double target = std::min(need, 1e-2);
            if (B.prec == MFD_MIXED) target = std::max(target, o.inner_tol);
            const int one = 1;
            MFD_HIP(hipMemcpyAsync(w.active, &one, sizeof(int), hipMemcpyHostToDevice, s));
            long long budget = o.maxit - B.it_total;
            if (B.prec == MFD_MIXED)
                k_init<float><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)w.r64, (c32 *)w.r, (c32 *)w.q, (c32 *)w.d,
                                                (c32 *)w.z, n, n, w.part + 2 * NPART * 4, w.cnt + 2, w.scal, target, w.active, budget);
            else
                k_init<double><<<gv, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)w.r64, (c64 *)w.r, (c64 *)w.q, (c64 *)w.d,
                                                 (c64 *)w.z, n, n, w.part + 2 * NPART * 4, w.cnt + 2, w.scal, target, w.active, budget);
            MFD_HIP(hipGetLastError());
            MFD_HIP(hipStreamSynchronize(s));   // the pageable `one` above
            B.cycles += 1;
            B.best_probe = 1e300;
            B.stall = 0;
            B.next_probe = probe_every;
            B.next64 = 0;
            return MFD_OK;
        };
        // end of an inner cycle: e += (Q T⁻¹) z in FP64
        auto end_cycle = [&](BatchProb &B, long long it) {
            Workspace &w = B.op->ws;
            dim3 gv(ngrid(w.n), 1);
            if (B.prec == MFD_MIXED) k_upd_e<float><<<gv, BS, 0, s>>>(B.op->G.d64, B.op->d64, (c64 *)B.e, (const c32 *)w.z, w.n, w.n);
            else k_upd_e<double><<<gv, BS, 0, s>>>(B.op->G.d64, B.op->d64, (c64 *)B.e, (const c64 *)w.z, w.n, w.n);
            B.it_total += it;
            if (B.prec == MFD_MIXED) B.it32 += it;
        };
        for (int p = 0; p < count; ++p) {
            int r = start_cycle(P[p]);
            if (r) { rc = r; goto cleanup; }
        }
        std::vector<int> cap_prec;   // precision each problem was captured with (-1: not in the graph)
        int cap_live = 0;
        while (true) {
            int live = 0;
            bool recap = gexec == nullptr;
            for (int p = 0; p < count; ++p) {
                if (P[p].done) continue;
                ++live;
                if (cap_prec.empty() || cap_prec[p] != P[p].prec) recap = true;
            }
            if (live == 0) break;
            if (live < (3 * cap_live) / 4) recap = true;
            if (recap) {
                if (gexec) { hipGraphExecDestroy(gexec); gexec = nullptr; }
                cap_prec.assign(count, -1);
                hipGraph_t gr;
                BCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                hipEventRecord(fork, s);
                int k = 0;
                for (int p = 0; p < count; ++p) {
                    if (P[p].done) continue;
                    hipStream_t sp = streams[k % nstreams];
                    if (k < nstreams) hipStreamWaitEvent(sp, fork, 0);
                    ++k;
                    P[p].op->stream = sp;
                    for (int it = 0; it < chunk; ++it) {
                        if (P[p].prec == MFD_MIXED) launch_iter<float>(P[p].op, 1);
                        else launch_iter<double>(P[p].op, 1);
                    }
                    P[p].op->stream = s;
                    cap_prec[p] = P[p].prec;
                }
                for (int i = 0; i < std::min(k, nstreams); ++i) {
                    hipEventRecord(joins[i], streams[i]);
                    hipStreamWaitEvent(s, joins[i], 0);
                }
                hipError_t ce = hipStreamEndCapture(s, &gr);
                BCHK(ce);
                ce = hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0);
                hipGraphDestroy(gr);
                BCHK(ce);
                cap_live = live;
            }
            BCHK(hipGraphLaunch(gexec, s));
            k_gather_scal<<<(count + 63) / 64, 64, 0, s>>>(src_d, all_d, count);
            BCHK(hipMemcpyAsync(all_h, all_d, sizeof(Scal) * count, hipMemcpyDeviceToHost, s));
            BCHK(hipStreamSynchronize(s));
            for (int p = 0; p < count; ++p) {
                BatchProb &B = P[p];
                if (B.done || cap_prec[p] != B.prec) continue;
                const Scal &S = all_h[p];
                bool end = S.conv != 0;
                if (!end && B.prec == MFD_MIXED && S.it >= B.next_probe) {
                    // true FP64 residual of e + (Q T⁻¹) z: the FP32 cycle ends
                    // once two probes in a row fail to halve it
                    B.next_probe += probe_every;
                    Workspace &w = B.op->ws;
                    dim3 gv(ngrid(w.n), 1);
                    k_probe_x<float><<<gv, BS, 0, s>>>(B.op->G.d64, B.op->d64, (const c64 *)B.e, (const c32 *)w.z, (c64 *)w.et, w.n, w.n);
                    launch_apply<double, MODE_A, EPI_RESID>(B.op, (const c64 *)w.et, (c64 *)w.et2, (const c64 *)B.b, 1,
                                                            w.part + 2 * NPART * 4, nullptr, w.rn);
                    BCHK(hipMemcpyAsync(w.rn_h, w.rn, sizeof(double), hipMemcpyDeviceToHost, s));
                    BCHK(hipStreamSynchronize(s));
                    double pr = B.bn > 0 ? sqrt(w.rn_h[0] / B.bn) : 0.0;
                    if (pr > 0.5 * B.best_probe) {
                        end = ++B.stall >= 2;
                        if (end) B.stalled32 = true;
                    } else B.stall = 0;
                    B.best_probe = std::min(B.best_probe, pr);
                }
                if (!end && B.prec == MFD_FP64 && !B.op->gd && S.it >= B.next64 &&
                    S.tau * B.worst <= 3.0 * o.tol * S.cn) {
                    // FP64 cycle near the tolerance: it ends as soon as the true
                    // residual is <= 0.9 tol (the rule of the single-problem solve)
                    B.next64 = S.it + probe64;
                    Workspace &w = B.op->ws;
                    dim3 gv(ngrid(w.n), 1);
                    k_probe_x<double><<<gv, BS, 0, s>>>(B.op->G.d64, B.op->d64, (const c64 *)B.e, (const c64 *)w.z, (c64 *)w.et, w.n, w.n);
                    launch_apply<double, MODE_A, EPI_RESID>(B.op, (const c64 *)w.et, (c64 *)w.et2, (const c64 *)B.b, 1,
                                                            w.part + 2 * NPART * 4, nullptr, w.rn);
                    BCHK(hipMemcpyAsync(w.rn_h, w.rn, sizeof(double), hipMemcpyDeviceToHost, s));
                    BCHK(hipStreamSynchronize(s));
                    const double pr = B.bn > 0 ? sqrt(w.rn_h[0] / B.bn) : 0.0;
                    if (getenv("MFD_SOLVE_TRACE")) fprintf(stderr, "[mfd batch probe] it %lld true residual %.3e\n", S.it, pr);
                    end = pr <= 0.9 * o.tol;
                }
                if (end) {
                    end_cycle(B, S.it);
                    int r = start_cycle(B);
                    if (r) { rc = r; goto cleanup; }
                    if (B.done) {
                        // the captured graph may still hold this problem's
                        // branch: its kernels skip a problem marked done
                        const int one = 1;
                        BCHK(hipMemcpyAsync(B.op->ws.active, &one, sizeof(int), hipMemcpyHostToDevice, s));
                        k_set_done<<<1, 64, 0, s>>>(B.op->ws.scal, B.op->ws.active, 1);
                        BCHK(hipStreamSynchronize(s));   // the pageable `one`
                    }
                }
            }
        }
        hipEventRecord(tev1, s);
        BCHK(hipEventSynchronize(tev1));
        float ms = 0;
        hipEventElapsedTime(&ms, tev0, tev1);
        bool all = true;
        for (int p = 0; p < count; ++p) {
            if (!P[p].conv) all = false;
            if (st) {
                st[p].iterations = P[p].it_total;
                st[p].outer_cycles = P[p].cycles;
                st[p].relres = P[p].worst;
                st[p].seconds = ms * 1e-3;
                st[p].converged = P[p].conv;
                st[p].iterations_fp32 = P[p].it32;
            }
        }
        if (!all) { set_error("batched solve: some problems did not reach the tolerance within maxit"); rc = MFD_ERR_NOCONV; }
    }
cleanup:
#undef BCHK
    for (int p = 0; p < count; ++p) ops[p]->stream = s;
    hipStreamSynchronize(s);
    if (gexec) hipGraphExecDestroy(gexec);
    for (int i = 0; i < nstreams; ++i) {
        if (streams[i]) hipStreamDestroy(streams[i]);
        if (joins[i]) hipEventDestroy(joins[i]);
    }
    if (fork) hipEventDestroy(fork);
    if (tev0) hipEventDestroy(tev0);
    if (tev1) hipEventDestroy(tev1);
    if (src_d) hipFree(src_d);
    if (all_d) hipFree(all_d);
    if (all_h) hipHostFree(all_h);
    return rc;
}

// Host-pointer form of the batched solve: b[p], e[p] are host arrays of 3M_p values.
int solve_batch_impl(int count, mfd_op *const *ops, const mfd_c64 *const *b, mfd_c64 *const *e,
                     const mfd_solve_opts *o, mfd_solve_stats *st) {
    if (count < 1 || count > 1024 || !ops || !b || !e) { set_error("count must be in [1, 1024] with non-null arrays"); return MFD_ERR_ARG; }
    for (int p = 0; p < count; ++p)
        if (!ops[p] || !b[p] || !e[p]) { set_error("batched solve: null operator or array"); return MFD_ERR_ARG; }
    hipStream_t s = ops[0]->ctx->stream;
    std::vector<void *> db(count, nullptr), de(count, nullptr);
    int rc = MFD_OK;
    for (int p = 0; p < count && !rc; ++p) {
        size_t bytes = sizeof(c64) * ops[p]->n;
        if (hipMalloc(&db[p], bytes) != hipSuccess || hipMalloc(&de[p], bytes) != hipSuccess) { set_error("hipMalloc failed"); rc = MFD_ERR_ALLOC; break; }
        if (hipMemcpyAsync(db[p], b[p], bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(de[p], e[p], bytes, hipMemcpyHostToDevice, s) != hipSuccess) { set_error("copy to device failed"); rc = MFD_ERR_HIP; }
    }
    if (!rc) rc = solve_batch_dev_impl(count, ops, db.data(), de.data(), o, st);
    if (rc == MFD_OK || rc == MFD_ERR_NOCONV) {
        for (int p = 0; p < count; ++p) {
            size_t bytes = sizeof(c64) * ops[p]->n;
            if (hipMemcpyAsync(e[p], de[p], bytes, hipMemcpyDeviceToHost, s) != hipSuccess) rc = MFD_ERR_HIP;
        }
        if (hipStreamSynchronize(s) != hipSuccess) rc = MFD_ERR_HIP;
    }
    for (int p = 0; p < count; ++p) {
        if (db[p]) hipFree(db[p]);
        if (de[p]) hipFree(de[p]);
    }
    return rc;
}

int solve_impl(mfd_op *op, const mfd_c64 *b, mfd_c64 *e, int64_t nrhs, const mfd_solve_opts *o, mfd_solve_stats *st) {
    int64_t n = op->n;
    size_t bytes = sizeof(c64) * n * nrhs;
    void *db = nullptr, *de = nullptr;
    MFD_HIP(hipMalloc(&db, bytes));
    if (hipMalloc(&de, bytes) != hipSuccess) { hipFree(db); set_error("hipMalloc failed"); return MFD_ERR_ALLOC; }
    hipStream_t s = op->stream;
    int rc = MFD_OK;
    if (hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(de, e, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = MFD_ERR_HIP;
    if (!rc) rc = solve_dev_impl(op, db, de, nrhs, o, st);
    if (rc == MFD_OK || rc == MFD_ERR_NOCONV) {
        if (hipMemcpyAsync(e, de, bytes, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            rc = MFD_ERR_HIP;
    }
    hipFree(db);
    hipFree(de);
    return rc;
}

int solve_adjoint_impl(mfd_op *op, const mfd_c64 *gvec, mfd_c64 *lam, const mfd_solve_opts *o, mfd_solve_stats *st) {
    if (op->win) { set_error("adjoint solve: an operator with a component window (2D/1D model) is not supported"); return MFD_ERR_UNSUPPORTED; }
    if (!op->sym) {
        set_error("adjoint solve: the operator is not complex-symmetric (complex Bloch phase), so Aᵀ is not S A S⁻¹");
        return MFD_ERR_UNSUPPORTED;
    }
    int64_t n = 3 * op->G.h.M;
    size_t bytes = sizeof(c64) * n;
    void *db = nullptr, *de = nullptr;
    MFD_HIP(hipMalloc(&db, bytes));
    if (hipMalloc(&de, bytes) != hipSuccess) { hipFree(db); return MFD_ERR_ALLOC; }
    hipStream_t s = op->stream;
    int rc = MFD_OK;
    // Aᵀ λ = g  <=>  A (S⁻¹ λ) = S⁻¹ g
    if (hipMemcpyAsync(db, gvec, bytes, hipMemcpyHostToDevice, s) != hipSuccess || hipMemsetAsync(de, 0, bytes, s) != hipSuccess)
        rc = MFD_ERR_HIP;
    if (!rc) {
        k_scale_S<<<ngrid(n), BS, 0, s>>>(op->G.d64, (c64 *)db, n, 1);
        rc = solve_dev_impl(op, db, de, 1, o, st);
    }
    if (rc == MFD_OK || rc == MFD_ERR_NOCONV) {
        k_scale_S<<<ngrid(n), BS, 0, s>>>(op->G.d64, (c64 *)de, n, 0);
        if (hipMemcpyAsync(lam, de, bytes, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            rc = MFD_ERR_HIP;
    }
    hipFree(db);
    hipFree(de);
    return rc;
}

// ------------------------------------------------ field post-processing

// h = (i/ω) μ⁻¹ (Cₑ e + jₘ) at every H DOF (model.jl:278-279); jm may be null
template <bool MU>
__global__ __launch_bounds__(BS) void k_h_from_e(DGrid<double> g, OpData<double> od, const c64 *__restrict__ e,
                                                 const c64 *__restrict__ jm, c64 *__restrict__ h, c64 iw) {
    for (int64_t p = blockIdx.x * (int64_t)BS + threadIdx.x; p < g.M; p += (int64_t)gridDim.x * BS) {
        int i, j, k;
        cell3(g, p, i, j, k);
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            int64_t d = dofidx(g, u, p);
            c64 v = Hval<double, MODE_A, false>(g, od, e, u, i, j, k);
            if (jm) v = v + jm[d];
            if (MU) v = v * od.muinv[3 * p + u];
            h[d] = iw * v;
        }
    }
}

// H_c of a given H vector at (i,j,k) with one index possibly −1 or N: the
// ghost is the Bloch image (times the phase) or zero
__device__ inline c64 ldH(const DGrid<double> &g, const c64 *__restrict__ h, int c, int i, int j, int k) {
    c64 ph = cone<double>();
    int id[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (id[a] < 0) {
            if (!g.bloch[a]) return czero<double>();
            id[a] += g.N[a];
            ph = ph * g.phinv[a];
        } else if (id[a] >= g.N[a]) {
            if (!g.bloch[a]) return czero<double>();
            id[a] -= g.N[a];
            ph = ph * g.ph[a];
        }
    }
    int64_t p = id[0] + (int64_t)g.N[0] * (id[1] + (int64_t)g.N[1] * id[2]);
    return h[dofidx(g, c, p)] * ph;
}

// e = (−i/ω) ε⁻¹ (Cₘ h − jₑ) at every E DOF, diagonal ε (model.jl:283-284); je may be null
__global__ __launch_bounds__(BS) void k_e_from_h(DGrid<double> g, OpData<double> od, const c64 *__restrict__ h,
                                                 const c64 *__restrict__ je, c64 *__restrict__ e, c64 miw) {
    for (int64_t p = blockIdx.x * (int64_t)BS + threadIdx.x; p < g.M; p += (int64_t)gridDim.x * BS) {
        int i, j, k;
        cell3(g, p, i, j, k);
        int id[3] = {i, j, k};
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            int v = u == 0 ? 1 : (u == 1 ? 2 : 0), w = u == 0 ? 2 : (u == 1 ? 0 : 1);
            // ∂_a h_c in Cₘ's direction along a (backward where boundft is EE, forward where HH)
            auto dif = [&](int a, int c) -> c64 {
                int i1[3] = {i, j, k};
                const c64 h0 = ldH(g, h, c, i, j, k);
                if (g.bft[a] == EE) {
                    i1[a] -= 1;
                    return h0 - ldH(g, h, c, i1[0], i1[1], i1[2]);
                }
                i1[a] += 1;
                return ldH(g, h, c, i1[0], i1[1], i1[2]) - h0;
            };
            // (Cₘ h)_u = ∂_v h_w − ∂_w h_v
            c64 cc = g.sdlinv[g0type(g.bft, EE, v)][v][id[v]] * dif(v, w) -
                     g.sdlinv[g0type(g.bft, EE, w)][w][id[w]] * dif(w, v);
            int64_t d = dofidx(g, u, p);
            if (je) cc = cc - je[d];
            e[d] = miw * cdiv(cc, od.epsd[3 * p + u]);
        }
    }
}

// which = 0: out = h_from_e(in, src = jₘ); which = 1: out = e_from_h(in, src = jₑ)
int field_convert_impl(mfd_op *op, int which, const mfd_c64 *in, const mfd_c64 *src, mfd_c64 *out) {
    if (op->csr) {
        set_error("h_from_e/e_from_h: the operator holds only A; use mfd_h_from_e_csc / mfd_e_from_h_csc with Pμ, Cₑ / Pε, Cₘ");
        return MFD_ERR_UNSUPPORTED;
    }
    if (op->ft != MFD_EE) {
        set_error("h_from_e/e_from_h: operator of the EE formulation");
        return MFD_ERR_UNSUPPORTED;
    }
    if (op->win) { set_error("h_from_e/e_from_h: the field conversions take all three components (no component window)"); return MFD_ERR_UNSUPPORTED; }
    if (which == 1 && op->full) { set_error("e_from_h: only diagonal ε is supported (Pε \\ is a sparse solve otherwise)"); return MFD_ERR_UNSUPPORTED; }
    if (op->omega.re == 0.0 && op->omega.im == 0.0) { set_error("h_from_e/e_from_h need ω ≠ 0"); return MFD_ERR_ARG; }
    int64_t n = 3 * op->G.h.M;
    size_t bytes = sizeof(c64) * n;
    hipStream_t s = op->stream;
    void *din = nullptr, *dsrc = nullptr, *dout = nullptr;
    MFD_HIP(hipMalloc(&din, bytes));
    int rc = MFD_OK;
    if (hipMalloc(&dout, bytes) != hipSuccess || (src && hipMalloc(&dsrc, bytes) != hipSuccess)) { set_error("hipMalloc failed"); rc = MFD_ERR_ALLOC; }
    if (!rc && (hipMemcpyAsync(din, in, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
                (src && hipMemcpyAsync(dsrc, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess)))
        rc = MFD_ERR_HIP;
    if (!rc) {
        c64 iw = cdiv(mk<double>(0.0, 1.0), op->omega);   // i/ω
        unsigned gr = ngrid(op->G.h.M);
        if (which == 0) {
            if (op->has_mu) k_h_from_e<true><<<gr, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)din, (const c64 *)dsrc, (c64 *)dout, iw);
            else k_h_from_e<false><<<gr, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)din, (const c64 *)dsrc, (c64 *)dout, iw);
        } else {
            k_e_from_h<<<gr, BS, 0, s>>>(op->G.d64, op->d64, (const c64 *)din, (const c64 *)dsrc, (c64 *)dout,
                                         mk<double>(-iw.re, -iw.im));
        }
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = MFD_ERR_HIP;
    }
    hipFree(din);
    if (dout) hipFree(dout);
    if (dsrc) hipFree(dsrc);
    return rc;
}

int adjoint_grad_impl(mfd_op *op, const mfd_c64 *e, const mfd_c64 *lam, double *grad) {
    if (op->csr) { set_error("adjoint_grad: needs an operator built from a model (ω and the grid)"); return MFD_ERR_UNSUPPORTED; }
    if (op->win) { set_error("adjoint_grad: an operator with a component window (2D/1D model) is not supported"); return MFD_ERR_UNSUPPORTED; }
    int64_t n = 3 * op->G.h.M;
    void *d = nullptr;
    MFD_HIP(hipMalloc(&d, (2 * sizeof(c64) + sizeof(double)) * n));
    hipStream_t s = op->stream;
    c64 *de = (c64 *)d, *dl = de + n;
    double *dg = (double *)(dl + n);
    int rc = MFD_OK;
    if (hipMemcpyAsync(de, e, sizeof(c64) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dl, lam, sizeof(c64) * n, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = MFD_ERR_HIP;
    if (!rc) {
        k_grad<<<ngrid(n), BS, 0, s>>>(de, dl, op->omega * op->omega, dg, n);
        if (hipMemcpyAsync(grad, dg, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = MFD_ERR_HIP;
    }
    hipFree(d);
    return rc;
}

// ------------------------------------------------------------- create_b

namespace mfd {
// value of f_c / p_cc (f a current, p the divided parameter) at cell
// p + off·â (off ∈ {−1, 0, +1}): a Bloch wrap carries the phase, a non-Bloch
// ghost is zero
__device__ inline c64 div_at(const DGrid<double> &g, const c64 *__restrict__ f, const c64 *__restrict__ par, int c,
                             int i, int j, int k, int a, int off) {
    int r[3] = {i, j, k};
    c64 ph = mk<double>(1, 0);
    if (off) {
        r[a] += off;
        if (r[a] < 0 || r[a] >= g.N[a]) {
            if (!g.bloch[a]) return mk<double>(0, 0);
            r[a] -= off * g.N[a];
            ph = off > 0 ? g.ph[a] : g.phinv[a];
        }
    }
    int64_t q = r[0] + (int64_t)g.N[0] * (r[1] + (int64_t)g.N[1] * r[2]);
    return cdiv(f[dofidx(g, c, q)], par[q + g.M * (c + 3 * c)]) * ph;
}

// b = −Cₘ (Pμ \ jₘ) − iω jₑ (ft = EE, model.jl:263) or b = Cₑ (Pε \ jₑ) − iω jₘ
// (ft = HH, model.jl:266-268), one thread per cell.  Cₑ differences run
// forward along an axis whose boundft is EE and backward where it is HH, Cₘ
// the other way, each with the ∆l of its output grid.
__global__ __launch_bounds__(BS) void k_create_b(DGrid<double> g, int ft, const c64 *__restrict__ par,
                                                 const c64 *__restrict__ je, const c64 *__restrict__ jm, c64 iw,
                                                 int omz, c64 *__restrict__ b) {
    int64_t p = blockIdx.x * (int64_t)BS + threadIdx.x;
    if (p >= g.M) return;
    int i, j, k;
    cell3(g, p, i, j, k);
    int id[3] = {i, j, k};
    const c64 *f = ft == EE ? jm : je;   // the current under the curl
    const c64 *jd = ft == EE ? je : jm;  // the current in the −iω term
    // difference of f_c/p along a in the curl's direction
    auto dif = [&](int c, int a) -> c64 {
        const bool fwd = ft == EE ? g.bft[a] == HH : g.bft[a] == EE;   // Cₘ for EE, Cₑ for HH
        const c64 y0 = div_at(g, f, par, c, i, j, k, a, 0);
        return fwd ? div_at(g, f, par, c, i, j, k, a, 1) - y0 : y0 - div_at(g, f, par, c, i, j, k, a, -1);
    };
    const int gout = ft == EE ? EE : HH;   // the output field's grid (E for Cₘ, H for Cₑ)
    for (int u = 0; u < 3; ++u) {
        int v = u == 0 ? 1 : (u == 1 ? 2 : 0), w = u == 0 ? 2 : (u == 1 ? 0 : 1);
        c64 y = g.sdlinv[g0type(g.bft, gout, v)][v][id[v]] * dif(w, v) -
                g.sdlinv[g0type(g.bft, gout, w)][w][id[w]] * dif(v, w);
        int64_t e = dofidx(g, u, p);
        c64 out = ft == EE ? -y : y;
        if (!omz) out = out - iw * jd[e];
        b[e] = out;
    }
}
}  // namespace mfd

int create_b_impl(mfd_ctx *ctx, const mfd::DevGrid &G, int ft, mfd_c64 omega, const mfd_c64 *par, const mfd_c64 *je,
                  const mfd_c64 *jm, mfd_c64 *b) {
    using namespace mfd;
    int64_t M = G.h.M, n = 3 * M;
    hipStream_t s = ctx->stream;
    DevArray U;
    int pfull = 0;
    MFD_TRY(upload_param(ctx, G, par, U, &pfull));
    if (pfull) {
        set_error(ft == EE ? "create_b: Pμ must be diagonal (Pμ \\ jₘ)" : "create_b: Pε must be diagonal (Pε \\ jₑ)");
        return MFD_ERR_UNSUPPORTED;
    }
    void *d = nullptr;
    MFD_HIP(hipMalloc(&d, sizeof(c64) * 3 * n));
    c64 *dje = (c64 *)d, *djm = dje + n, *db = djm + n;
    int rc = MFD_OK;
    if (hipMemcpyAsync(dje, je, sizeof(c64) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(djm, jm, sizeof(c64) * n, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = MFD_ERR_HIP;
    if (!rc) {
        c64 iw = mk<double>(-omega.im, omega.re);
        int omz = omega.re == 0.0 && omega.im == 0.0;
        k_create_b<<<(unsigned)((M + BS - 1) / BS), BS, 0, s>>>(G.d64, ft, (const c64 *)U.p, dje, djm, iw, omz, db);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(b, db, sizeof(c64) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = MFD_ERR_HIP;
    }
    if (rc) set_error("create_b failed");
    hipFree(d);
    return rc;
}
