// Matrix-free operator and device-resident Krylov solver.
#pragma once
#include "grid.hpp"

namespace mfd {

// Per-RHS solver scalars, kept on the device.
struct Scal {
    double rho_re, rho_im;     // rᵀr (bilinear)
    double sig_re, sig_im;     // qᵀ P q
    double alpha_re, alpha_im;
    double beta_re, beta_im;
    double tau, theta, cn, target;
    double c2t2, c2a_re, c2a_im;
    double rnorm2;             // scratch: last reduced norm²
    int conv, done, breakdown, stag;
    long long it, itmax, it_ref;
    double tau_ref;
    double apend_re, apend_im;   // COCG: α of the step whose iterate update is still pending
    double om_re, om_im;         // BiCGSTAB: ω
};

template <typename R>
struct OpData {
    const cx<R> *epsd;    // [M][3] diagonal ε at E_c locations
    const cx<R> *epso;    // [M][3] (xy, xz, yz) at primal corners, or null
    const cx<R> *muinv;   // [M][3] 1/μ_cc at H_c locations, or null
    const cx<R> *q;       // [M][3] Jacobi scaling diag(A)^-1/2 (of A' when the grad-div form is on)
    const cx<R> *qs;      // [M][3] q·T⁻¹: the tiled stencil's input scaling of P = Q T A T⁻¹ Q
    const cx<R> *wnode;   // [M] grad-div weight W = 1/ε̄² at the primal nodes
    cx<R> om2;            // ω²
};

struct Workspace {
    int64_t nrhs = 0, n = 0;
    int prec = -1;
    void *buf = nullptr;
    size_t bytes = 0;
    // inner vectors (precision of the inner solve), nrhs each
    void *r = nullptr, *q = nullptr, *t = nullptr, *d = nullptr, *z = nullptr;
    void *r64 = nullptr;      // outer FP64 residual
    void *psi = nullptr;      // FP64 node scratch for the grad-div residual transform
    void *et = nullptr, *et2 = nullptr;   // FP64 scratch for true-residual probes
    void *ebest = nullptr;    // FP64 iterate with the smallest residual at a cycle start
    double *part = nullptr;   // reduction partials [slot][rhs][NPART][4]
    unsigned *cnt = nullptr;  // [slot][rhs]
    Scal *scal = nullptr;     // [rhs]
    Scal *scal_h = nullptr;   // pinned host mirror
    double *bn = nullptr;     // [rhs] ||b||²
    double *rn = nullptr;     // [rhs] ||r||²
    double *rn_h = nullptr;   // pinned
    int *active = nullptr;    // [3 rhs] per-RHS masks and flags
    hipGraphExec_t graph = nullptr;
    int graph_iters = 0;
    int graph_kind = 0;       // 1/3 = SQMR FP32/FP64, 2 = COCG, 4 = BiCGSTAB iteration
    bool fixed = false;       // memory the solver must not free or reallocate (caller-provided)
    void release();
};

// Solver workspace layout (op.hip): its size for nrhs right-hand sides of n
// unknowns, and the carving of its device part from a 256-byte aligned base.
size_t ws_bytes(int64_t n, int64_t nrhs);
void ws_carve(Workspace &w, void *base, int64_t n, int64_t nrhs);

}  // namespace mfd

struct mfd_op {
    mfd_ctx *ctx;
    hipStream_t stream = nullptr;   // the context's stream; a batched solve re-points it while capturing
    mfd::DevGrid G;
    mfd::c64 omega;
    int full = 0, has_mu = 0;
    void *buf = nullptr;
    mfd::OpData<double> d64;
    mfd::OpData<float> d32;
    mfd::Workspace ws;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // tiled stencil shape (TY rows per block, KZ planes per chunk, and
    // XCD-contiguous tile order (1) or plain dispatch order (0)); TY = 0
    // selects the direct kernel
    int tile64[3] = {4, 16, 1};
    int tile32[3] = {4, 16, 1};
    int sym = 1;              // complex-symmetric up to S_E (no complex Bloch phase)
    int gd = 0;               // grad-div augmented inner operator in use
    int gd_ok = 0;            // eligible (diagonal ε, μ = 1)
    int ft = 0;               // formulation the caller asked for (MFD_EE, or MFD_HH served as the dual EE operator)
    void *hbuf = nullptr;     // H = μ⁻¹ Cₑ x scratch of the two-pass direct kernel (3M values per RHS)
    size_t hbuf_bytes = 0;
    int ee_all = 1;           // boundft = EE on every axis (the tiled stencils and field conversions need it)
    int rev = 0;              // axes with boundft = HH (bit a): the tiled stencil runs on G.r64, reflected along them
    int64_t n = 0;            // unknowns per right-hand side (3M, or nc·M with a component window)
    int win = 0;              // a component window is set (mfd_op_set_components)
    // an operator built from a caller's CSC matrix A (mfd_op_create_csc):
    // row-major copy of A, applied by a sparse product; its grid is a 1D unit
    // grid of n/3 cells that only provides the DOF numbering (T = 1)
    int csr = 0;
    void *csr_buf = nullptr;
    int64_t *csr_rowptr = nullptr;
    int32_t *csr_colind = nullptr;
    mfd::c64 *csr_val = nullptr;
    const mfd::c64 *csr_qo = nullptr;   // output scaling Q T of the sparse product
};
