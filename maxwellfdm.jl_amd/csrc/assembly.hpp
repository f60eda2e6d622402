// CSC assembly entry points (device kernels, host buffers out).
#pragma once
#include "grid.hpp"

namespace mfd {

struct DevArray {
    void *p = nullptr;
    ~DevArray();
};

// Uploads a (Nx,Ny,Nz,3,3) complex array; the int after the array says
// whether any off-diagonal entry is nonzero (copied to *full_h if given).
int upload_param(mfd_ctx *ctx, const DevGrid &G, const mfd_c64 *parr, DevArray &d, int *full_h);

// colptr == nullptr: only *nnz is computed.
int curl_csc(mfd_ctx *ctx, const DevGrid &G, int which, int64_t *nnz, int64_t *colptr, int64_t *rowval,
             mfd_c64 *nzval);
// corner-mean operator of create_Mcs: which 0 = Mcₑ, 1 = Mcₘ
int mean_csc(mfd_ctx *ctx, const DevGrid &G, int which, int64_t *nnz, int64_t *colptr, int64_t *rowval,
             mfd_c64 *nzval);
int param_csc(mfd_ctx *ctx, const DevGrid &G, int which, const mfd_c64 *parr, int64_t *nnz, int64_t *colptr,
              int64_t *rowval, mfd_c64 *nzval);
// ft = EE: A = Cₘ (Pμ \ Cₑ) − ω² Pε; ft = HH: A = Cₑ (Pε \ Cₘ) − ω² Pμ
int A_csc(mfd_ctx *ctx, const DevGrid &G, int ft, mfd_c64 omega, const mfd_c64 *eps, const mfd_c64 *mu,
          int64_t *nnz, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval);

// A caller's CSC matrix on the device (1-based colptr/rowval as given).
struct DevCscV {
    int64_t m, n, nnz;
    const int64_t *colptr, *rowval;
    const c64 *nzval;
};
struct DevCsc {
    int64_t m = 0, n = 0, nnz = 0;
    int64_t *colptr = nullptr, *rowval = nullptr;
    c64 *nzval = nullptr;
    void *buf = nullptr;
    DevCscV view() const { return DevCscV{m, n, nnz, colptr, rowval, nzval}; }
    ~DevCsc();
};
// Row-major copy (0-based rowptr, 32-bit column indices ascending per row).
struct DevCsr {
    int64_t m = 0, n = 0, nnz = 0;
    int64_t *rowptr = nullptr;
    int32_t *colind = nullptr;
    c64 *val = nullptr;
    void *buf = nullptr;
    ~DevCsr();
};

// Uploads and validates a caller's CSC matrix (name for error messages).
int csc_upload(mfd_ctx *ctx, const mfd_csc *h, const char *name, DevCsc &d);
// Transpose of the storage: the CSR form of A.
int csr_of(mfd_ctx *ctx, const DevCsc &A, DevCsr &R);

// create_A / create_b / h_from_e / e_from_h of the reference (model.jl:222-284)
// from the caller's assembled matrices.  colptr == nullptr: only *nnz.
int A_from_csc(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
               const mfd_csc *Cm, int64_t *nnz, int64_t *colptr, int64_t *rowval, mfd_c64 *nzval);
int b_from_csc(mfd_ctx *ctx, int ft, mfd_c64 omega, const mfd_csc *Pe, const mfd_csc *Pm, const mfd_csc *Ce,
               const mfd_csc *Cm, const mfd_c64 *je, const mfd_c64 *jm, mfd_c64 *b);
// which 0: h = (i/ω) (P \ (C in + src)) with P = Pμ, C = Cₑ (h_from_e);
// which 1: e = (−i/ω) (P \ (C in − src)) with P = Pε, C = Cₘ (e_from_h); src may be null
int field_from_csc(mfd_ctx *ctx, int which, mfd_c64 omega, const mfd_csc *P, const mfd_csc *Cx,
                   const mfd_c64 *in, const mfd_c64 *src, mfd_c64 *out);

}  // namespace mfd
