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

}  // namespace mfd
