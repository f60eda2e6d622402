// Shared definitions of libmfdfd: complex arithmetic, the device view of the
// staggered grid, and HIP error plumbing.  gfx950 only.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/mfdfd.h"

namespace mfd {

template <typename R>
struct cx {
    R re, im;
};
using c64 = cx<double>;
using c32 = cx<float>;

template <typename R>
__host__ __device__ inline cx<R> mk(R a, R b) { cx<R> z; z.re = a; z.im = b; return z; }
template <typename R>
__host__ __device__ inline cx<R> operator+(cx<R> a, cx<R> b) { return mk<R>(a.re + b.re, a.im + b.im); }
template <typename R>
__host__ __device__ inline cx<R> operator-(cx<R> a, cx<R> b) { return mk<R>(a.re - b.re, a.im - b.im); }
template <typename R>
__host__ __device__ inline cx<R> operator-(cx<R> a) { return mk<R>(-a.re, -a.im); }
template <typename R>
__host__ __device__ inline cx<R> operator*(cx<R> a, cx<R> b) {
    return mk<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__host__ __device__ inline cx<R> operator*(R s, cx<R> a) { return mk<R>(s * a.re, s * a.im); }
template <typename R>
__host__ __device__ inline cx<R> cdiv(cx<R> a, cx<R> b) {
    // Smith's algorithm (scaled) for a / b
    if (fabs(b.re) >= fabs(b.im)) {
        R t = b.im / b.re, den = b.re + b.im * t;
        return mk<R>((a.re + a.im * t) / den, (a.im - a.re * t) / den);
    } else {
        R t = b.re / b.im, den = b.re * t + b.im;
        return mk<R>((a.re * t + a.im) / den, (a.im * t - a.re) / den);
    }
}
template <typename R>
__host__ __device__ inline cx<R> cinv(cx<R> b) { return cdiv(mk<R>(1, 0), b); }
template <typename R>
__host__ __device__ inline bool iszero(cx<R> a) { return a.re == 0 && a.im == 0; }
template <typename R>
__host__ __device__ inline cx<R> cconj(cx<R> a) { return mk<R>(a.re, -a.im); }
template <typename R>
__host__ __device__ inline R cabs2(cx<R> a) { return a.re * a.re + a.im * a.im; }
// principal square root
template <typename R>
__host__ __device__ inline cx<R> csqrt_(cx<R> z) {
    R a = z.re, b = z.im;
    if (a == 0 && b == 0) return mk<R>(0, 0);
    R m = sqrt(a * a + b * b);
    R t = sqrt((m + fabs(a)) / 2);
    if (a >= 0) return mk<R>(t, b / (2 * t));
    return mk<R>(fabs(b) / (2 * t), b >= 0 ? t : -t);
}
template <typename R, typename S>
__host__ __device__ inline cx<R> cvt(cx<S> a) { return mk<R>((R)a.re, (R)a.im); }

enum { PRIM = 0, DUAL = 1, EE = 0, HH = 1 };

// Device view of the grid: 1D tables per (grid type, axis), all of length N[a].
template <typename R>
struct DGrid {
    int N[3];
    int64_t M;
    int bloch[3];
    int bft[3];
    int cmpfirst;
    cx<R> ph[3];           // e^{-i k L}
    cx<R> phinv[3];        // 1 / e^{-i k L}
    const cx<R> *sdl[2][3];     // stretched ∆l
    const cx<R> *sdlinv[2][3];  // 1 / stretched ∆l
    const cx<R> *ssq[2][3];     // principal sqrt of stretched ∆l
    const cx<R> *ssqinv[2][3];  // 1 / sqrt
    // per-DOF T and T⁻¹ ([3p + c]) where the metric is not separable (an
    // operator built from an assembled matrix); null: T from the tables above
    const cx<R> *tdof = nullptr;
    const cx<R> *tdofinv = nullptr;
    // axes along which this view is the reflection of the operator's grid
    // (bit a: index i of axis a is the caller's N_a−1−i); 0 for the grid itself
    int rmask = 0;
    // component window of the DOF vectors: components c0 .. c0+nc−1 only
    // (nc < 3 with the component-last order: the 2D and 1D models' unknowns)
    int c0 = 0, nc = 3;
};

__host__ __device__ inline int g0type(const int *bft, int ft, int a) {
    return (ft == EE) ? (bft[a] == EE ? PRIM : DUAL) : (bft[a] == HH ? PRIM : DUAL);
}
__host__ __device__ inline int gtype(const int *bft, int ft, int w, int a) {
    int g0 = g0type(bft, ft, a);
    return a == w ? 1 - g0 : g0;
}

// Host-side grid: positions, ∆l, ghosted coordinates (double) built from lprim.
struct HGrid {
    int64_t N[3], M;
    std::vector<double> l[2][3], lg[2][3], tlg[2][3], dl[2][3];
    double bnd[2][3], L[3];
    int bloch[3], bft[3], cmpfirst;
    int64_t npml[2][3];
    double om[2], m, lnR, kb[3];
};

int build_hgrid(const mfd_grid *g, HGrid *h);

// error plumbing
void set_error(const std::string &msg);
#define MFD_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            mfd::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                           __FILE__ + ":" + std::to_string(__LINE__));                  \
            return MFD_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)
#define MFD_TRY(call)              \
    do {                           \
        int rc_ = (call);          \
        if (rc_ != MFD_OK) return rc_; \
    } while (0)

}  // namespace mfd

struct mfd_ctx {
    int device;
    hipStream_t stream;
};
