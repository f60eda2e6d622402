// Device-resident grid tables (1D per grid type and axis).
#pragma once
#include "common.hpp"

namespace mfd {

const char *last_error();

struct DevGrid {
    HGrid h;
    void *buf = nullptr;
    int64_t total = 0;            // 2 * (Nx + Ny + Nz)
    const c64 *sdl64 = nullptr;   // [gt][axis][N] stretched ∆l, complex double
    DGrid<double> d64;
    DGrid<float> d32;
    int create(const mfd_grid *g, hipStream_t stream);
    void release();
    ~DevGrid();
    template <typename R> const DGrid<R> &view() const;
  private:
    void fill(DGrid<double> &d, c64 *const *t);
    void fill(DGrid<float> &d, c32 *const *t);
};
template <> inline const DGrid<double> &DevGrid::view<double>() const { return d64; }
template <> inline const DGrid<float> &DevGrid::view<float>() const { return d32; }

// DOF index of component c at cell p
template <typename R>
__host__ __device__ inline int64_t dofidx(const DGrid<R> &g, int c, int64_t p) {
    return g.cmpfirst ? (int64_t)c + 3 * p : p + g.M * c;
}

}  // namespace mfd
