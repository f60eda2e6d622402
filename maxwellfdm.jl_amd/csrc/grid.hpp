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
    // The grid seen through the reflection i -> N_a−1−i of the axes in mask:
    // along those axes every 1D table reversed with its primal and dual types
    // swapped and the Bloch phase inverted; every boundft EE.  An operator
    // whose boundft is HH on exactly those axes is, in these coordinates and
    // with the E components along the reflected axes negated (E is a polar
    // vector), the all-EE operator: along a reflected axis each curl's
    // difference direction and ghost side flip, and the sign each derivative
    // picks up is absorbed by those components and by H (an axial vector).
    // Built by reflect(); null tables until then.
    DGrid<double> r64;
    void *rbuf = nullptr;
    int create(const mfd_grid *g, hipStream_t stream);
    int reflect(int mask, hipStream_t stream);
    void release();
    ~DevGrid();
    template <typename R> const DGrid<R> &view() const;
  private:
    void fill(DGrid<double> &d, c64 *const *t);
    void fill(DGrid<float> &d, c32 *const *t);
};
template <> inline const DGrid<double> &DevGrid::view<double>() const { return d64; }
template <> inline const DGrid<float> &DevGrid::view<float>() const { return d32; }

// DOF index of component c at cell p (component-last vectors start at the
// window's first component)
template <typename R>
__host__ __device__ inline int64_t dofidx(const DGrid<R> &g, int c, int64_t p) {
    return g.cmpfirst ? (int64_t)c + 3 * p : p + g.M * (c - g.c0);
}

}  // namespace mfd
