// Solve entry point that allocates nothing: the solver workspace lives in a
// fixed per-device array instead of memory allocated per operator.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the
// operator is built beforehand (mfd_op_create allocates), and each
// mfd_solve_dev_noalloc call then runs the same SQMR / COCG / BiCGSTAB solve
// as mfd_solve_dev, launch for launch, on a workspace carved from `scratch`.
//
// `scratch` is one array per GPU, loaded with the library's code object and
// held until the process ends.  In host code its name is only a host-side
// stand-in: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device when a context is created on it
// (with that device current) and used for everything that touches the array.
// Every call on a GPU shares that GPU's array and finds the previous call's
// contents in it; calls that overlap on one GPU race.
#include "op.hpp"
#include <cstdio>
#include <mutex>
#include <vector>

void solve_opts_default(mfd_solve_opts *o);
int solve_dev_impl(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *oin, mfd_solve_stats *st);

namespace mfd {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;   // 64 MiB per GPU
constexpr int NOALLOC_MAX_DEV = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static void *scratch_dev[NOALLOC_MAX_DEV];
static std::mutex scratch_mu;

// Records the device address of `scratch` on the current device (`device`);
// called by mfd_ctx_create with that device current.
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEV) { set_error("device index beyond the scratch table"); return MFD_ERR_ARG; }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_dev[device]) return MFD_OK;
    void *ptr = nullptr;
    MFD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) { set_error("hipGetSymbolAddress returned no address for the scratch array"); return MFD_ERR_HIP; }
    scratch_dev[device] = ptr;
    return MFD_OK;
}

static void *scratch_of(int device) {
    std::lock_guard<std::mutex> lk(scratch_mu);
    return device >= 0 && device < NOALLOC_MAX_DEV ? scratch_dev[device] : nullptr;
}

}  // namespace mfd

using namespace mfd;

int solve_dev_noalloc_impl(mfd_op *op, const void *b, void *e, int64_t nrhs, const mfd_solve_opts *oin,
                           mfd_solve_stats *st) {
    if (nrhs < 1 || nrhs > 65535) { set_error("nrhs out of range"); return MFD_ERR_ARG; }
    const int64_t n = op->n;
    const size_t need = ws_bytes(n, nrhs);
    if (need > NOALLOC_SCRATCH_BYTES) {
        char msg[160];
        snprintf(msg, sizeof msg, "mfd_solve_dev_noalloc: the workspace needs %zu bytes, above the %zu-byte scratch array",
                 need, NOALLOC_SCRATCH_BYTES);
        set_error(msg);
        return MFD_ERR_ARG;
    }
    void *base = scratch_of(op->ctx->device);
    if (!base) { set_error("scratch array address not set up for this device"); return MFD_ERR_HIP; }
    mfd_solve_opts o;
    if (oin) o = *oin; else solve_opts_default(&o);
    o.use_graph = 0;   // instantiating a graph may allocate

    // host mirrors of the solver scalars (pageable host memory)
    std::vector<Scal> scal_h((size_t)nrhs);
    std::vector<double> rn_h((size_t)nrhs);
    Workspace w;
    ws_carve(w, base, n, nrhs);
    w.fixed = true;
    w.scal_h = scal_h.data();
    w.rn_h = rn_h.data();
    // the last-block reduction counters must start at zero; the array keeps
    // whatever the previous call (possibly of another size) left there
    MFD_HIP(hipMemsetAsync(w.cnt, 0, sizeof(unsigned) * 3 * nrhs, op->stream));

    Workspace saved = op->ws;
    op->ws = w;
    const int rc = solve_dev_impl(op, b, e, nrhs, &o, st);
    op->ws = saved;
    return rc;
}
