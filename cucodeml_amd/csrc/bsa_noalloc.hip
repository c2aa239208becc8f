// Likelihood of one gene without any device allocation, for callers that
// cannot call hipMalloc: code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation.  The computation is bsa_lnl's (same
// schedule, same kernels, bitwise the same results); every device buffer it
// needs (P(t) matrices, tip tables, partials, exponents, per-site values,
// sums) lies in `scratch`, an array of kScratchBytes that each GPU the
// library runs on holds until the process ends.
//
// `scratch` names a host-side stand-in in host code: its device address on
// the current device comes from hipGetSymbolAddress, taken once per device at
// the entry point's first call there and kept per device; only that address
// is ever handed to kernels and copies.  Calls on one device are serialized
// by a lock held until their work on the caller's stream has finished, so
// calls from several threads do not overlap (their buffers would collide:
// each call lays its gene out from the start of the array, and the array
// keeps the previous call's contents).
#include "bsa_forest.h"

#include <hip/hip_runtime.h>

#include <mutex>

using namespace bsa;

// the device memory of every call (one copy per GPU)
static __device__ __align__(256) char scratch[BSA_NOALLOC_BYTES];

namespace {

constexpr size_t kScratchBytes = BSA_NOALLOC_BYTES;
constexpr int kMaxDevices = 64;

struct DeviceScratch {
    char *ptr = nullptr;   // device address of `scratch` on this device
    std::mutex mu;         // held for the whole of a call on this device
};
DeviceScratch g_scratch[kMaxDevices];
std::mutex g_scratch_init;

int scratch_of(int dev, DeviceScratch **out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(BSA_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_init);
    DeviceScratch &d = g_scratch[dev];
    if (!d.ptr) {
        HIPCHK(hipSetDevice(dev));
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(BSA_EDEVICE, "no device address for the scratch array");
        d.ptr = static_cast<char *>(p);
    }
    *out = &d;
    return BSA_OK;
}

}  // namespace

extern "C" int bsa_lnl_noalloc(const bsa_problem *p, const bsa_model_opts *opts, const double *x,
                               int np, double *neg_lnl, double *fhK, double *site_lnf,
                               void *stream)
{
    if (!p || !x || !neg_lnl) return fail(BSA_EINVAL, "null argument");
    if (p->precision != BSA_FP64 && p->precision != BSA_FP32)
        return fail(BSA_EINVAL, "unknown precision %d", p->precision);
    if (opts && opts->fix_omega && !(opts->omega_fix > 0))
        return fail(BSA_EINVAL, "omega_fix must be > 0");
    Gene g;
    int rc = gene_init(g, p, p->precision == BSA_FP32 ? 4 : 8);
    if (rc) return rc;
    if (opts) g.opts = *opts;
    if (np != gene_num_params(g))
        return fail(BSA_EINVAL, "np=%d, the model expects %d", np, gene_num_params(g));
    Forest f;
    if ((rc = forest_plan(f, {&g}, p->precision, p->kernel, p->tile))) return rc;
    if ((rc = forest_check(f))) return rc;
    // refuse what does not fit before touching the device
    const size_t need = forest_place_lnl(f, nullptr);
    if (need > kScratchBytes)
        return fail(BSA_EINVAL,
                    "the gene needs %zu bytes of device scratch (%d taxa, %d patterns, %s partials); "
                    "bsa_lnl_noalloc has %zu",
                    need, g.ns, g.npatt, p->precision == BSA_FP32 ? "FP32" : "FP64",
                    kScratchBytes);
    HIP_SHARED;
    DeviceScratch *d = nullptr;
    if ((rc = scratch_of(p->device, &d))) return rc;
    std::lock_guard<std::mutex> lk(d->mu);
    HIPCHK(hipSetDevice(p->device));
    f.dev = p->device;
    double lnl = 0;
    rc = forest_lnl_in_place(f, d->ptr, kScratchBytes, static_cast<hipStream_t>(stream), x, &lnl,
                             fhK, site_lnf);
    if (rc) return rc;
    *neg_lnl = -lnl;
    return BSA_OK;
}
