// ComputeDt + CaculateMass of a solve's current state that allocates nothing
// and synchronises nothing (ibtfo_compute_dt_noalloc, include/ibtfo.h), for
// callers that cannot call hipMalloc: code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation.
//
// It runs the kernels of ibtfo_compute_dt / ibtfo_mass (k_reduce and the
// one-block finish) with the same partial-record layout, so the results are
// the same bit for bit, but keeps the per-block partials and the control
// record the finish writes in `scratch`, a fixed array in device memory, and
// picks the current state buffer on the device. The solve's own control block
// and partials are left untouched.
//
// `scratch` is a file-scope __device__ array: each GPU that runs the library
// holds its own copy, at its own address, until the process ends. Host code
// never uses the array's name as a device address: on the first call on a
// device, with that device current, hipGetSymbolAddress gives that device's
// address, kept per device, and that pointer is what the kernels get. Every
// call on a GPU shares that GPU's array: calls that overlap on one GPU (on
// different streams, or from different threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ibtfo_ctx.h"

namespace ibtfo {

constexpr int NOALLOC_PARTS = 4096;  // partial records: the grids ibtfo_compute_dt reduces in <= 4096 blocks
constexpr size_t NOALLOC_CTL = (sizeof(Ctl) + 255) / 256 * 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_CTL + sizeof(double) * PART_STRIDE * NOALLOC_PARTS;  // about 360 KB
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

// The fields of the scratch control record that the finish reads before it
// writes them (status is OR-ed) start at zero on every call.
__global__ void k_noalloc_begin(Ctl* c) {
    c->status = 0;
    c->dt = 0.0;
    c->mass_now = 0.0;
    c->done = 0;
}
__global__ void k_noalloc_end(const Ctl* c, double* out) {
    out[0] = c->dt;
    out[1] = c->mass_now;
    out[2] = c->status ? 1.0 : 0.0;
}

}  // namespace ibtfo

using namespace ibtfo;

extern "C" int ibtfo_compute_dt_noalloc(ibtfo_ctx* x, double* out) {
    if (!x || !out) { set_error("ibtfo_compute_dt_noalloc: null argument"); return IBTFO_E_ARG; }
    if (!x->prepared) { set_error("ibtfo_compute_dt_noalloc before ibtfo_prepare"); return IBTFO_E_STATE; }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != x->device || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("ibtfo_compute_dt_noalloc: the current device (%d) is not the solve's (%d)", dev, x->device);
        return IBTFO_E_ARG;
    }
    const ReduceArgs& base = x->ra;  // the arguments ibtfo_compute_dt launches with
    if (base.nblocks > NOALLOC_PARTS) {
        set_error("ibtfo_compute_dt_noalloc: %d x %d interior cells need %d partial records, more than the %d "
                  "this entry point holds (at most %d interior cells)", x->L.nx, x->L.ny, base.nblocks,
                  NOALLOC_PARTS, NOALLOC_PARTS * 256);
        return IBTFO_E_ARG;
    }
    unsigned char* p;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* q = nullptr;
            hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !q) {
                set_error("ibtfo_compute_dt_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
                return IBTFO_E_HIP;
            }
            g_scratch_ptr[dev] = (unsigned char*)q;
        }
        p = g_scratch_ptr[dev];
    }
    Ctl* ctl = (Ctl*)p;
    ReduceArgs r = base;
    r.ctl = ctl;
    r.mass_part = (double*)(p + NOALLOC_CTL);
    r.reset = 0;
    r.U = nullptr;
    r.ubuf[0] = x->state[0];
    r.ubuf[1] = x->state[1];
    r.cur_ctl = x->ctl;
    hipLaunchKernelGGL(k_noalloc_begin, dim3(1), dim3(1), 0, x->stream, ctl);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_reduce(r, x->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_noalloc_end, dim3(1), dim3(1), 0, x->stream, (const Ctl*)ctl, out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        set_error("ibtfo_compute_dt_noalloc: launch failed: %s", hipGetErrorString(e));
        return IBTFO_E_HIP;
    }
    return IBTFO_OK;
}
