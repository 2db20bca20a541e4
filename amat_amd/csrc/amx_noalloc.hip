// amx_noalloc.hip -- the FP64 CG solve of amx_cg_solve_dev with its
// workspace in a fixed device array instead of buffers the library
// allocates, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The array holds, for problems of up to AMX_NOALLOC_MAX_DOFS DoFs: the
// reduction partials and counters, the two device scalar blocks, and the five
// CG vectors (r, z, p, q, 1/diag).  Every GPU the library runs on holds its
// own copy of it until the process ends; its device address is looked up once
// per device (hipGetSymbolAddress) on the first call there.  Every call on a
// GPU shares that GPU's array: calls that overlap on one GPU race, and each
// call clears the partials and counters it needs at zero before it starts.
#include <mutex>

#include "amx_op.hpp"

namespace {

constexpr size_t kScalarsBytes = 256;  // two CgScalars, padded to keep the vectors aligned
constexpr size_t kVecBytes = (size_t)AMX_NOALLOC_MAX_DOFS * sizeof(double);
constexpr size_t kScratchBytes = amx::kPartialsBytes + kScalarsBytes + 5 * kVecBytes;
static_assert(2 * sizeof(amx::CgScalars) <= kScalarsBytes, "scalar block");
static_assert(amx::kPartialsBytes % 256 == 0, "vector alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// the current device's copy of the array (the device must be current)
unsigned char* scratch_on(int device)
{
    AMX_REQUIRE(device >= 0 && device < kMaxDevices, AMX_INVALID_ARGUMENT, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        AMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        AMX_REQUIRE(p != nullptr, AMX_DEVICE_ERROR, "no device address for the solver scratch array");
        g_scratch[device] = static_cast<unsigned char*>(p);
    }
    return g_scratch[device];
}

}  // namespace

extern "C" int amx_cg_solve_noalloc_dev(amx_op* h, const double* b, double* x, double rtol, int max_it,
                                        amx_solve_stats* stats, void* stream)
{
    return amx::guarded(__func__, [&] {
        AMX_REQUIRE(h && b && x, AMX_INVALID_ARGUMENT, "NULL argument");
        AMX_REQUIRE(rtol > 0.0 && max_it > 0, AMX_INVALID_ARGUMENT, "rtol and max_it must be positive");
        amx::Op& op = h->op;
        AMX_REQUIRE(op.finalized, AMX_INVALID_STATE, "finalize() has not been called");
        AMX_REQUIRE(op.size == 1, AMX_NOT_IMPLEMENTED, "device solver is one-rank");
        AMX_REQUIRE(op.n_total <= AMX_NOALLOC_MAX_DOFS, AMX_INVALID_ARGUMENT,
                    "operator has " + std::to_string(op.n_total) + " DoFs; the fixed workspace holds at most " +
                        std::to_string(AMX_NOALLOC_MAX_DOFS));
        amx::DeviceGuard dg(op.device);
        hipStream_t s = op.pick(stream);
        unsigned char* base = scratch_on(op.device);
        amx::CgWork w;
        w.partials = reinterpret_cast<double*>(base);
        w.sc = reinterpret_cast<amx::CgScalars*>(base + amx::kPartialsBytes);
        double* vec = reinterpret_cast<double*>(base + amx::kPartialsBytes + kScalarsBytes);
        w.r = vec;
        w.z = vec + AMX_NOALLOC_MAX_DOFS;
        w.p = vec + 2 * (size_t)AMX_NOALLOC_MAX_DOFS;
        w.q = vec + 3 * (size_t)AMX_NOALLOC_MAX_DOFS;
        w.dinv = vec + 4 * (size_t)AMX_NOALLOC_MAX_DOFS;
        w.host = nullptr;  // scalars read back through a local, no pinned buffer
        // the reduction counters must start at zero whatever the last call left
        amx::launch_fill(w.partials, amx::kPartialsBytes / sizeof(double), true, 0.0, s);
        amx::fp64_cg(op, b, x, rtol, max_it, w, stats, s);
    });
}
