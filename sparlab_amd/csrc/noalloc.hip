// K_ff x without any device allocation: the variant of slx_kff_apply for callers
// that cannot allocate (code under stream capture, tight serving loops, hosts
// that forbid hipMalloc). The staged copies of x and y, the per-block partial
// sums of x.Kx and the result live in one file-scope device array; its address
// on each device is taken once, when a context is created on that device.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <sstream>

#include "slx_internal.hpp"
#include "slx_kernels.hpp"
#include "slx_solver.hpp"

namespace slx {
namespace {

constexpr long long kScratchRows = 1LL << 21;     // unknowns staged per call (x and y)
constexpr long long kScratchPartials = 1LL << 18;  // per-block partial sums
constexpr long long kScratchDoubles = 2 * kScratchRows + kScratchPartials + 32;

// One copy per device, held until the process ends (35.7 MB).
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::array<double*, kMaxDevices> g_scratch{};
std::mutex g_scratch_mutex;

}  // namespace

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) fail(SLX_CONFIG_ERROR, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[(size_t)device]) {
    // the caller has made `device` current; the array's device address there
    void* p = nullptr;
    SLX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[(size_t)device] = static_cast<double*>(p);
  }
  return g_scratch[(size_t)device];
}

}  // namespace slx

extern "C" slx_status slx_kff_apply_noalloc(slx_context* ctx, const double* x, double* y, double* xkx,
                                            int* matrix_free) {
  using namespace slx;
  Context* c = reinterpret_cast<Context*>(ctx);
  if (!c) return SLX_CONFIG_ERROR;
  try {
    SLX_HIP(hipSetDevice(c->device));
    if (c->K.rows == 0) fail(SLX_SOLVER_ERROR, "K_ff is not assembled");
    const long long n = c->K.rows;
    if (n > kScratchRows) {
      std::ostringstream os;
      os << "slx_kff_apply_noalloc: " << n << " unknowns exceed the staging limit of " << kScratchRows
         << "; use slx_kff_apply";
      fail(SLX_CONFIG_ERROR, os.str());
    }
    const int nb = op_dot_blocks(*c);
    if (nb > kScratchPartials) fail(SLX_CONFIG_ERROR, "slx_kff_apply_noalloc: the product's block count exceeds "
                                                      "the scratch partials; use slx_kff_apply");
    double* base = g_scratch[(size_t)c->device];
    if (!base) fail(SLX_DEVICE_ERROR, "slx_kff_apply_noalloc: the scratch address of this device is unknown");
    double* dx = base;
    double* dy = base + kScratchRows;
    double* part = dy + kScratchRows;
    double* res = part + kScratchPartials;
    hipStream_t s = c->stream;
    // every block writes its partial and the reduction writes res: nothing to clear
    SLX_HIP(hipMemcpyAsync(dx, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, s));
    op_apply_dot(*c, dx, dy, part);
    k::reduce(nb, part, res, 0, s);
    SLX_HIP(hipMemcpyAsync(y, dy, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s));
    if (xkx) SLX_HIP(hipMemcpyAsync(xkx, res, sizeof(double), hipMemcpyDeviceToHost, s));
    SLX_HIP(hipStreamSynchronize(s));
    if (matrix_free) *matrix_free = c->mf ? 1 : 0;
    c->last_error.clear();
    return SLX_OK;
  } catch (const Error& e) {
    c->last_error = e.msg;
    return (slx_status)e.code;
  } catch (const std::exception& e) {
    c->last_error = e.what();
    return SLX_DEVICE_ERROR;
  }
}
