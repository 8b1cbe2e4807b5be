// selfcheck.hip -- device-side comparison of the shortened division used by
// the fast-division stage variants (div_inrange, mgad_common.hpp) with the
// IEEE division the compiler emits for '/', on caller-chosen operands.
#include <hip/hip_runtime.h>

#include "mgad_common.hpp"

namespace mgad {

// mode 0: a/b.  mode 1, 2: the WENO-Z weight 0.3*(1 + (a/b)^p), p = mode.
__global__ __launch_bounds__(256) void k_check_division(int mode, const double *__restrict__ a,
                                                        const double *__restrict__ b,
                                                        double *__restrict__ q_fast,
                                                        double *__restrict__ q_ieee, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const double x = a[i], y = b[i];
    if (mode == 0) {
        q_fast[i] = div_inrange(x, y);
        q_ieee[i] = x / y;
    } else if (mode == 1) {
        q_fast[i] = zweight<1, true>(kD2N, x, y, 0.0);
        q_ieee[i] = zweight<1, false>(kD2N, x, y, 0.0);
    } else {
        q_fast[i] = zweight<2, true>(kD2N, x, y, 0.0);
        q_ieee[i] = zweight<2, false>(kD2N, x, y, 0.0);
    }
}

}  // namespace mgad

extern "C" int mgad_check_division(hipStream_t stream, int mode, const double *a, const double *b,
                                   double *q_fast, double *q_ieee, size_t n) {
    using namespace mgad;
    if (mode < 0 || mode > 2 || !a || !b || !q_fast || !q_ieee || n == 0 || n > (1ull << 31)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    k_check_division<<<blocks, 256, 0, stream>>>(mode, a, b, q_fast, q_ieee, n);
    return check_launch();
}
