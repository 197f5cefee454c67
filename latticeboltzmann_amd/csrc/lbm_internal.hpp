// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_d2q9.h"

namespace lbm {

constexpr int kBlock = 256;          // threads per block of the grid kernels
constexpr int kReduceBlocks = 1024;  // fixed block count of the two-pass reductions
constexpr int64_t kMaxSide = (int64_t(1) << 30);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Launch status: LBM_OK or the positive hipError_t of the launch.
inline int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LBM_OK : int(e);
}

inline bool overlaps(const void *a, int64_t abytes, const void *b, int64_t bbytes) {
    const char *pa = static_cast<const char *>(a), *pb = static_cast<const char *>(b);
    return pa < pb + bbytes && pb < pa + abytes;
}

// Second passes of the two-pass reductions (defined in d2q9_kernels.hip):
// fold np partials into *out with one block (sum: *out = bias + sum).
void launch_sum_pass2(const double *partial, int np, double bias, double *out, hipStream_t s);
void launch_max_pass2(const double *partial, int np, double *out, hipStream_t s);

inline bool valid_layout(int layout) {
    return layout == LBM_LAYOUT_SOA || layout == LBM_LAYOUT_ROWS;
}

inline bool valid_grid(int64_t nx, int64_t ny) {
    if (nx <= 0 || ny <= 0 || nx > kMaxSide || ny > kMaxSide) return false;
    // total element count must stay far from int64 overflow
    return nx <= (int64_t(1) << 40) / ny;
}

}  // namespace lbm
