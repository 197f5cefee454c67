// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

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

// Compute units of the current device (cached per device ordinal).
inline int device_cus() {
    static std::atomic<int> cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cus[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cus[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

// Resident waves of a kernel launched with 256-thread blocks on the current
// device: CUs x resident blocks per CU (register-limited) x 4, the blocks per
// CU queried once per device into the caller's cache.
template <typename F>
inline int64_t resident_waves(F kernel, std::atomic<int> (&per_cu)[64]) {
    int dev = 0, nb = 1;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64) {
        nb = per_cu[dev].load(std::memory_order_relaxed);
        if (nb == 0) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb <= 0) {
                (void)hipGetLastError();  // not a launch error of the caller
                nb = 1;
            }
            per_cu[dev].store(nb, std::memory_order_relaxed);
        }
    }
    return int64_t(device_cus()) * nb * 4;
}

// Rows per segment of a balanced sweep.  A launch of W waves runs in rounds of
// R resident waves (CUs x blocks per CU x 4); with 64-row segments the last
// round can be mostly empty (8192^2, K = 5, FP64 at 244 VGPRs, two waves per
// SIMD: 18 816 waves = 9.2 rounds of 2 048).  Around the round count m0 of
// 64-row segments, each m in [m0 - 1, m0 + 1] takes the most segments that
// fit m rounds, and the m with the least m x (segment rows + lead-in rows)
// wins (there: 125 segments of 66 rows, 18 375 waves in 9 rounds, the grid a
// kernel trace of the launch shows).  Segments stay at least 16 rows long.
inline int balanced_seg_rows(int64_t nx, int64_t nstrips, int64_t resident, int lead) {
    const int64_t min_rows = nx < 16 ? nx : 16;
    const int64_t w64 = nstrips * ((nx + 63) / 64);
    const int64_t m0 = (w64 + resident - 1) / resident;
    int64_t best_rows = 64, best_cost = -1;
    for (int64_t m = m0 > 1 ? m0 - 1 : 1; m <= m0 + 1; ++m) {
        const int64_t segs = (m * resident) / nstrips;
        if (segs < 1) continue;
        int64_t rows = (nx + segs - 1) / segs;
        if (rows < min_rows) rows = min_rows;
        const int64_t waves = nstrips * ((nx + rows - 1) / rows);
        const int64_t rounds = (waves + resident - 1) / resident;
        const int64_t cost = rounds * (rows + lead);
        if (best_cost < 0 || cost < best_cost) best_cost = cost, best_rows = rows;
    }
    return int(best_rows);
}

inline bool valid_layout(int layout) {
    return layout == LBM_LAYOUT_SOA || layout == LBM_LAYOUT_ROWS;
}

inline bool valid_grid(int64_t nx, int64_t ny) {
    if (nx <= 0 || ny <= 0 || nx > kMaxSide || ny > kMaxSide) return false;
    // total element count must stay far from int64 overflow
    return nx <= (int64_t(1) << 40) / ny;
}

}  // namespace lbm
