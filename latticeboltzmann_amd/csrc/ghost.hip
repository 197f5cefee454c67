// Ghost-layer D2Q9 flows: one fused pull step for the reference's Couette,
// Poiseuille, sliding-lid and obstacle-canal scripts
// (simulators/simple_flows/PoiseuilleFlow.py, slidingLid.py;
// simulators/experimantal_flows/obstacle_canal.py).
//
// Those scripts run, per step and on the whole array including its ghost
// layer: [pressure-periodic columns] -> np.roll streaming -> [pressure-periodic
// columns] -> override rules (bounce-back, moving walls, obstacle faces) ->
// moments and collision.  Here one thread owns one node (x, y) and computes
// its post-step populations directly from the previous state:
//
//   P(x, y)  pre-stream value: the state, or on columns 0 and nx-1 the
//            pressure-periodic inlet / outlet value (mode BEFORE_STREAM);
//   G_i(x,y) streamed value P_i(x - cx_i, y - cy_i) (periodic), or on columns
//            0 and nx-1 the inlet / outlet value built from the streamed
//            column (mode AFTER_STREAM);
//   H(x, y)  G with every rule whose target holds (x, y) applied in order; a
//            rule reads G at (x + dx, y + dy) (the host rejects rule tables in
//            which a rule reads what an earlier one writes, so G is what the
//            scripts read too);
//   out      H collided with the scripts' equilibrium.
//
// Interior nodes read their nine populations once (whole rows shifted by at
// most one element, coalesced along y); nodes next to the ghost layer or an
// obstacle read a few more.  Either population layout works; the
// row-interleaved one (the nine channel segments of an x-row adjacent) with
// 512-thread blocks streams about 8 % faster than f[i][k][l] with 256.  Every expression keeps the evaluation order of
// the scripts' numpy code, and the unit is built with -ffp-contract=off, so the
// FP64 and FP32 results are those of the scripts bit for bit.

#include "d2q9_grid.hpp"
#include "lbm_internal.hpp"

#pragma clang fp contract(off)

namespace lbm {
namespace {

constexpr int kGhostBlock = 512;

struct GhostParams {
    int nx, ny;
    int nrules, pbc_mode, flags;
    double omega, rho_in, rho_out;
    lbm_ghost_rule rules[LBM_GHOST_MAX_RULES];
};

__device__ __forceinline__ int cx_of(int i) {
    return int(i == 1 || i == 5 || i == 8) - int(i == 3 || i == 6 || i == 7);
}
__device__ __forceinline__ int cy_of(int i) {
    return int(i == 2 || i == 5 || i == 6) - int(i == 4 || i == 7 || i == 8);
}

// rho = sum g (sequential), u = (sum of the + channels - sum of the - channels) / rho
// (PoiseuilleFlow.py:49-53).
template <typename T>
__device__ __forceinline__ void sf_moments(const T g[9], T &rho, T &ux, T &uy) {
    rho = g[0];
#pragma unroll
    for (int i = 1; i < 9; ++i) rho = rho + g[i];
    ux = ((g[1] + g[5]) + g[8]) - ((g[3] + g[6]) + g[7]);
    ux = ux / rho;
    uy = ((g[2] + g[5]) + g[6]) - ((g[4] + g[7]) + g[8]);
    uy = uy / rho;
}

// The scripts' equilibrium (PoiseuilleFlow.py:25-43) with the density factors
// a0 = 2 rho / 9, a1 = rho / 18, a2 = rho / 36 given.
template <typename T>
__device__ __forceinline__ void sf_eq(T a0, T a1, T a2, T ux, T uy, T e[9]) {
    const T p3 = T(3) * (ux + uy);
    const T m3 = T(3) * (ux - uy);
    const T uu = T(3) * (ux * ux + uy * uy);
    const T x6 = T(6) * ux;
    const T y6 = T(6) * uy;
    const T xx9 = T(9) * ux * ux;
    const T yy9 = T(9) * uy * uy;
    const T xy9 = T(9) * ux * uy;
    e[0] = a0 * (T(2) - uu);
    e[1] = a1 * (T(2) + x6 + xx9 - uu);
    e[2] = a1 * (T(2) + y6 + yy9 - uu);
    e[3] = a1 * (T(2) - x6 + xx9 - uu);
    e[4] = a1 * (T(2) - y6 + yy9 - uu);
    e[5] = a2 * (T(1) + p3 + xy9 + uu);
    e[6] = a2 * (T(1) - m3 - xy9 + uu);
    e[7] = a2 * (T(1) - p3 + xy9 + uu);
    e[8] = a2 * (T(1) + m3 - xy9 + uu);
}

// Density factors of an array density (computed in T) ...
template <typename T>
__device__ __forceinline__ void eq_of_node(T rho, T ux, T uy, T e[9]) {
    sf_eq<T>(T(2) * rho / T(9), rho / T(18), rho / T(36), ux, uy, e);
}
// ... and of the scalar inlet / outlet density, a Python float: the factors
// are computed in double and then meet the arrays in T.
template <typename T>
__device__ __forceinline__ void eq_of_scalar(double rho, T ux, T uy, T e[9]) {
    sf_eq<T>(T(2.0 * rho / 9.0), T(rho / 18.0), T(rho / 36.0), ux, uy, e);
}

// v[i] and v[i] = x for a wave-uniform channel i, as branches on i, so the
// arrays stay in registers (an indexed access would put them in scratch).
template <typename T>
__device__ __forceinline__ T pick(const T v[9], int i) {
    switch (i) {
    case 1: return v[1];
    case 2: return v[2];
    case 3: return v[3];
    case 4: return v[4];
    case 5: return v[5];
    case 6: return v[6];
    case 7: return v[7];
    case 8: return v[8];
    default: return v[0];
    }
}
template <typename T>
__device__ __forceinline__ void put(T v[9], int i, T x) {
    switch (i) {
    case 0: v[0] = x; break;
    case 1: v[1] = x; break;
    case 2: v[2] = x; break;
    case 3: v[3] = x; break;
    case 4: v[4] = x; break;
    case 5: v[5] = x; break;
    case 6: v[6] = x; break;
    case 7: v[7] = x; break;
    default: v[8] = x; break;
    }
}

// Element (i, x, y) of an nx x ny population array: f[x][i][y] (Rows) or
// f[i][x][y]; the layout is a template parameter so the index folds.
template <bool Rows>
__device__ __forceinline__ int64_t index_of(int i, int x, int y, int nx, int ny) {
    return Rows ? (int64_t(x) * 9 + i) * ny + y : (int64_t(i) * nx + x) * ny + y;
}

template <typename T, bool Rows>
struct Grid {
    using value_type = T;
    const T *s;
    int nx, ny;
    __device__ __forceinline__ T at(int i, int x, int y) const {
        return s[index_of<Rows>(i, x, y, nx, ny)];
    }
};

// Inlet (x == 0) or outlet (x == nx-1) populations from the nine values g of
// the opposite interior column (nx-2 or 1): f_eq(rho_b, u) + (g - f_eq(g)),
// summed as eq_b + (g - eq) (BEFORE) or (eq_b + g) - eq (AFTER).
template <typename T, bool After>
__device__ __forceinline__ T pbc_value(const T g[9], double rho_b, int i) {
    T rho, ux, uy, e[9], eb[9];
    sf_moments<T>(g, rho, ux, uy);
    eq_of_node<T>(rho, ux, uy, e);
    eq_of_scalar<T>(rho_b, ux, uy, eb);
    const T gi = pick(g, i), ei = pick(e, i), bi = pick(eb, i);
    return After ? (bi + gi) - ei : bi + (gi - ei);
}

// Pre-stream value of channel i at (x, y).
template <typename Gr, typename T = typename Gr::value_type>
__device__ T load_p(const Gr &g, const GhostParams &p, int i, int x, int y) {
    if (p.pbc_mode == LBM_PBC_BEFORE_STREAM && (x == 0 || x == p.nx - 1)) {
        const int xs = x == 0 ? p.nx - 2 : 1;
        T v[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) v[j] = g.at(j, xs, y);
        return pbc_value<T, false>(v, x == 0 ? p.rho_in : p.rho_out, i);
    }
    return g.at(i, x, y);
}

// Streamed value of channel i at (x, y), periodic in x and y.
template <typename Gr, typename T = typename Gr::value_type>
__device__ __forceinline__ T pull(const Gr &g, const GhostParams &p, int i, int x, int y) {
    int xs = x - cx_of(i), ys = y - cy_of(i);
    xs = xs < 0 ? xs + p.nx : (xs >= p.nx ? xs - p.nx : xs);
    ys = ys < 0 ? ys + p.ny : (ys >= p.ny ? ys - p.ny : ys);
    return load_p(g, p, i, xs, ys);
}

// Streamed value with the after-stream inlet / outlet columns.
template <typename Gr, typename T = typename Gr::value_type>
__device__ T load_g(const Gr &g, const GhostParams &p, int i, int x, int y) {
    if (p.pbc_mode == LBM_PBC_AFTER_STREAM && (x == 0 || x == p.nx - 1)) {
        const int xs = x == 0 ? p.nx - 2 : 1;
        T v[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) v[j] = pull(g, p, j, xs, y);
        return pbc_value<T, true>(v, x == 0 ? p.rho_in : p.rho_out, i);
    }
    return pull(g, p, i, x, y);
}

template <typename Gr, typename T = typename Gr::value_type>
__device__ T rule_value(const Gr &g, const GhostParams &p, const lbm_ghost_rule &r, int x,
                        int y) {
    if (r.kind == LBM_RULE_ZERO) return T(0);
    const int xs = x + r.dx, ys = y + r.dy;
    const T v = load_g(g, p, r.src_ch, xs, ys);
    if (r.kind == LBM_RULE_LID) return v + T(r.coef);
    if (r.kind == LBM_RULE_LID_RHO) {
        // 2.0 * (g2 + g5 + g6) + g0 + g1 + g3 at the source node (slidingLid.py:86-87)
        T rw = (load_g(g, p, 2, xs, ys) + load_g(g, p, 5, xs, ys)) + load_g(g, p, 6, xs, ys);
        rw = T(2) * rw;
        rw = rw + load_g(g, p, 0, xs, ys);
        rw = rw + load_g(g, p, 1, xs, ys);
        rw = rw + load_g(g, p, 3, xs, ys);
        return v + T(r.coef) * rw;
    }
    return v;
}

template <typename T, bool Rows>
__global__ __launch_bounds__(kGhostBlock) void k_ghost_step(const T *__restrict__ src, T *__restrict__ dst,
                                                       const GhostParams p, T *__restrict__ rho_out,
                                                       T *__restrict__ ux_out, T *__restrict__ uy_out) {
    const int x = blockIdx.y;
    const int y = blockIdx.x * kGhostBlock + threadIdx.x;
    if (y >= p.ny) return;
    const Grid<T, Rows> g{src, p.nx, p.ny};
    T h[9];
    if (p.flags & LBM_GHOST_COLLIDE_ONLY) {
#pragma unroll
        for (int i = 0; i < 9; ++i) h[i] = g.at(i, x, y);
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) h[i] = load_g(g, p, i, x, y);
        for (int r = 0; r < p.nrules; ++r) {
            const lbm_ghost_rule &rule = p.rules[r];
            if (unsigned(x - rule.x0) < unsigned(rule.x1 - rule.x0) &&
                unsigned(y - rule.y0) < unsigned(rule.y1 - rule.y0)) {
                put(h, rule.ch, rule_value(g, p, rule, x, y));
            }
        }
    }
    if (p.flags & LBM_GHOST_NO_COLLIDE) {
#pragma unroll
        for (int i = 0; i < 9; ++i) dst[index_of<Rows>(i, x, y, p.nx, p.ny)] = h[i];
        return;
    }
    T rho, ux, uy, e[9];
    sf_moments<T>(h, rho, ux, uy);
    eq_of_node<T>(rho, ux, uy, e);
    const T om = T(p.omega);
#pragma unroll
    for (int i = 0; i < 9; ++i) dst[index_of<Rows>(i, x, y, p.nx, p.ny)] = h[i] - om * (h[i] - e[i]);
    if (rho_out) {
        const int64_t node = int64_t(x) * p.ny + y;
        rho_out[node] = rho;
        ux_out[node] = ux;
        uy_out[node] = uy;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ghost_equilibrium(const T *__restrict__ rho,
                                                              const T *__restrict__ ux,
                                                              const T *__restrict__ uy,
                                                              T *__restrict__ f, int64_t n) {
    const int64_t k = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (k >= n) return;
    T e[9];
    eq_of_node<T>(rho[k], ux[k], uy[k], e);
#pragma unroll
    for (int i = 0; i < 9; ++i) f[int64_t(i) * n + k] = e[i];
}

template <typename T>
int ghost_equilibrium(const T *rho, const T *ux, const T *uy, T *f, int64_t n, void *stream) {
    if (!rho || !ux || !uy || !f) return LBM_E_NULL;
    if (n <= 0 || n > (int64_t(1) << 40)) return LBM_E_SIZE;
    const int64_t fb = 9 * n * int64_t(sizeof(T)), mb = n * int64_t(sizeof(T));
    if (overlaps(rho, mb, f, fb) || overlaps(ux, mb, f, fb) || overlaps(uy, mb, f, fb))
        return LBM_E_ALIAS;
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > INT32_MAX) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_ghost_equilibrium<T>, dim3(unsigned(blocks)), dim3(kBlock), 0,
                       as_stream(stream), rho, ux, uy, f, n);
    return launch_status();
}

bool rect_hits(int ax0, int ax1, int ay0, int ay1, int bx0, int bx1, int by0, int by1) {
    return ax0 < bx1 && bx0 < ax1 && ay0 < by1 && by0 < ay1;
}

// Validates the rule table: ranges inside the grid, sources inside the grid,
// and no rule reading a population an earlier rule writes.
int check_rules(const lbm_ghost_rule *rules, int nrules, int nx, int ny) {
    for (int r = 0; r < nrules; ++r) {
        const lbm_ghost_rule &a = rules[r];
        if (a.kind < LBM_RULE_COPY || a.kind > LBM_RULE_ZERO) return LBM_E_ARG;
        if (a.ch < 0 || a.ch > 8) return LBM_E_ARG;
        if (a.x0 < 0 || a.x1 > nx || a.x0 >= a.x1 || a.y0 < 0 || a.y1 > ny || a.y0 >= a.y1)
            return LBM_E_ARG;
        if (a.kind == LBM_RULE_ZERO) continue;
        if (a.src_ch < 0 || a.src_ch > 8) return LBM_E_ARG;
        if (a.x0 + a.dx < 0 || a.x1 + a.dx > nx || a.y0 + a.dy < 0 || a.y1 + a.dy > ny)
            return LBM_E_ARG;
        if ((a.kind == LBM_RULE_LID || a.kind == LBM_RULE_LID_RHO) && !(a.coef == a.coef))
            return LBM_E_ARG;
        const int sx0 = a.x0 + a.dx, sx1 = a.x1 + a.dx, sy0 = a.y0 + a.dy, sy1 = a.y1 + a.dy;
        for (int q = 0; q < r; ++q) {
            const lbm_ghost_rule &b = rules[q];
            if (!rect_hits(sx0, sx1, sy0, sy1, b.x0, b.x1, b.y0, b.y1)) continue;
            if (b.ch == a.src_ch) return LBM_E_ARG;
            if (a.kind == LBM_RULE_LID_RHO &&
                (b.ch == 0 || b.ch == 1 || b.ch == 2 || b.ch == 3 || b.ch == 5 || b.ch == 6))
                return LBM_E_ARG;
        }
    }
    return LBM_OK;
}

template <typename T>
int ghost_step(const T *src, T *dst, int64_t nx, int64_t ny, int layout, double omega,
               const lbm_ghost_rule *rules, int nrules, const lbm_ghost_pbc *pbc, int flags, T *rho,
               T *ux, T *uy, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if (nx < 3 || ny < 3 || nx > 65535 || ny > (int64_t(1) << 30) || nx * ny > INT32_MAX)
        return LBM_E_SIZE;
    if (!valid_layout(layout)) return LBM_E_ARG;
    if (nrules < 0 || nrules > LBM_GHOST_MAX_RULES || (nrules > 0 && !rules)) return LBM_E_ARG;
    if ((flags & ~(LBM_GHOST_NO_COLLIDE | LBM_GHOST_COLLIDE_ONLY)) ||
        flags == (LBM_GHOST_NO_COLLIDE | LBM_GHOST_COLLIDE_ONLY))
        return LBM_E_ARG;
    if (!(omega == omega)) return LBM_E_ARG;
    const bool any = rho || ux || uy;
    if (any && !(rho && ux && uy)) return LBM_E_NULL;
    const int64_t bytes = 9 * nx * ny * int64_t(sizeof(T));
    if (overlaps(src, bytes, dst, bytes)) return LBM_E_ALIAS;
    if (any) {
        const int64_t mb = nx * ny * int64_t(sizeof(T));
        if (overlaps(rho, mb, dst, bytes) || overlaps(ux, mb, dst, bytes) ||
            overlaps(uy, mb, dst, bytes) || overlaps(rho, mb, src, bytes) ||
            overlaps(ux, mb, src, bytes) || overlaps(uy, mb, src, bytes))
            return LBM_E_ALIAS;
    }
    GhostParams p{};
    p.nx = int(nx);
    p.ny = int(ny);
    p.flags = flags;
    p.omega = omega;
    p.pbc_mode = LBM_PBC_NONE;
    if (pbc) {
        if (pbc->mode < LBM_PBC_NONE || pbc->mode > LBM_PBC_AFTER_STREAM) return LBM_E_ARG;
        if (!(pbc->rho_in == pbc->rho_in) || !(pbc->rho_out == pbc->rho_out)) return LBM_E_ARG;
        p.pbc_mode = pbc->mode;
        p.rho_in = pbc->rho_in;
        p.rho_out = pbc->rho_out;
    }
    if (flags & LBM_GHOST_COLLIDE_ONLY) {
        p.nrules = 0;
        p.pbc_mode = LBM_PBC_NONE;
    } else {
        const int rc = check_rules(rules, nrules, p.nx, p.ny);
        if (rc) return rc;
        p.nrules = nrules;
        for (int r = 0; r < nrules; ++r) p.rules[r] = rules[r];
    }
    const dim3 grid(unsigned((ny + kGhostBlock - 1) / kGhostBlock), unsigned(nx));
    if (layout == LBM_LAYOUT_ROWS)
        hipLaunchKernelGGL((k_ghost_step<T, true>), grid, dim3(kGhostBlock), 0, as_stream(stream),
                           src, dst, p, rho, ux, uy);
    else
        hipLaunchKernelGGL((k_ghost_step<T, false>), grid, dim3(kGhostBlock), 0, as_stream(stream),
                           src, dst, p, rho, ux, uy);
    return launch_status();
}

}  // namespace
}  // namespace lbm

extern "C" {

int lbm_ghost_equilibrium_f64(const double *rho, const double *ux, const double *uy, double *f,
                              int64_t n, void *stream) {
    return lbm::ghost_equilibrium<double>(rho, ux, uy, f, n, stream);
}

int lbm_ghost_equilibrium_f32(const float *rho, const float *ux, const float *uy, float *f,
                              int64_t n, void *stream) {
    return lbm::ghost_equilibrium<float>(rho, ux, uy, f, n, stream);
}

int lbm_ghost_step_f64(const double *src, double *dst, int64_t nx, int64_t ny, int layout,
                       double omega, const lbm_ghost_rule *rules, int nrules,
                       const lbm_ghost_pbc *pbc, int flags, double *rho, double *ux, double *uy,
                       void *stream) {
    return lbm::ghost_step<double>(src, dst, nx, ny, layout, omega, rules, nrules, pbc, flags,
                                   rho, ux, uy, stream);
}

int lbm_ghost_step_f32(const float *src, float *dst, int64_t nx, int64_t ny, int layout,
                       double omega, const lbm_ghost_rule *rules, int nrules,
                       const lbm_ghost_pbc *pbc, int flags, float *rho, float *ux, float *uy,
                       void *stream) {
    return lbm::ghost_step<float>(src, dst, nx, ny, layout, omega, rules, nrules, pbc, flags, rho,
                                  ux, uy, stream);
}

}  // extern "C"
