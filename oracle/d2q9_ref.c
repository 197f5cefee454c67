/*
 * Single-thread C restatement of the reference's D2Q9 BGK path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and by the CPU-baseline leg of
 * bench.py through oracle/ctypes_ref.py; never linked into the product.
 *
 * Layout: f[i*nx*ny + k*ny + l], channel i, x index k (slow), y index l (fast),
 * the layout of the reference's numpy arrays f_ikl (cavity_opt1.py:81-85).
 *
 * Two formulas are restated:
 *   opt1 order (cavity_opt1.py:100-245, Appendix-B order): the full step
 *     stream -> walls/lid -> collide as one pull loop, bit-identical to the
 *     numpy restatement when built with -ffp-contract=off;
 *   _lbkernels order (c/d2q9.h:59-131): equilibrium and collide as the
 *     reference's C++ extension writes them (prescaled velocities, rho/9
 *     weights), for the "opt2" CPU baseline (numpy stream + C collide, as
 *     cavity_opt2.py:272-277).
 * Both are built for float and double from one macro body.
 */
#include <stdint.h>
#include <stddef.h>

static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

static inline int64_t wrap(int64_t a, int64_t n) { a %= n; return a < 0 ? a + n : a; }

#define DEFINE_ORACLE(T, SUF)                                                        \
/* opt1 equilibrium + relaxation on one node, g[] in, out[] written.           */   \
static inline void opt1_collide_node_##SUF(const T g[9], T omega, T out[9],          \
                                           T *rho_o, T *ux_o, T *uy_o) {             \
    const T w0 = (T)(4.0 / 9.0), w1 = (T)(1.0 / 9.0), w5 = (T)(1.0 / 36.0);         \
    T rho = g[0];                                                                    \
    for (int i = 1; i < 9; ++i) rho = rho + g[i];                                    \
    T ux = (((((g[1] - g[3]) + g[5]) - g[6]) - g[7]) + g[8]) / rho;                 \
    T uy = (((((g[2] - g[4]) + g[5]) + g[6]) - g[7]) - g[8]) / rho;                 \
    T cu5 = ux + uy, cu6 = (-ux) + uy, cu7 = (-ux) - uy, cu8 = ux - uy;              \
    T uu = ux * ux + uy * uy;                                                        \
    T uu15 = (T)1.5 * uu;                                                            \
    T feq[9];                                                                        \
    feq[0] = (w0 * rho) * ((T)1 - uu15);                                             \
    feq[1] = (w1 * rho) * ((((T)1 + (T)3 * ux) + (T)4.5 * (ux * ux)) - uu15);        \
    feq[2] = (w1 * rho) * ((((T)1 + (T)3 * uy) + (T)4.5 * (uy * uy)) - uu15);        \
    feq[3] = (w1 * rho) * ((((T)1 - (T)3 * ux) + (T)4.5 * (ux * ux)) - uu15);        \
    feq[4] = (w1 * rho) * ((((T)1 - (T)3 * uy) + (T)4.5 * (uy * uy)) - uu15);        \
    feq[5] = (w5 * rho) * ((((T)1 + (T)3 * cu5) + (T)4.5 * (cu5 * cu5)) - uu15);     \
    feq[6] = (w5 * rho) * ((((T)1 + (T)3 * cu6) + (T)4.5 * (cu6 * cu6)) - uu15);     \
    feq[7] = (w5 * rho) * ((((T)1 + (T)3 * cu7) + (T)4.5 * (cu7 * cu7)) - uu15);     \
    feq[8] = (w5 * rho) * ((((T)1 + (T)3 * cu8) + (T)4.5 * (cu8 * cu8)) - uu15);     \
    for (int i = 0; i < 9; ++i) out[i] = g[i] + omega * (feq[i] - g[i]);            \
    if (rho_o) { *rho_o = rho; *ux_o = ux; *uy_o = uy; }                             \
}                                                                                    \
                                                                                     \
/* One full step src -> dst. bc = 0 periodic, 1 cavity (walls + lid u0).        */  \
/* rho/ux/uy (nullable) receive the pre-collision moments.                      */  \
void oracle_step_##SUF(const T *src, T *dst, int64_t nx, int64_t ny, T omega,       \
                       int bc, T u0, T *rho, T *ux, T *uy) {                         \
    const int64_t P = nx * ny;                                                       \
    const T lidc = (T)6 * (T)(1.0 / 36.0);                                           \
    for (int64_t k = 0; k < nx; ++k) {                                               \
        for (int64_t l = 0; l < ny; ++l) {                                           \
            T g[9], o[9];                                                            \
            for (int i = 0; i < 9; ++i)                                              \
                g[i] = src[i * P + wrap(k - CX[i], nx) * ny + wrap(l - CY[i], ny)];  \
            if (bc == 1) {                                                           \
                const T *c = src + k * ny + l; /* pre-stream node values c[i*P] */   \
                if (l == 0) { g[2] = c[4 * P]; g[5] = c[7 * P]; g[6] = c[8 * P]; }   \
                if (l == ny - 1) {                                                   \
                    T rl = c[6 * P] + c[2 * P];                                      \
                    rl = rl + c[5 * P]; rl = rl + g[6]; rl = rl + g[2];              \
                    rl = rl + g[5]; rl = rl + g[3]; rl = rl + g[0]; rl = rl + g[1];  \
                    g[4] = c[2 * P];                                                 \
                    g[8] = c[6 * P] + (lidc * rl) * u0;                              \
                    g[7] = c[5 * P] - (lidc * rl) * u0;                              \
                }                                                                    \
                if (k == 0) { g[1] = c[3 * P]; g[5] = c[7 * P]; g[8] = c[6 * P]; }   \
                if (k == nx - 1) { g[3] = c[1 * P]; g[6] = c[8 * P]; g[7] = c[5 * P]; } \
            }                                                                        \
            const int64_t n = k * ny + l;                                            \
            if (rho) opt1_collide_node_##SUF(g, omega, o, rho + n, ux + n, uy + n);  \
            else opt1_collide_node_##SUF(g, omega, o, NULL, NULL, NULL);             \
            for (int i = 0; i < 9; ++i) dst[i * P + n] = o[i];                       \
        }                                                                            \
    }                                                                                \
}                                                                                    \
                                                                                     \
/* nsteps steps, ping-ponging between a and b; the result ends in a.            */  \
void oracle_run_##SUF(T *a, T *b, int64_t nx, int64_t ny, T omega, int bc, T u0,    \
                      int64_t nsteps) {                                              \
    const int64_t total = 9 * nx * ny;                                               \
    for (int64_t s = 0; s < nsteps; ++s) {                                           \
        oracle_step_##SUF(a, b, nx, ny, omega, bc, u0, NULL, NULL, NULL);            \
        T *t = a; a = b; b = t;                                                      \
    }                                                                                \
    if (nsteps & 1) for (int64_t i = 0; i < total; ++i) b[i] = a[i];                 \
}                                                                                    \
                                                                                     \
/* opt1 collide in place on f[9][n] (cavity_opt1.py:134-160).                  */   \
void oracle_collide_##SUF(T *f, int64_t n, T omega) {                                \
    for (int64_t j = 0; j < n; ++j) {                                                \
        T g[9], o[9];                                                                \
        for (int i = 0; i < 9; ++i) g[i] = f[i * n + j];                             \
        opt1_collide_node_##SUF(g, omega, o, NULL, NULL, NULL);                      \
        for (int i = 0; i < 9; ++i) f[i * n + j] = o[i];                             \
    }                                                                                \
}                                                                                    \
                                                                                     \
/* _lbkernels d2q9_equilibrium1 (c/d2q9.h:59-81), one node.                     */  \
static inline void lbk_eq_node_##SUF(T rho, T ux, T uy, T out[9]) {                  \
    T a0 = (4 * rho) / 9, a1 = rho / 9, a5 = rho / 36;                               \
    ux = ux * 3; uy = uy * 3;                                                        \
    T cu5 = ux + uy, cu6 = -ux + uy, cu7 = -ux - uy, cu8 = ux - uy;                  \
    T uu = (ux * ux + uy * uy) / 6;                                                  \
    out[0] = a0 * (1 - uu);                                                          \
    out[1] = a1 * (((1 + ux) + (ux * ux) / 2) - uu);                                 \
    out[2] = a1 * (((1 + uy) + (uy * uy) / 2) - uu);                                 \
    out[3] = a1 * (((1 - ux) + (ux * ux) / 2) - uu);                                 \
    out[4] = a1 * (((1 - uy) + (uy * uy) / 2) - uu);                                 \
    out[5] = a5 * (((1 + cu5) + (cu5 * cu5) / 2) - uu);                              \
    out[6] = a5 * (((1 + cu6) + (cu6 * cu6) / 2) - uu);                              \
    out[7] = a5 * (((1 + cu7) + (cu7 * cu7) / 2) - uu);                              \
    out[8] = a5 * (((1 + cu8) + (cu8 * cu8) / 2) - uu);                              \
}                                                                                    \
                                                                                     \
/* _lbkernels d2q9_equilibriumn (c/d2q9.h:98-108): f[9][n] from rho, ux, uy.    */  \
void oracle_lbk_equilibrium_##SUF(const T *rho, const T *ux, const T *uy, T *f,     \
                                  int64_t n) {                                       \
    for (int64_t j = 0; j < n; ++j) {                                                \
        T o[9];                                                                      \
        lbk_eq_node_##SUF(rho[j], ux[j], uy[j], o);                                  \
        for (int i = 0; i < 9; ++i) f[i * n + j] = o[i];                             \
    }                                                                                \
}                                                                                    \
                                                                                     \
/* _lbkernels d2q9_colliden (c/d2q9.h:121-131), in place on f[9][n].            */  \
/* The density is summed pairwise (an unrolled reduction of 9 terms).           */  \
void oracle_lbk_collide_##SUF(T *f, int64_t n, T omega) {                            \
    for (int64_t j = 0; j < n; ++j) {                                                \
        T g[9], e[9];                                                                \
        for (int i = 0; i < 9; ++i) g[i] = f[i * n + j];                             \
        T rho = ((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + (g[7] + g[8]))); \
        T ux = (g[1] - g[3] + g[5] - g[6] - g[7] + g[8]) / rho;                      \
        T uy = (g[2] - g[4] + g[5] + g[6] - g[7] - g[8]) / rho;                      \
        lbk_eq_node_##SUF(rho, ux, uy, e);                                           \
        for (int i = 0; i < 9; ++i) f[i * n + j] = g[i] + omega * (e[i] - g[i]);     \
    }                                                                                \
}

DEFINE_ORACLE(double, f64)
DEFINE_ORACLE(float, f32)
