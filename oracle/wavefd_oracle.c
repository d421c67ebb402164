/* C oracle for the finite-difference wave stepping.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded (through ctypes) by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker; the product library never links or calls it.
 *
 * It restates the damped leapfrog update of the reference
 *   u2 = -u0 + 2 u1 + dt^2 c^2 (Lx u1 + Ly u1)      (projects/18_simple_wave_fd/step2D_wave.cu:72-79,
 *                                                    2D_wave_cupy.py:54-62)
 * extended to order 2K, a per-node coefficient kc = v^2 dt^2, the sponge
 *   (1 + b) u^{n+1} = 2 u^n - (1 - b) u^{n-1} + kc (L u^n + s^n)
 * and point sources / receivers, in scalar loops over zero-padded arrays.  The arithmetic is
 * written in the operation order of the HIP kernels (one fma per stencil pair, explicit fma for
 * the coefficient product), so built with -ffp-contract=off it is bit-identical to them on
 * IEEE hardware.  The slice-style NumPy restatement in slice_ref.py is the independent writing
 * it is checked against.
 *
 * Layout of every array argument: logical, C order, (n0, n1, n2) nodes including the boundary
 * ring (ndim < 3 drops the leading axes).  Nodes are flat logical indices.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PAD 4 /* half-width of the widest stencil (order 8) */

typedef struct {
    int ndim, K;
    int64_t n[3];   /* nodes per axis, leading axes of size 1 when ndim < 3 */
    int64_t pad[3]; /* PAD on real axes, 0 on the leading size-1 axes */
    int64_t pn[3];  /* padded extents */
    int64_t ps[3];  /* padded strides */
} geo;

static int geo_init(geo* g, int ndim, const int64_t* n, int order)
{
    if (ndim < 1 || ndim > 3 || order < 2 || order > 8 || order % 2) return 1;
    g->ndim = ndim;
    g->K = order / 2;
    for (int d = 0; d < 3; d++) {
        int real = d >= 3 - ndim;
        g->n[d] = real ? n[d - (3 - ndim)] : 1;
        if (real && g->n[d] < 3) return 1;
        g->pad[d] = real ? PAD : 0;
        g->pn[d] = g->n[d] + 2 * g->pad[d];
    }
    g->ps[2] = 1;
    g->ps[1] = g->pn[2];
    g->ps[0] = g->pn[1] * g->pn[2];
    return 0;
}

/* logical flat node -> padded offset, -1 if the node is not interior */
static int64_t node_to_padded(const geo* g, int64_t node)
{
    int64_t i2 = node % g->n[2];
    int64_t r = node / g->n[2];
    int64_t i1 = r % g->n[1];
    int64_t i0 = r / g->n[1];
    int64_t idx[3] = {i0, i1, i2};
    if (node < 0 || i0 >= g->n[0]) return -1;
    for (int d = 3 - g->ndim; d < 3; d++)
        if (idx[d] < 1 || idx[d] > g->n[d] - 2) return -1;
    return (i0 + g->pad[0]) * g->ps[0] + (i1 + g->pad[1]) * g->ps[1] + i2 + g->pad[2];
}

static int64_t gcd64(int64_t a, int64_t b)
{
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a;
}

/* Central second-derivative weights a_0..a_K, each rounded once from the exact rational:
 * a_k = 2 (-1)^{k+1} (K!)^2 / (k^2 (K-k)! (K+k)!),  a_0 = -2 sum a_k. */
static void fd_coeffs(int order, double* a)
{
    int K = order / 2;
    int64_t f[17];
    f[0] = 1;
    for (int i = 1; i <= 16; i++) f[i] = f[i - 1] * i;
    int64_t sn = 0, sd = 1; /* running exact sum of a_k */
    for (int k = 1; k <= K; k++) {
        int64_t num = 2 * f[K] * f[K];
        int64_t den = (int64_t)k * k * f[K - k] * f[K + k];
        int64_t q = gcd64(num, den);
        num /= q;
        den /= q;
        if (k % 2 == 0) num = -num;
        a[k] = (double)num / (double)den;
        int64_t nn = sn * den + num * sd, nd = sd * den;
        q = gcd64(nn, nd);
        sn = nn / q;
        sd = nd / q;
    }
    a[0] = (double)(-2 * sn) / (double)sd;
}

/* exported for the tests: the weights as doubles */
int wfo_fd_coeffs(int order, double* a)
{
    if (order < 2 || order > 8 || order % 2) return -1;
    fd_coeffs(order, a);
    return 0;
}

#define T float
#define FMA fmaf
#define SFX _f32
#include "wavefd_oracle_body.h"
#undef T
#undef FMA
#undef SFX

#define T double
#define FMA fma
#define SFX _f64
#include "wavefd_oracle_body.h"
#undef T
#undef FMA
#undef SFX

/* exported for the tests: weights in the working precision */
int wfo_weights_f32(int ndim, const double* h, int order, float* wc, float* w)
{
    geo g;
    int64_t n[3] = {3, 3, 3};
    if (geo_init(&g, ndim, n, order)) return -1;
    weights_f32(&g, h, order, wc, w);
    return 0;
}

int wfo_weights_f64(int ndim, const double* h, int order, double* wc, double* w)
{
    geo g;
    int64_t n[3] = {3, 3, 3};
    if (geo_init(&g, ndim, n, order)) return -1;
    weights_f64(&g, h, order, wc, w);
    return 0;
}
