/* Precision-generic body of the C oracle, included once per scalar type by wavefd_oracle.c.
 * TEST INFRASTRUCTURE ONLY (see wavefd_oracle.c).
 *
 * Expects: T (scalar type), FMA (fused multiply-add of T, correctly rounded), SFX (name
 * suffix).  Every expression below is written in the operation order of the HIP kernels so
 * that, built with -ffp-contract=off, the two produce identical bits.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* logical (n0, n1, n2) array <-> zero-padded working array */
static void FN(to_padded)(const geo* g, const T* src, T* dst)
{
    for (int64_t i0 = 0; i0 < g->n[0]; i0++)
        for (int64_t i1 = 0; i1 < g->n[1]; i1++)
            memcpy(dst + (i0 + g->pad[0]) * g->ps[0] + (i1 + g->pad[1]) * g->ps[1] + g->pad[2],
                   src + (i0 * g->n[1] + i1) * g->n[2], sizeof(T) * (size_t)g->n[2]);
}

static void FN(from_padded)(const geo* g, const T* src, T* dst)
{
    for (int64_t i0 = 0; i0 < g->n[0]; i0++)
        for (int64_t i1 = 0; i1 < g->n[1]; i1++)
            memcpy(dst + (i0 * g->n[1] + i1) * g->n[2],
                   src + (i0 + g->pad[0]) * g->ps[0] + (i1 + g->pad[1]) * g->ps[1] + g->pad[2],
                   sizeof(T) * (size_t)g->n[2]);
}

/* Stencil weights in T: w[d*K + m-1] = a_m / h_d^2, wc = a_0 * sum_d 1/h_d^2. */
static void FN(weights)(const geo* g, const double* h, int order, T* wc, T* w)
{
    int K = order / 2;
    double a[5];
    fd_coeffs(order, a);
    double s = 0.0;
    for (int d = 0; d < g->ndim; d++) s += 1.0 / (h[d] * h[d]);
    *wc = (T)(a[0] * s);
    for (int d = 0; d < g->ndim; d++)
        for (int m = 1; m <= K; m++) w[d * K + m - 1] = (T)(a[m] / (h[d] * h[d]));
}

/* Laplacian at padded offset x, kernel order: lap = wc*u; per axis, per m: fma(w, u- + u+, lap). */
static inline T FN(lap_at)(const geo* g, const T* u, int64_t x, T wc, const T* w)
{
    int K = g->K;
    T lap = wc * u[x];
    for (int d = 0; d < g->ndim; d++) {
        int64_t s = g->ps[3 - g->ndim + d];
        for (int m = 1; m <= K; m++) lap = FMA(w[d * K + m - 1], u[x - m * s] + u[x + m * s], lap);
    }
    return lap;
}

static inline T FN(beta_at)(const geo* g, const T* beta, int64_t i0, int64_t i1, int64_t i2)
{
    /* profiles are concatenated over the real axes, slowest first; summed left to right */
    int64_t idx[3] = {i0, i1, i2};
    int64_t off = 0;
    T b = (T)0;
    int first = 1;
    for (int d = 3 - g->ndim; d < 3; d++) {
        T v = beta[off + idx[d]];
        b = first ? v : b + v;
        first = 0;
        off += g->n[d];
    }
    return b;
}

/* One step on padded arrays: un = step(up, uc).  Sources and receivers are padded offsets. */
static void FN(step)(const geo* g, const T* wts, T wc, const T* up, const T* uc, T* un,
                     const T* kc, T kconst, const T* beta,
                     int nsrc, const int64_t* src, const T* sig_n,
                     int nrec, const int64_t* rec, T* trace_n, int nthreads)
{
    const int64_t n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
    const int64_t lo0 = g->ndim >= 3 ? 1 : 0, hi0 = g->ndim >= 3 ? n0 - 1 : 1;
    const int64_t lo1 = g->ndim >= 2 ? 1 : 0, hi1 = g->ndim >= 2 ? n1 - 1 : 1;
#pragma omp parallel for collapse(2) num_threads(nthreads) schedule(static)
    for (int64_t i0 = lo0; i0 < hi0; i0++)
        for (int64_t i1 = lo1; i1 < hi1; i1++) {
            int64_t row = (i0 + g->pad[0]) * g->ps[0] + (i1 + g->pad[1]) * g->ps[1] + g->pad[2];
            for (int64_t i2 = 1; i2 < n2 - 1; i2++) {
                int64_t x = row + i2;
                T lap = FN(lap_at)(g, uc, x, wc, wts);
                T k = kc ? kc[x] : kconst;
                T t;
                if (beta) {
                    T b = FN(beta_at)(g, beta, i0, i1, i2);
                    t = (T)2 * uc[x] - ((T)1 - b) * up[x];
                } else {
                    t = (T)2 * uc[x] - up[x];
                }
                un[x] = FMA(k, lap, t);
            }
        }
    for (int s = 0; sig_n && s < nsrc; s++) {
        int64_t x = src[s];
        if (x < 0) continue; /* on the boundary ring or outside: never injected */
        T k = kc ? kc[x] : kconst;
        un[x] = FMA(k, sig_n[s], un[x]);
    }
    if (beta) {
#pragma omp parallel for collapse(2) num_threads(nthreads) schedule(static)
        for (int64_t i0 = lo0; i0 < hi0; i0++)
            for (int64_t i1 = lo1; i1 < hi1; i1++) {
                int64_t row = (i0 + g->pad[0]) * g->ps[0] + (i1 + g->pad[1]) * g->ps[1] + g->pad[2];
                for (int64_t i2 = 1; i2 < n2 - 1; i2++) {
                    T b = FN(beta_at)(g, beta, i0, i1, i2);
                    un[row + i2] = un[row + i2] / ((T)1 + b);
                }
            }
    }
    for (int r = 0; trace_n && r < nrec; r++) trace_n[r] = rec[r] < 0 ? (T)0 : un[rec[r]];
}

/* Forward run on logical arrays (n0 x n1 x n2, no padding).  State in/out: u_prev, u_cur.
 * traces[n*nrec + r] = u^{n+1}(rec_r); src_sig[n*nsrc + s].  Returns 0 or a negative code. */
int FN(wfo_forward)(int ndim, const int64_t* n, const double* h, int order,
                    const T* kc, double kconst, const T* beta,
                    int nsrc, const int64_t* src_node, const T* src_sig,
                    int nrec, const int64_t* rec_node, T* traces,
                    T* u_prev, T* u_cur, int64_t nsteps, int nthreads)
{
    geo g;
    if (geo_init(&g, ndim, n, order)) return -1;
    T wc, wts[3 * 4];
    FN(weights)(&g, h, order, &wc, wts);
    int64_t tot = g.pn[0] * g.pn[1] * g.pn[2];
    T* a = (T*)calloc((size_t)tot, sizeof(T));
    T* b = (T*)calloc((size_t)tot, sizeof(T));
    T* kp = kc ? (T*)calloc((size_t)tot, sizeof(T)) : NULL;
    int64_t* src = (int64_t*)malloc(sizeof(int64_t) * (nsrc + 1));
    int64_t* rec = (int64_t*)malloc(sizeof(int64_t) * (nrec + 1));
    if (!a || !b || (kc && !kp) || !src || !rec) return -2;
    FN(to_padded)(&g, u_prev, a);
    FN(to_padded)(&g, u_cur, b);
    if (kc) FN(to_padded)(&g, kc, kp);
    for (int s = 0; s < nsrc; s++) src[s] = node_to_padded(&g, src_node[s]);
    for (int r = 0; r < nrec; r++) rec[r] = node_to_padded(&g, rec_node[r]);
    for (int64_t step = 0; step < nsteps; step++) {
        /* in place: the new level overwrites the oldest one */
        FN(step)(&g, wts, wc, a, b, a, kp, (T)kconst, beta, nsrc, src,
                 src_sig ? src_sig + step * nsrc : NULL, nrec, rec,
                 traces ? traces + step * nrec : NULL, nthreads);
        T* t = a; a = b; b = t;
    }
    FN(from_padded)(&g, a, u_prev);
    FN(from_padded)(&g, b, u_cur);
    free(a); free(b); free(kp); free(src); free(rec);
    return 0;
}

/* Adjoint sweep and imaging on logical arrays.
 * adj_sig[n*nrec + r] is the value injected at receiver r at adjoint step m = n+1 (that is
 * -residual[n]); u_hist[n*cells] = u^n for n = 0..nsteps-1.  Produces
 *   G = sum_n mu^{n+1} * (L u^n)            (float64, flushed from a T partial every `seg` steps
 *                                            when (m-1) % seg == 0)
 *   mu_src[n*nsrc + s] = mu^{n+1}(src_s).
 */
/* Adjoint sweeps of nshots shots over one medium, with the imaging sum of all of them.
 *
 * Per shot s: adj_sig[s][n][r] is injected at receiver r at adjoint step m = n + 1, mu_src[s][m-1]
 * records the field at the forward sources, u_hist[s][n] = u^n.  Node lists are [nshots][count]
 * when *_per_shot is set, else one list for every shot.  The imaging partials (type T) are one per
 * group of `group` consecutive shots: at every step m and interior node x, for the shots of the
 * group in ascending order, p = fma(mu_s^m[x], (L u_s^{m-1})[x], p).  Whenever (m-1) % seg == 0
 * the partials are added, in group order, into the float64 gradient (acc = G[x]; acc += p_0[x];
 * acc += p_1[x] ...) and cleared.  group = 1 is one partial per shot.  mu_hist (optional,
 * [nshots][nsteps][cells]) receives mu_s^m at index m - 1.
 */
int FN(wfo_adjoint_shots)(int ndim, const int64_t* n, const double* h, int order,
                          const T* kc, double kconst, const T* beta, int nshots, int group,
                          int nrec, const int64_t* rec_node, int rec_per_shot, const T* adj_sig,
                          int nsrc, const int64_t* src_node, int src_per_shot, T* mu_src,
                          const T* u_hist, int64_t nsteps, int64_t seg, double* G, T* mu_hist,
                          int nthreads)
{
    geo g;
    if (geo_init(&g, ndim, n, order) || seg <= 0 || nshots < 1 || group < 1) return -1;
    T wc, wts[3 * 4];
    FN(weights)(&g, h, order, &wc, wts);
    const int64_t tot = g.pn[0] * g.pn[1] * g.pn[2];
    const int64_t cells = g.n[0] * g.n[1] * g.n[2];
    const int parts = (nshots + group - 1) / group;
    const int nsl = src_per_shot ? nshots : 1, nrl = rec_per_shot ? nshots : 1;
    T* a = (T*)calloc((size_t)(tot * nshots), sizeof(T)); /* per shot: mu^{m+2}, then mu^m */
    T* b = (T*)calloc((size_t)(tot * nshots), sizeof(T)); /* per shot: mu^{m+1} */
    T* up = (T*)calloc((size_t)tot, sizeof(T));
    T* gp = (T*)calloc((size_t)(tot * parts), sizeof(T));
    double* Gp = (double*)calloc((size_t)tot, sizeof(double));
    T* kp = kc ? (T*)calloc((size_t)tot, sizeof(T)) : NULL;
    int64_t* src = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nsrc * nsl + 1));
    int64_t* rec = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nrec * nrl + 1));
    int rc = -2;
    if (!a || !b || !up || !gp || !Gp || (kc && !kp) || !src || !rec) goto done;
    if (kc) FN(to_padded)(&g, kc, kp);
    for (int64_t s = 0; s < (int64_t)nsrc * nsl; s++) src[s] = node_to_padded(&g, src_node[s]);
    for (int64_t r = 0; r < (int64_t)nrec * nrl; r++) rec[r] = node_to_padded(&g, rec_node[r]);
    const int64_t n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const int64_t lo0 = g.ndim >= 3 ? 1 : 0, hi0 = g.ndim >= 3 ? n0 - 1 : 1;
    const int64_t lo1 = g.ndim >= 2 ? 1 : 0, hi1 = g.ndim >= 2 ? n1 - 1 : 1;
    for (int64_t m = nsteps; m >= 1; m--) {
        for (int s = 0; s < nshots; s++) {
            T* as = a + s * tot;
            T* bs = b + s * tot;
            T* gs = gp + (int64_t)(s / group) * tot;
            /* receivers of the forward run are the sources here; the forward sources record */
            FN(step)(&g, wts, wc, as, bs, as, kp, (T)kconst, beta, nrec, rec + (rec_per_shot ? s * nrec : 0),
                     adj_sig + ((int64_t)s * nsteps + m - 1) * nrec, nsrc, src + (src_per_shot ? s * nsrc : 0),
                     mu_src ? mu_src + ((int64_t)s * nsteps + m - 1) * nsrc : NULL, nthreads);
            if (mu_hist) FN(from_padded)(&g, as, mu_hist + ((int64_t)s * nsteps + m - 1) * cells);
            FN(to_padded)(&g, u_hist + ((int64_t)s * nsteps + m - 1) * cells, up);
#pragma omp parallel for collapse(2) num_threads(nthreads) schedule(static)
            for (int64_t i0 = lo0; i0 < hi0; i0++)
                for (int64_t i1 = lo1; i1 < hi1; i1++) {
                    int64_t row = (i0 + g.pad[0]) * g.ps[0] + (i1 + g.pad[1]) * g.ps[1] + g.pad[2];
                    for (int64_t i2 = 1; i2 < n2 - 1; i2++) {
                        int64_t x = row + i2;
                        T lapu = FN(lap_at)(&g, up, x, wc, wts);
                        gs[x] = FMA(as[x], lapu, gs[x]);
                    }
                }
        }
        if ((m - 1) % seg == 0) {
            for (int64_t x = 0; x < tot; x++) {
                double acc = Gp[x];
                for (int q = 0; q < parts; q++) {
                    acc += (double)gp[q * tot + x];
                    gp[q * tot + x] = (T)0;
                }
                Gp[x] = acc;
            }
        }
        T* t = a; a = b; b = t;
    }
    /* un-pad G */
    for (int64_t i0 = 0; i0 < n0; i0++)
        for (int64_t i1 = 0; i1 < n1; i1++)
            for (int64_t i2 = 0; i2 < n2; i2++)
                G[(i0 * n1 + i1) * n2 + i2] =
                    Gp[(i0 + g.pad[0]) * g.ps[0] + (i1 + g.pad[1]) * g.ps[1] + i2 + g.pad[2]];
    rc = 0;
done:
    free(a); free(b); free(up); free(gp); free(Gp); free(kp); free(src); free(rec);
    return rc;
}

/* One shot: the single partial flushed into G whenever (m-1) % seg == 0. */
int FN(wfo_adjoint)(int ndim, const int64_t* n, const double* h, int order,
                    const T* kc, double kconst, const T* beta,
                    int nrec, const int64_t* rec_node, const T* adj_sig,
                    int nsrc, const int64_t* src_node, T* mu_src,
                    const T* u_hist, int64_t nsteps, int64_t seg, double* G, int nthreads)
{
    return FN(wfo_adjoint_shots)(ndim, n, h, order, kc, kconst, beta, 1, 1, nrec, rec_node, 0, adj_sig,
                                 nsrc, src_node, 0, mu_src, u_hist, nsteps, seg, G, NULL, nthreads);
}

/* The Laplacian L u the imaging sum uses, at interior nodes (zero on the boundary ring). */
int FN(wfo_lap)(int ndim, const int64_t* n, const double* h, int order, const T* u, T* out)
{
    geo g;
    if (geo_init(&g, ndim, n, order)) return -1;
    T wc, wts[3 * 4];
    FN(weights)(&g, h, order, &wc, wts);
    const int64_t tot = g.pn[0] * g.pn[1] * g.pn[2];
    T* up = (T*)calloc((size_t)tot, sizeof(T));
    T* lp = (T*)calloc((size_t)tot, sizeof(T));
    if (!up || !lp) { free(up); free(lp); return -2; }
    FN(to_padded)(&g, u, up);
    const int64_t n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const int64_t lo0 = g.ndim >= 3 ? 1 : 0, hi0 = g.ndim >= 3 ? n0 - 1 : 1;
    const int64_t lo1 = g.ndim >= 2 ? 1 : 0, hi1 = g.ndim >= 2 ? n1 - 1 : 1;
    for (int64_t i0 = lo0; i0 < hi0; i0++)
        for (int64_t i1 = lo1; i1 < hi1; i1++) {
            int64_t row = (i0 + g.pad[0]) * g.ps[0] + (i1 + g.pad[1]) * g.ps[1] + g.pad[2];
            for (int64_t i2 = 1; i2 < n2 - 1; i2++) lp[row + i2] = FN(lap_at)(&g, up, row + i2, wc, wts);
        }
    FN(from_padded)(&g, lp, out);
    free(up); free(lp);
    return 0;
}

#undef FN
#undef CAT
#undef CAT2
