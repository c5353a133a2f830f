/*
 * qvti_oracle_impl.h -- body of the CPU oracle, included twice by
 * qvti_oracle.c with REAL = float (SFX _f32) and REAL = double (SFX _f64).
 * TEST INFRASTRUCTURE ONLY (see qvti_oracle.h).
 *
 * Every block cites the reference lines it restates.
 */

#define CAT2_(a, b) a##b
#define CAT_(a, b) CAT2_(a, b)
#define FN(x) CAT_(x, SFX)

/* ---------------------------------------------------------------- FFT ---- */
/* Mixed-radix Stockham autosort, radices 4,2,3,5 and any other prime by a
 * direct DFT.  Twiddles are computed in FP64 and rounded to REAL. */

#ifndef QO_MAXR
#define QO_MAXR 1024
#endif
typedef struct {
    int n, nf, fac[40];
    REAL *twr, *twi; /* exp(-2 pi i t / n), t in [0, n) */
    int maxr;
} FN(plan_t);

static int FN(plan_init)(FN(plan_t) *p, int n)
{
    int m = n, nf = 0;
    if (n < 1) return -1;
    p->n = n;
    /* 4s first, then 2, 3, 5, then other primes */
    while (m % 4 == 0) { p->fac[nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[nf++] = 2; m /= 2; }
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { p->fac[nf++] = q; m /= q; }
    for (int i = 0; i < nf; i++)
        if (p->fac[i] > QO_MAXR) return -3; /* prime factor too large for the direct DFT */
    p->nf = nf;
    p->maxr = 1;
    for (int i = 0; i < nf; i++) if (p->fac[i] > p->maxr) p->maxr = p->fac[i];
    p->twr = (REAL *)malloc(sizeof(REAL) * n);
    p->twi = (REAL *)malloc(sizeof(REAL) * n);
    if (!p->twr || !p->twi) return -2;
    for (int t = 0; t < n; t++) {
        double a = -2.0 * M_PI * (double)t / (double)n;
        p->twr[t] = (REAL)cos(a);
        p->twi[t] = (REAL)sin(a);
    }
    return 0;
}

static void FN(plan_free)(FN(plan_t) *p)
{
    free(p->twr);
    free(p->twi);
}

/* In-register DFT of size R: X[q] = sum_r v[r] exp(sign 2 pi i q r / R). */
static void FN(dft_small)(int R, int sign, REAL *vr, REAL *vi, const FN(plan_t) *p)
{
    if (R == 2) {
        REAL ar = vr[0], ai = vi[0];
        vr[0] = ar + vr[1]; vi[0] = ai + vi[1];
        vr[1] = ar - vr[1]; vi[1] = ai - vi[1];
    } else if (R == 4) {
        REAL s = (REAL)sign;
        REAL t0r = vr[0] + vr[2], t0i = vi[0] + vi[2];
        REAL t1r = vr[0] - vr[2], t1i = vi[0] - vi[2];
        REAL t2r = vr[1] + vr[3], t2i = vi[1] + vi[3];
        REAL t3r = vr[1] - vr[3], t3i = vi[1] - vi[3];
        /* t3 * (sign i) */
        REAL ur = -s * t3i, ui = s * t3r;
        vr[0] = t0r + t2r; vi[0] = t0i + t2i;
        vr[2] = t0r - t2r; vi[2] = t0i - t2i;
        vr[1] = t1r + ur;  vi[1] = t1i + ui;
        vr[3] = t1r - ur;  vi[3] = t1i - ui;
    } else {
        REAL xr[QO_MAXR], xi[QO_MAXR];
        for (int q = 0; q < R; q++) {
            double accr = 0.0, acci = 0.0;
            for (int r = 0; r < R; r++) {
                /* exp(sign 2 pi i q r / R) via the length-n table */
                int t = (int)(((long long)q * r % R) * (p->n / R));
                /* twi holds sin(-2 pi t/n): forward uses it, inverse its negation */
                REAL wr = p->twr[t], wi = (sign < 0) ? p->twi[t] : -p->twi[t];
                accr += (double)(vr[r] * wr - vi[r] * wi);
                acci += (double)(vr[r] * wi + vi[r] * wr);
            }
            xr[q] = (REAL)accr;
            xi[q] = (REAL)acci;
        }
        for (int q = 0; q < R; q++) { vr[q] = xr[q]; vi[q] = xi[q]; }
    }
}

/* One pencil: x (re, im arrays) transformed in place using y as scratch. */
static void FN(fft_pencil)(const FN(plan_t) *p, int sign, REAL *xr, REAL *xi, REAL *yr, REAL *yi)
{
    int n = p->n, Ns = 1;
    REAL *sr = xr, *si = xi, *dr = yr, *di = yi;
    for (int f = 0; f < p->nf; f++) {
        int R = p->fac[f], m = n / R, ts = n / (Ns * R);
        for (int j = 0; j < m; j++) {
            int k = j % Ns;
            REAL vr[QO_MAXR], vi[QO_MAXR];
            for (int r = 0; r < R; r++) { vr[r] = sr[j + r * m]; vi[r] = si[j + r * m]; }
            for (int r = 1; r < R; r++) {
                int t = k * r * ts;
                REAL wr = p->twr[t], wi = (sign < 0) ? p->twi[t] : -p->twi[t];
                REAL ar = vr[r] * wr - vi[r] * wi;
                REAL ai = vr[r] * wi + vi[r] * wr;
                vr[r] = ar; vi[r] = ai;
            }
            FN(dft_small)(R, sign, vr, vi, p);
            int d = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) { dr[d + r * Ns] = vr[r]; di[d + r * Ns] = vi[r]; }
        }
        REAL *t;
        t = sr; sr = dr; dr = t;
        t = si; si = di; di = t;
        Ns *= R;
    }
    if (sr != xr) {
        memcpy(xr, sr, sizeof(REAL) * n);
        memcpy(xi, si, sizeof(REAL) * n);
    }
}

/* Blocked variant: QO_B pencils side by side, element (i, b) at i*QO_B + b,
 * so the inner loop over b runs with one twiddle set (vectorisable).  Same
 * Stockham factorisation and twiddles as fft_pencil. */
#ifndef QO_B
#define QO_B 16
#endif
static void FN(fft_block)(const FN(plan_t) *p, int sign, REAL *xr, REAL *xi, REAL *yr, REAL *yi)
{
    const int n = p->n, B = QO_B;
    int Ns = 1;
    REAL *sr = xr, *si = xi, *dr = yr, *di = yi;
    for (int f = 0; f < p->nf; f++) {
        const int R = p->fac[f], m = n / R, ts = n / (Ns * R);
        for (int j = 0; j < m; j++) {
            const int k = j % Ns;
            const int d = (j - k) * R + k;
            REAL wr[QO_MAXR], wi[QO_MAXR];
            wr[0] = 1; wi[0] = 0;
            for (int r = 1; r < R; r++) {
                const int t = k * r * ts;
                wr[r] = p->twr[t];
                wi[r] = (sign < 0) ? p->twi[t] : -p->twi[t];
            }
            if (R == 2) {
                for (int b = 0; b < B; b++) {
                    REAL ar = sr[j * B + b], ai = si[j * B + b];
                    REAL br0 = sr[(j + m) * B + b], bi0 = si[(j + m) * B + b];
                    REAL br = br0 * wr[1] - bi0 * wi[1], bi = br0 * wi[1] + bi0 * wr[1];
                    dr[d * B + b] = ar + br; di[d * B + b] = ai + bi;
                    dr[(d + Ns) * B + b] = ar - br; di[(d + Ns) * B + b] = ai - bi;
                }
            } else if (R == 4) {
                const REAL sg = (REAL)sign;
                for (int b = 0; b < B; b++) {
                    REAL v0r = sr[j * B + b], v0i = si[j * B + b];
                    REAL u, q;
                    u = sr[(j + m) * B + b]; q = si[(j + m) * B + b];
                    REAL v1r = u * wr[1] - q * wi[1], v1i = u * wi[1] + q * wr[1];
                    u = sr[(j + 2 * m) * B + b]; q = si[(j + 2 * m) * B + b];
                    REAL v2r = u * wr[2] - q * wi[2], v2i = u * wi[2] + q * wr[2];
                    u = sr[(j + 3 * m) * B + b]; q = si[(j + 3 * m) * B + b];
                    REAL v3r = u * wr[3] - q * wi[3], v3i = u * wi[3] + q * wr[3];
                    REAL t0r = v0r + v2r, t0i = v0i + v2i, t1r = v0r - v2r, t1i = v0i - v2i;
                    REAL t2r = v1r + v3r, t2i = v1i + v3i, t3r = v1r - v3r, t3i = v1i - v3i;
                    REAL ur = -sg * t3i, ui = sg * t3r;
                    dr[d * B + b] = t0r + t2r; di[d * B + b] = t0i + t2i;
                    dr[(d + 2 * Ns) * B + b] = t0r - t2r; di[(d + 2 * Ns) * B + b] = t0i - t2i;
                    dr[(d + Ns) * B + b] = t1r + ur; di[(d + Ns) * B + b] = t1i + ui;
                    dr[(d + 3 * Ns) * B + b] = t1r - ur; di[(d + 3 * Ns) * B + b] = t1i - ui;
                }
            } else if (R == 3) {
                const REAL h = (REAL)(sign * 0.86602540378443864676);
                for (int b = 0; b < B; b++) {
                    REAL v0r = sr[j * B + b], v0i = si[j * B + b], u, q;
                    u = sr[(j + m) * B + b]; q = si[(j + m) * B + b];
                    REAL v1r = u * wr[1] - q * wi[1], v1i = u * wi[1] + q * wr[1];
                    u = sr[(j + 2 * m) * B + b]; q = si[(j + 2 * m) * B + b];
                    REAL v2r = u * wr[2] - q * wi[2], v2i = u * wi[2] + q * wr[2];
                    REAL t1r = v1r + v2r, t1i = v1i + v2i;
                    REAL t2r = v0r - (REAL)0.5 * t1r, t2i = v0i - (REAL)0.5 * t1i;
                    REAL t3r = -h * (v1i - v2i), t3i = h * (v1r - v2r);
                    dr[d * B + b] = v0r + t1r; di[d * B + b] = v0i + t1i;
                    dr[(d + Ns) * B + b] = t2r + t3r; di[(d + Ns) * B + b] = t2i + t3i;
                    dr[(d + 2 * Ns) * B + b] = t2r - t3r; di[(d + 2 * Ns) * B + b] = t2i - t3i;
                }
            } else if (R == 5) {
                const REAL c1 = (REAL)0.30901699437494742410, c2 = (REAL)-0.80901699437494742410;
                const REAL s1 = (REAL)(sign * 0.95105651629515357212), s2 = (REAL)(sign * 0.58778525229247312917);
                for (int b = 0; b < B; b++) {
                    REAL xr5[5], xi5[5];
                    for (int r = 0; r < 5; r++) {
                        REAL u = sr[(j + r * m) * B + b], q = si[(j + r * m) * B + b];
                        xr5[r] = u * wr[r] - q * wi[r];
                        xi5[r] = u * wi[r] + q * wr[r];
                    }
                    REAL t1r = xr5[1] + xr5[4], t1i = xi5[1] + xi5[4], t2r = xr5[2] + xr5[3], t2i = xi5[2] + xi5[3];
                    REAL t3r = xr5[1] - xr5[4], t3i = xi5[1] - xi5[4], t4r = xr5[2] - xr5[3], t4i = xi5[2] - xi5[3];
                    REAL a1r = xr5[0] + c1 * t1r + c2 * t2r, a1i = xi5[0] + c1 * t1i + c2 * t2i;
                    REAL a2r = xr5[0] + c2 * t1r + c1 * t2r, a2i = xi5[0] + c2 * t1i + c1 * t2i;
                    REAL e1r = s1 * t3r + s2 * t4r, e1i = s1 * t3i + s2 * t4i;
                    REAL e2r = s2 * t3r - s1 * t4r, e2i = s2 * t3i - s1 * t4i;
                    /* multiply by i: (re, im) -> (-im, re) */
                    dr[d * B + b] = xr5[0] + t1r + t2r; di[d * B + b] = xi5[0] + t1i + t2i;
                    dr[(d + Ns) * B + b] = a1r - e1i; di[(d + Ns) * B + b] = a1i + e1r;
                    dr[(d + 4 * Ns) * B + b] = a1r + e1i; di[(d + 4 * Ns) * B + b] = a1i - e1r;
                    dr[(d + 2 * Ns) * B + b] = a2r - e2i; di[(d + 2 * Ns) * B + b] = a2i + e2r;
                    dr[(d + 3 * Ns) * B + b] = a2r + e2i; di[(d + 3 * Ns) * B + b] = a2i - e2r;
                }
            } else {
                for (int b = 0; b < B; b++) {
                    REAL vr[QO_MAXR], vi[QO_MAXR];
                    for (int r = 0; r < R; r++) {
                        REAL u = sr[(j + r * m) * B + b], q = si[(j + r * m) * B + b];
                        vr[r] = u * wr[r] - q * wi[r];
                        vi[r] = u * wi[r] + q * wr[r];
                    }
                    FN(dft_small)(R, sign, vr, vi, p);
                    for (int r = 0; r < R; r++) { dr[(d + r * Ns) * B + b] = vr[r]; di[(d + r * Ns) * B + b] = vi[r]; }
                }
            }
        }
        REAL *t;
        t = sr; sr = dr; dr = t;
        t = si; si = di; di = t;
        Ns *= R;
    }
    if (sr != xr) {
        memcpy(xr, sr, sizeof(REAL) * n * B);
        memcpy(xi, si, sizeof(REAL) * n * B);
    }
}

/* 3-D transform of interleaved complex data [n0][n1][n2], axis by axis,
 * QO_B pencils at a time.  Reference: cufftExecC2C on a
 * cufftPlan3d(ntx,nty,ntz) plan (cuviscoelas_3D.cu:1703-1704), unnormalised. */
typedef struct {
    int n[3];
    FN(plan_t) p[3];
} FN(plan3_t);

static int FN(plan3_init)(FN(plan3_t) *q, int n0, int n1, int n2)
{
    q->n[0] = n0; q->n[1] = n1; q->n[2] = n2;
    for (int a = 0; a < 3; a++)
        if (FN(plan_init)(&q->p[a], q->n[a])) return -1;
    return 0;
}

static void FN(plan3_free)(FN(plan3_t) *q)
{
    for (int a = 0; a < 3; a++) FN(plan_free)(&q->p[a]);
}

static void FN(fft3d_exec)(const FN(plan3_t) *q, int sign, REAL *data)
{
    const int n0 = q->n[0], n1 = q->n[1], n2 = q->n[2], B = QO_B;
    for (int axis = 2; axis >= 0; axis--) {
        const int n = q->n[axis];
        /* pencils are grouped B at a time: along z, B consecutive rows; along
         * y and x, B consecutive z-columns of one plane */
        long nblk;
        if (axis == 2) nblk = ((long)n0 * n1 + B - 1) / B;
        else if (axis == 1) nblk = (long)n0 * ((n2 + B - 1) / B);
        else nblk = (long)n1 * ((n2 + B - 1) / B);
#pragma omp parallel
        {
            REAL *buf = (REAL *)malloc(sizeof(REAL) * 4 * (size_t)n * B);
            REAL *xr = buf, *xi = buf + (size_t)n * B, *yr = buf + 2 * (size_t)n * B, *yi = buf + 3 * (size_t)n * B;
#pragma omp for schedule(static)
            for (long blk = 0; blk < nblk; blk++) {
                long base, stride, pstep;
                int cnt;
                if (axis == 2) {
                    const long r0 = blk * B, nr = (long)n0 * n1;
                    cnt = (int)(nr - r0 < B ? nr - r0 : B);
                    base = r0 * n2; stride = 1; pstep = n2;
                } else {
                    const int zb = (n2 + B - 1) / B;
                    const long other = blk / zb;
                    const int z0 = (int)(blk % zb) * B;
                    cnt = n2 - z0 < B ? n2 - z0 : B;
                    base = (axis == 1 ? other * (long)n1 * n2 : other * (long)n2) + z0;
                    stride = axis == 1 ? n2 : (long)n1 * n2;
                    pstep = 1;
                }
                for (int i = 0; i < n; i++)
                    for (int b = 0; b < B; b++) {
                        if (b < cnt) {
                            const long g = 2 * (base + i * stride + b * pstep);
                            xr[i * B + b] = data[g];
                            xi[i * B + b] = data[g + 1];
                        } else {
                            xr[i * B + b] = 0;
                            xi[i * B + b] = 0;
                        }
                    }
                FN(fft_block)(&q->p[axis], sign, xr, xi, yr, yi);
                for (int i = 0; i < n; i++)
                    for (int b = 0; b < cnt; b++) {
                        const long g = 2 * (base + i * stride + b * pstep);
                        data[g] = xr[i * B + b];
                        data[g + 1] = xi[i * B + b];
                    }
            }
            free(buf);
        }
    }
}

int FN(qo_fft1d)(int n, int sign, int howmany, REAL *data)
{
    FN(plan_t) p;
    if (n < 1 || howmany < 0 || (sign != 1 && sign != -1)) return -1;
    if (FN(plan_init)(&p, n)) return -2;
    REAL *buf = (REAL *)malloc(sizeof(REAL) * 4 * n);
    for (int h = 0; h < howmany; h++) {
        REAL *x = data + 2L * n * h;
        for (int i = 0; i < n; i++) { buf[i] = x[2 * i]; buf[n + i] = x[2 * i + 1]; }
        FN(fft_pencil)(&p, sign, buf, buf + n, buf + 2 * n, buf + 3 * n);
        for (int i = 0; i < n; i++) { x[2 * i] = buf[i]; x[2 * i + 1] = buf[n + i]; }
    }
    free(buf);
    FN(plan_free)(&p);
    return 0;
}

int FN(qo_fft3d)(int n0, int n1, int n2, int sign, REAL *data)
{
    FN(plan3_t) q;
    if (n0 < 1 || n1 < 1 || n2 < 1 || (sign != 1 && sign != -1)) return -1;
    if (FN(plan3_init)(&q, n0, n1, n2)) return -2;
    FN(fft3d_exec)(&q, sign, data);
    FN(plan3_free)(&q);
    return 0;
}

/* ------------------------------------------------------- setup tables ---- */

/* Folded |K| and the D+/D- symbols of one axis, cuviscoelas_3D.cu:44-149.
 * The reference's float intermediates (dk, tmp, K = tmp*m) are kept as REAL;
 * sin/cos of m*pi/n are FP64 with the truncated pi.  D+ = iK exp(+iK d/2)
 * (kvz_z, kvx_x, kvy_y), D- = iK exp(-iK d/2) (the other six). */
int FN(qo_axis_tables)(int n, float d, REAL *kfold, REAL *dplus, REAL *dminus)
{
    if (n < 1) return -1;
    REAL dk = (REAL)(1.0 / n / d);       /* dkz=1.0/ntz/dz;  :45 */
    REAL tmp = (REAL)(2 * QO_PI * dk);   /* tmpz=2*pi*dkz;   :52 */
    for (int i = 0; i < n; i++) {
        int m = (i <= n / 2) ? i : n - i;
        if (kfold) kfold[i] = (REAL)(2 * QO_PI / n / d * m); /* :55,57 */
        REAL K = tmp * (REAL)m;
        double s = sin(m * QO_PI / n), c = cos(m * QO_PI / n);
        if (i <= n / 2) {
            if (dplus)  { dplus[2 * i] = (REAL)(-(double)K * s);  dplus[2 * i + 1] = (REAL)((double)K * c); }
            if (dminus) { dminus[2 * i] = (REAL)((double)K * s);  dminus[2 * i + 1] = (REAL)((double)K * c); }
        } else {
            if (dplus)  { dplus[2 * i] = (REAL)(-(double)K * s);  dplus[2 * i + 1] = (REAL)(-(double)K * c); }
            if (dminus) { dminus[2 * i] = (REAL)((double)K * s);  dminus[2 * i + 1] = (REAL)(-(double)K * c); }
        }
    }
    return 0;
}

/* C-PML a(x), b(x) along one axis, cuviscoelas_3D.cu:213-309.  vmax is
 * overridden to 5000 (:235); kappa (gamma*) is 1 everywhere (:245,254). */
int FN(qo_pml_profile)(int n, int pml, float d, float f0, REAL *a, REAL *b)
{
    if (pml < 2 || n < 2 * pml) return -1;
    const int n1 = 2, n2 = 1, n3 = 2;
    const float R = 1e-2f;
    const float vmax = 5000.0f;
    REAL alpha_max = (REAL)(QO_PI * f0);
    REAL Om_max = (REAL)((1 + n1 + n2) * vmax * log(1.0 / R) / ((n1 + n2 - 1) * pml * d));
    for (int i = 0; i < n; i++) {
        REAL alpha, Om, kappa = (REAL)1.0;
        int j = (i < pml) ? i : ((i > n - 1 - pml) ? n - 1 - i : -1);
        if (j >= 0) {
#if QO_IS_F32
            alpha = alpha_max * powf((float)(1.0 * j / (pml - 1)), (float)n3);
            Om = Om_max * powf((float)(1.0 * (pml - 1 - j) / pml), (float)(n1 + n2));
#else
            alpha = alpha_max * pow(1.0 * j / (pml - 1), (double)n3);
            Om = Om_max * pow(1.0 * (pml - 1 - j) / pml, (double)(n1 + n2));
#endif
        } else {
            alpha = alpha_max;
            Om = (REAL)0.0;
        }
        a[i] = alpha + Om / kappa;
        b[i] = Om / (kappa * kappa);
    }
    return 0;
}

/* ------------------------------------------------------------ the run ---- */

typedef struct { REAL e1, e2, e3, t1, t2; } FN(qset_t);

/* S_M[g] at node arithmetic (float expression tree of calculate_p,
 * cuviscoelas_3D.cu:847-898): c*part/ntp summed left to right. */
static inline REAL FN(sterm)(REAL c1, REAL c2, REAL c3, REAL c4, REAL c5,
                             REAL p1, REAL p2, REAL p3, REAL p4, REAL p5, REAL ntp)
{
    return c1 * p1 / ntp + c2 * p2 / ntp + c3 * p3 / ntp + c4 * p4 / ntp + c5 * p5 / ntp;
}

/* S-bar term with the 8-corner cube averages (cuviscoelas_3D.cu:902-1005);
 * 0.125 is a double literal in the reference, so the sum is FP64. */
static inline double FN(sbar)(const REAL *const *cs, long ip, long sy, long sx,
                              REAL p1, REAL p2, REAL p3, REAL p4, REAL p5, REAL ntp)
{
    const REAL p[5] = {p1, p2, p3, p4, p5};
    double acc = 0.0;
    for (int j = 0; j < 5; j++) {
        const REAL *c = cs[j];
        REAL s8 = c[ip] + c[ip + 1] + c[ip + sy] + c[ip + 1 + sy] + c[ip + sx] + c[ip + sx + 1]
                + c[ip + sx + sy] + c[ip + sx + 1 + sy];
        acc += 0.125 * s8 * p[j] / ntp;
    }
    return acc;
}

/* Opt-in k-space factor of the stress operators, F(|k|) = sinc^2(vh |k|) W(|k|)
 * with vh = vref dt / 2 and sinc(x) = sin(x)/x: the reference's disabled
 * sinc2nd = pow(sin(velp_max |k| dt/2) / (velp_max |k| dt/2), 2)
 * (cuviscoelas_3D.cu:183-184), multiplied into A1, A2, A3 as at :188-190;
 * W is a radial Tukey taper, 1 for |k| <= k0 = (1 - r) kc, half cosine
 * (1 + cos(pi (|k| - k0) / (kc - k0))) / 2 down to 0 at kc = 2 pi fcut / vmax,
 * 0 beyond (the filter declared but never built: Myfunctions.h:69,
 * viscoelas_3D.cpp:111-112).  vref = 0 / fcut = 0 drop either part. */
typedef struct { REAL vh, k0, kc, pw; int on; } FN(ks_t);

static int FN(ks_init)(FN(ks_t) *k, float dt, float vref, float fcut, float vmax, float taper)
{
    memset(k, 0, sizeof(*k));
    if (!(vref >= 0) || !(fcut >= 0)) return -1;
    if (fcut > 0 && (!(vmax > 0) || !(taper >= 0 && taper <= 1))) return -1;
    k->vh = (REAL)(0.5 * (double)vref * (double)dt);
    if (fcut > 0) {
        double kc = 2.0 * QO_PI * (double)fcut / (double)vmax, k0 = (1.0 - (double)taper) * kc;
        k->kc = (REAL)kc;
        k->k0 = (REAL)k0;
        k->pw = kc > k0 ? (REAL)(M_PI / (kc - k0)) : (REAL)0;
    }
    k->on = vref > 0 || fcut > 0;
    return 0;
}

static REAL FN(ks_factor)(const FN(ks_t) *k, REAL ka)
{
    REAL f = 1;
    if (k->vh > 0) {
        REAL x = k->vh * ka;
#if QO_IS_F32
        REAL sx = x > 0 ? sinf(x) / x : (REAL)1;
#else
        REAL sx = x > 0 ? sin(x) / x : (REAL)1;
#endif
        f = sx * sx;
    }
    if (k->kc > 0) {
#if QO_IS_F32
        if (ka > k->kc) f = 0;
        else if (ka > k->k0) f *= (REAL)0.5 * ((REAL)1 + cosf(k->pw * (ka - k->k0)));
#else
        if (ka > k->kc) f = 0;
        else if (ka > k->k0) f *= 0.5 * (1.0 + cos(k->pw * (ka - k->k0)));
#endif
    }
    return f;
}

int FN(qo_kspace_factor)(int n, const REAL *kabs, float dt, float vref, float fcut, float vmax, float taper,
                         REAL *out)
{
    FN(ks_t) k;
    if (n < 0 || FN(ks_init)(&k, dt, vref, fcut, vmax, taper)) return -1;
    for (int i = 0; i < n; i++) out[i] = FN(ks_factor)(&k, kabs[i]);
    return 0;
}

int FN(qo_run)(const qo_config *c, const float *rik,
               const float *vp, const float *gp, const float *vs, const float *gs,
               const float *rho_in, const float *eps, const float *del,
               REAL *rec_vx, REAL *rec_vy, REAL *rec_vz, REAL *snaps, REAL *state)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, pml = c->pml, nt = c->nt;
    if (nx < 4 || ny < 4 || nz < 4 || pml < 2 || nx < 2 * pml || ny < 2 * pml || nz < 2 * pml || nt < 0)
        return -1;
    if (!rik || !vp || !gp || !vs || !gs || !rho_in) return -1;
    if (c->vti && (!eps || !del)) return -1;
    FN(ks_t) ks;
    if (FN(ks_init)(&ks, c->dt, c->ks_vref, c->ks_fcut, c->ks_vmax, c->ks_taper)) return -1;
    if (c->sx < 0 || c->sx >= nx || c->sy < 0 || c->sy >= ny || c->sz < 0 || c->sz >= nz) return -1;
    for (int r = 0; r < c->nrec; r++)
        if (c->rx[r] < 0 || c->rx[r] >= nx || c->ry[r] < 0 || c->ry[r] >= ny || c->rz[r] < 0 || c->rz[r] >= nz)
            return -1;
    for (int r = 0; r < c->ninj; r++)
        if (c->inj_x[r] < 0 || c->inj_x[r] >= nx || c->inj_y[r] < 0 || c->inj_y[r] >= ny || c->inj_z[r] < 0 ||
            c->inj_z[r] >= nz)
            return -1;
    if (c->hist_field > QO_PRESSURE) return -1;
#ifdef _OPENMP
    if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#endif
    FN(plan3_t) plan;
    memset(&plan, 0, sizeof(plan));
    if (FN(plan3_init)(&plan, nx, ny, nz)) { FN(plan3_free)(&plan); return -2; }
    const long N = (long)nx * ny * nz;
    const long sx = (long)ny * nz, sy = nz;
    const REAL ntp = (REAL)N;
    const REAL dt = (REAL)c->dt;

    /* -------- tables: wavenumber (cuviscoelas_3D.cu:22-152) */
    REAL *kxf = malloc(sizeof(REAL) * nx), *kyf = malloc(sizeof(REAL) * ny), *kzf = malloc(sizeof(REAL) * nz);
    REAL *dpx = malloc(sizeof(REAL) * 2 * nx), *dmx = malloc(sizeof(REAL) * 2 * nx);
    REAL *dpy = malloc(sizeof(REAL) * 2 * ny), *dmy = malloc(sizeof(REAL) * 2 * ny);
    REAL *dpz = malloc(sizeof(REAL) * 2 * nz), *dmz = malloc(sizeof(REAL) * 2 * nz);
    FN(qo_axis_tables)(nx, c->dx, kxf, dpx, dmx);
    FN(qo_axis_tables)(ny, c->dy, kyf, dpy, dmy);
    FN(qo_axis_tables)(nz, c->dz, kzf, dpz, dmz);
    /* A1 = k^-1/2, A2 = 1, A3 = k^1/2 with k = |k|^2, k[0] = 1e-10 (:72-74, 188-190) */
    REAL *A1 = malloc(sizeof(REAL) * N), *A3 = malloc(sizeof(REAL) * N);
    /* k-space mode only: A2 = F, and F multiplies A1 and A3 */
    REAL *A2 = ks.on ? malloc(sizeof(REAL) * N) : NULL;
#pragma omp parallel for schedule(static)
    for (long ip = 0; ip < N; ip++) {
        int ix = (int)(ip / sx), iy = (int)((ip / sy) % ny), iz = (int)(ip % nz);
        REAL k2 = kxf[ix] * kxf[ix] + kzf[iz] * kzf[iz] + kyf[iy] * kyf[iy];
        if (ip == 0) k2 = (REAL)1e-10;
#if QO_IS_F32
        A1[ip] = powf(k2, -0.5f);
        A3[ip] = powf(k2, 0.5f);
#else
        A1[ip] = pow(k2, -0.5);
        A3[ip] = pow(k2, 0.5);
#endif
        if (A2) {
            REAL F = FN(ks_factor)(&ks, A3[ip]);
            A1[ip] *= F;
            A3[ip] *= F;
            A2[ip] = F;
        }
    }

    /* -------- Q coefficients (cuviscoelas_3D.cu:195-205), FP64 then REAL.
     * Sets: P (vp, gp), S (vs, gs); VTI adds H (vp*sqrt(1+2 eps), gp) for C11
     * and D (v_delta, gp) for C13 + 2 C44, where
     * rho v_delta^2 = sqrt((C33-C44)(C33(1+2 delta)-C44)) + C44. */
    REAL *cP[5], *cS[5], *cH[5], *cD[5];
    for (int j = 0; j < 5; j++) {
        cP[j] = malloc(sizeof(REAL) * N);
        cS[j] = malloc(sizeof(REAL) * N);
        cH[j] = c->vti ? malloc(sizeof(REAL) * N) : NULL;
        cD[j] = c->vti ? malloc(sizeof(REAL) * N) : NULL;
    }
    /* planted mistakes (slip tests only): applied to every coefficient set */
#define QSLIP(q)                                                     \
    do {                                                             \
        if (c->media == 1) { (q)[3] = 0.0; (q)[4] = 0.0; }            \
        if (c->media == 3) { (q)[3] = -(q)[3]; (q)[4] = -(q)[4]; }    \
        if (c->slip & QO_SLIP_ETA_SWAP) { double t_ = (q)[0]; (q)[0] = (q)[2]; (q)[2] = t_; } \
        if (c->slip & QO_SLIP_NO_TAU) { (q)[3] = 0.0; (q)[4] = 0.0; } \
        if (c->slip & QO_SLIP_ETA_SIGN) { (q)[0] = -(q)[0]; (q)[2] = -(q)[2]; } \
    } while (0)
#pragma omp parallel for schedule(static)
    for (long ip = 0; ip < N; ip++) {
        double q[5];
        qo_q_coeffs(rho_in[ip], vp[ip], c->media == 2 ? 0.0 : gp[ip], c->w0, q);
        QSLIP(q);
        for (int j = 0; j < 5; j++) cP[j][ip] = (REAL)q[j];
        qo_q_coeffs(rho_in[ip], vs[ip], c->media == 2 ? 0.0 : gs[ip], c->w0, q);
        QSLIP(q);
        for (int j = 0; j < 5; j++) cS[j][ip] = (REAL)q[j];
        if (c->vti) {
            double vh = (eps[ip] == 0.0f) ? (double)vp[ip] : (double)vp[ip] * sqrt(1.0 + 2.0 * (double)eps[ip]);
            qo_q_coeffs(rho_in[ip], vh, c->media == 2 ? 0.0 : gp[ip], c->w0, q);
            QSLIP(q);
            for (int j = 0; j < 5; j++) cH[j][ip] = (REAL)q[j];
            double vd;
            if (del[ip] == 0.0f) {
                vd = vp[ip];
            } else {
                double r = rho_in[ip], c33 = r * (double)vp[ip] * vp[ip], c44 = r * (double)vs[ip] * vs[ip];
                double prod = (c33 - c44) * (c33 * (1.0 + 2.0 * (double)del[ip]) - c44);
                vd = sqrt((sqrt(prod > 0 ? prod : 0.0) + c44) / r);
            }
            qo_q_coeffs(rho_in[ip], vd, c->media == 2 ? 0.0 : gp[ip], c->w0, q);
            QSLIP(q);
            for (int j = 0; j < 5; j++) cD[j][ip] = (REAL)q[j];
        }
    }
#undef QSLIP

    /* -------- C-PML profiles */
    REAL *ax = malloc(sizeof(REAL) * nx), *bx = malloc(sizeof(REAL) * nx);
    REAL *ay = malloc(sizeof(REAL) * ny), *by = malloc(sizeof(REAL) * ny);
    REAL *az = malloc(sizeof(REAL) * nz), *bz = malloc(sizeof(REAL) * nz);
    FN(qo_pml_profile)(nx, pml, c->dx, c->f0, ax, bx);
    FN(qo_pml_profile)(ny, pml, c->dy, c->f0, ay, by);
    FN(qo_pml_profile)(nz, pml, c->dz, c->f0, az, bz);

    /* -------- state (zeroed as cuda_kernel_initialization, :312-410) */
    REAL *F[9];
    for (int f = 0; f < 9; f++) F[f] = calloc(N, sizeof(REAL));
    REAL *vx = F[QO_VX], *vy = F[QO_VY], *vz = F[QO_VZ];
    REAL *pxx = F[QO_PXX], *pyy = F[QO_PYY], *pzz = F[QO_PZZ];
    REAL *pxy = F[QO_PXY], *pxz = F[QO_PXZ], *pyz = F[QO_PYZ];
    /* 24 memory variables, full grid like the reference */
    enum { PVXXX, PVZZX, PVYYX, PVXXY, PVZZY, PVYYY, PVXXZ, PVZZZ, PVYYZ,
           PVXZ, PVZX, PVXY, PVYX, PVYZ, PVZY,
           PPXXX, PPXYY, PPXZZ, PPXYX, PPYYY, PPYZZ, PPXZX, PPYZY, PPZZZ, NPHI };
    REAL *phi[NPHI];
    for (int f = 0; f < NPHI; f++) phi[f] = calloc(N, sizeof(REAL));
    /* spectra: 6 stress, 3 velocity, 3 previous velocity (zero at step 0) */
    REAL *spec[9], *dv[3];
    for (int f = 0; f < 9; f++) spec[f] = malloc(sizeof(REAL) * 2 * N);
    for (int f = 0; f < 3; f++) dv[f] = calloc(2 * N, sizeof(REAL));
    REAL *work = malloc(sizeof(REAL) * 2 * N);
    /* 9 velocity-step parts and 45 stress-step parts (only .x kept) */
    REAL *part[9], *pv[45];
    for (int f = 0; f < 9; f++) part[f] = malloc(sizeof(REAL) * N);
    for (int f = 0; f < 45; f++) pv[f] = malloc(sizeof(REAL) * N);


    long snap_n = (long)(nx - 2 * pml) * (ny - 2 * pml) * (nz - 2 * pml);
    long snap_k = 0;

    /* operator symbol along one axis, broadcast: axis 0 = x, 1 = y, 2 = z */
#define SYM(tab, axis, ip) ((axis) == 0 ? (tab) + 2 * ((ip) / sx) : ((axis) == 1 ? (tab) + 2 * (((ip) / sy) % ny) : (tab) + 2 * ((ip) % nz)))

    for (int it = 0; it < nt; it++) {
        /* ===== velocity half-step (:1265-1325) */
        REAL *sig[6] = {pxx, pyy, pzz, pxy, pxz, pyz};
        for (int f = 0; f < 6; f++) {
#pragma omp parallel for schedule(static)
            for (long ip = 0; ip < N; ip++) { spec[f][2 * ip] = sig[f][ip]; spec[f][2 * ip + 1] = 0; }
            FN(fft3d_exec)(&plan, -1, spec[f]);
        }
        /* (spectrum, symbol table, axis) for partx1..partz3; :1279-1306 */
        struct { int s; const REAL *tab; int axis; } vop[9] = {
            {0, dmx, 0}, {3, dpy, 1}, {4, dpz, 2},   /* partx1 D-x pxx, party1 D+y pxy, partz1 D+z pxz */
            {4, dpx, 0}, {5, dpy, 1}, {2, dmz, 2},   /* partx2 D+x pxz, party2 D+y pyz, partz2 D-z pzz */
            {3, dpx, 0}, {1, dmy, 1}, {5, dpz, 2},   /* partx3 D+x pxy, party3 D-y pyy, partz3 D+z pyz */
        };
        for (int o = 0; o < 9; o++) {
            const REAL *S = spec[vop[o].s];
#pragma omp parallel for schedule(static)
            for (long ip = 0; ip < N; ip++) {
                const REAL *k = SYM(vop[o].tab, vop[o].axis, ip);
                work[2 * ip] = k[0] * S[2 * ip] - k[1] * S[2 * ip + 1];
                work[2 * ip + 1] = k[0] * S[2 * ip + 1] + k[1] * S[2 * ip];
            }
            FN(fft3d_exec)(&plan, +1, work);
#pragma omp parallel for schedule(static)
            for (long ip = 0; ip < N; ip++) part[o][ip] = work[2 * ip];
        }
        /* calculate_v (:1018-1101); kappa = 1 */
        {
            REAL *px1 = part[0], *py1 = part[1], *pz1 = part[2];
            REAL *px2 = part[3], *py2 = part[4], *pz2 = part[5];
            REAL *px3 = part[6], *py3 = part[7], *pz3 = part[8];
#pragma omp parallel for schedule(static)
            for (long ip = 0; ip < N; ip++) {
                int ix = (int)(ip / sx), iy = (int)((ip / sy) % ny), iz = (int)(ip % nz);
                if (ix >= nx - 1 || iy >= ny - 1 || iz >= nz - 1) continue;
                REAL *ph;
                ph = phi[PPXXX];
                if (c->slip & QO_SLIP_PML_NODE_V) ph[ip] = ph[ip] + dt * (-ax[ix] * ph[ip] - bx[ix] * px1[ip] / ntp);
                else ph[ip] = ph[ip] + dt * (-0.5 * (ax[ix] + ax[ix + 1]) * ph[ip] - 0.5 * (bx[ix] + bx[ix + 1]) * px1[ip] / ntp);
                ph = phi[PPXYY]; ph[ip] = ph[ip] + dt * (-ay[iy] * ph[ip] - by[iy] * py1[ip] / ntp);
                ph = phi[PPXZZ]; ph[ip] = ph[ip] + dt * (-az[iz] * ph[ip] - bz[iz] * pz1[ip] / ntp);
                ph = phi[PPXYX]; ph[ip] = ph[ip] + dt * (-ax[ix] * ph[ip] - bx[ix] * px3[ip] / ntp);
                ph = phi[PPYYY]; ph[ip] = ph[ip] + dt * (-0.5 * (ay[iy] + ay[iy + 1]) * ph[ip] - 0.5 * (by[iy] + by[iy + 1]) * py3[ip] / ntp);
                ph = phi[PPYZZ]; ph[ip] = ph[ip] + dt * (-az[iz] * ph[ip] - bz[iz] * pz3[ip] / ntp);
                ph = phi[PPXZX]; ph[ip] = ph[ip] + dt * (-ax[ix] * ph[ip] - bx[ix] * px2[ip] / ntp);
                ph = phi[PPYZY]; ph[ip] = ph[ip] + dt * (-ay[iy] * ph[ip] - by[iy] * py2[ip] / ntp);
                ph = phi[PPZZZ]; ph[ip] = ph[ip] + dt * (-0.5 * (az[iz] + az[iz + 1]) * ph[ip] - 0.5 * (bz[iz] + bz[iz + 1]) * pz2[ip] / ntp);

                vx[ip] = vx[ip] + dt / (0.5 * ((REAL)rho_in[ip] + (REAL)rho_in[ip + sx])) *
                    (1.0 / (0.5 * (1.0 + 1.0)) * px1[ip] / ntp + phi[PPXXX][ip] + 1.0 * py1[ip] / ntp + phi[PPXYY][ip] + 1.0 * pz1[ip] / ntp + phi[PPXZZ][ip]);
                vy[ip] = vy[ip] + dt / (0.5 * ((REAL)rho_in[ip] + (REAL)rho_in[ip + sy])) *
                    (1.0 * px3[ip] / ntp + phi[PPXYX][ip] + 1.0 / (0.5 * (1.0 + 1.0)) * py3[ip] / ntp + phi[PPYYY][ip] + 1.0 * pz3[ip] / ntp + phi[PPYZZ][ip]);
                vz[ip] = vz[ip] + dt / (0.5 * ((REAL)rho_in[ip] + (REAL)rho_in[ip + 1])) *
                    (1.0 * px2[ip] / ntp + phi[PPXZX][ip] + 1.0 * py2[ip] / ntp + phi[PPYZY][ip] + 1.0 / (0.5 * (1.0 + 1.0)) * pz2[ip] / ntp + phi[PPZZZ][ip]);
            }
        }

        /* ===== stress half-step (:1331-1591) */
        REAL *vel[3] = {vx, vy, vz};
        for (int f = 0; f < 3; f++) {
            REAL *S = spec[6 + f];
#pragma omp parallel for schedule(static)
            for (long ip = 0; ip < N; ip++) { S[2 * ip] = vel[f][ip]; S[2 * ip + 1] = 0; }
            FN(fft3d_exec)(&plan, -1, S);
        }
        /* 45 parts: index = 5*g + (j-1), g in
         * 0 vx_x(D+x vx) 1 vy_y(D+y vy) 2 vz_z(D+z vz) 3 vx_z(D-z vx) 4 vz_x(D-x vz)
         * 5 vx_y(D-y vx) 6 vy_x(D-x vy) 7 vy_z(D-z vy) 8 vz_y(D-y vz);
         * j = 1,2,3: A_j D V;  j = 4,5: A_{j-3} D (V - Vprev)/dt */
        struct { int comp; const REAL *tab; int axis; } sop[9] = {
            {0, dpx, 0}, {1, dpy, 1}, {2, dpz, 2}, {0, dmz, 2}, {2, dmx, 0},
            {0, dmy, 1}, {1, dmx, 0}, {1, dmz, 2}, {2, dmy, 1},
        };
        if (c->slip & QO_SLIP_SHEAR_DPLUS) sop[3].tab = dpz;
        for (int g = 0; g < 9; g++) {
            const REAL *S = spec[6 + sop[g].comp], *P = dv[sop[g].comp];
            for (int j = 1; j <= 5; j++) {
                const REAL *A = (j == 1 || j == 4) ? A1 : ((j == 3) ? A3 : A2);
                int AorB = (j >= 4);
#pragma omp parallel for schedule(static)
                for (long ip = 0; ip < N; ip++) {
                    const REAL *k = SYM(sop[g].tab, sop[g].axis, ip);
                    REAL k2 = A ? A[ip] : (REAL)1.0;
                    REAL ur, ui;
                    if (AorB) { ur = (S[2 * ip] - P[2 * ip]) / dt; ui = (S[2 * ip + 1] - P[2 * ip + 1]) / dt; }
                    else { ur = S[2 * ip]; ui = S[2 * ip + 1]; }
                    work[2 * ip] = k2 * (k[0] * ur - k[1] * ui);
                    work[2 * ip + 1] = k2 * (k[0] * ui + k[1] * ur);
                }
                FN(fft3d_exec)(&plan, +1, work);
                REAL *dst = pv[5 * g + j - 1];
#pragma omp parallel for schedule(static)
                for (long ip = 0; ip < N; ip++) dst[ip] = work[2 * ip];
            }
        }
        /* calculate_p (:800-1015) generalised to VTI moduli:
         * iso: C11 = C33 = P, C12 = C13 = P - 2S; VTI: C11 = H, C33 = P,
         * C12 = H - 2S, C13 = D - 2S.  Shear moduli S-bar in every case. */
        {
            REAL **c11 = c->vti ? cH : cP, **c33 = cP, **c13b = c->vti ? cD : cP;
            const REAL *const sb[5] = {cS[0], cS[1], cS[2], cS[3], cS[4]};
#pragma omp parallel for schedule(static)
            for (long ip = 0; ip < N; ip++) {
                int ix = (int)(ip / sx), iy = (int)((ip / sy) % ny), iz = (int)(ip % nz);
                if (ix >= nx - 1 || iy >= ny - 1 || iz >= nz - 1) continue;
#define PV(g, j) pv[5 * (g) + (j) - 1][ip]
#define NODE(C, g) FN(sterm)(C[0][ip], C[1][ip], C[2][ip], C[3][ip], C[4][ip], PV(g, 1), PV(g, 2), PV(g, 3), PV(g, 4), PV(g, 5), ntp)
#define DIFF(C, g) FN(sterm)(C[0][ip] - 2 * cS[0][ip], C[1][ip] - 2 * cS[1][ip], C[2][ip] - 2 * cS[2][ip], C[3][ip] - 2 * cS[3][ip], C[4][ip] - 2 * cS[4][ip], PV(g, 1), PV(g, 2), PV(g, 3), PV(g, 4), PV(g, 5), ntp)
#define SBAR(g) FN(sbar)(sb, ip, sy, sx, PV(g, 1), PV(g, 2), PV(g, 3), PV(g, 4), PV(g, 5), ntp)
                REAL s11x = NODE(c11, 0), s12x = DIFF(c11, 0), s13x = DIFF(c13b, 0);
                REAL s12y = DIFF(c11, 1), s11y = NODE(c11, 1), s13y = DIFF(c13b, 1);
                REAL s13z = DIFF(c13b, 2), s33z = NODE(c33, 2);
                REAL *ph;
                ph = phi[PVXXX]; ph[ip] = ph[ip] + dt * (-ax[ix] * ph[ip] - bx[ix] * s11x);
                ph = phi[PVZZX]; ph[ip] = ph[ip] + dt * (-az[iz] * ph[ip] - bz[iz] * s13z);
                ph = phi[PVYYX]; ph[ip] = ph[ip] + dt * (-ay[iy] * ph[ip] - by[iy] * s12y);
                ph = phi[PVZZZ]; ph[ip] = ph[ip] + dt * (-az[iz] * ph[ip] - bz[iz] * s33z);
                ph = phi[PVXXZ]; ph[ip] = ph[ip] + dt * (-ax[ix] * ph[ip] - bx[ix] * s13x);
                ph = phi[PVYYZ]; ph[ip] = ph[ip] + dt * (-ay[iy] * ph[ip] - by[iy] * s13y);
                ph = phi[PVYYY]; ph[ip] = ph[ip] + dt * (-ay[iy] * ph[ip] - by[iy] * s11y);
                ph = phi[PVXXY]; ph[ip] = ph[ip] + dt * (-ax[ix] * ph[ip] - bx[ix] * s12x);
                ph = phi[PVZZY]; ph[ip] = ph[ip] + dt * (-az[iz] * ph[ip] - bz[iz] * s13z);

                pxx[ip] = pxx[ip] + dt * (1.0 * s11x + 1.0 * s12y + 1.0 * s13z + (phi[PVXXX][ip] + phi[PVYYX][ip] + phi[PVZZX][ip]));
                pyy[ip] = pyy[ip] + dt * (1.0 * s12x + 1.0 * s11y + 1.0 * s13z + (phi[PVXXY][ip] + phi[PVYYY][ip] + phi[PVZZY][ip]));
                pzz[ip] = pzz[ip] + dt * (1.0 * s13x + 1.0 * s13y + 1.0 * s33z + (phi[PVXXZ][ip] + phi[PVYYZ][ip] + phi[PVZZZ][ip]));

                double sb_vxz = SBAR(3), sb_vzx = SBAR(4), sb_vxy = SBAR(5), sb_vyx = SBAR(6), sb_vyz = SBAR(7), sb_vzy = SBAR(8);
                ph = phi[PVXZ];
                if (c->slip & QO_SLIP_PML_NODE_S) ph[ip] = ph[ip] + dt * (-az[iz] * ph[ip] - bz[iz] * sb_vxz);
                else ph[ip] = ph[ip] + dt * (-0.5 * (az[iz] + az[iz + 1]) * ph[ip] - 0.5 * (bz[iz] + bz[iz + 1]) * sb_vxz);
                ph = phi[PVZX]; ph[ip] = ph[ip] + dt * (-0.5 * (ax[ix] + ax[ix + 1]) * ph[ip] - 0.5 * (bx[ix] + bx[ix + 1]) * sb_vzx);
                ph = phi[PVXY]; ph[ip] = ph[ip] + dt * (-0.5 * (ay[iy] + ay[iy + 1]) * ph[ip] - 0.5 * (by[iy] + by[iy + 1]) * sb_vxy);
                ph = phi[PVYX]; ph[ip] = ph[ip] + dt * (-0.5 * (ax[ix] + ax[ix + 1]) * ph[ip] - 0.5 * (bx[ix] + bx[ix + 1]) * sb_vyx);
                ph = phi[PVYZ]; ph[ip] = ph[ip] + dt * (-0.5 * (az[iz] + az[iz + 1]) * ph[ip] - 0.5 * (bz[iz] + bz[iz + 1]) * sb_vyz);
                ph = phi[PVZY]; ph[ip] = ph[ip] + dt * (-0.5 * (ay[iy] + ay[iy + 1]) * ph[ip] - 0.5 * (by[iy] + by[iy + 1]) * sb_vzy);

                pxz[ip] = pxz[ip] + dt * (1.0 / (0.5 * (1.0 + 1.0)) * sb_vxz + phi[PVXZ][ip] + 1.0 / (0.5 * (1.0 + 1.0)) * sb_vzx + phi[PVZX][ip]);
                pyz[ip] = pyz[ip] + dt * (1.0 / (0.5 * (1.0 + 1.0)) * sb_vyz + phi[PVYZ][ip] + 1.0 / (0.5 * (1.0 + 1.0)) * sb_vzy + phi[PVZY][ip]);
                pxy[ip] = pxy[ip] + dt * (1.0 / (0.5 * (1.0 + 1.0)) * sb_vxy + phi[PVXY][ip] + 1.0 / (0.5 * (1.0 + 1.0)) * sb_vyx + phi[PVYX][ip]);
#undef PV
#undef NODE
#undef DIFF
#undef SBAR
            }
        }

        /* ===== source + receivers (cuviscoelas_3D.cu:771-793): source first */
        {
            long is = (long)c->sx * sx + (long)c->sy * sy + c->sz;
            REAL s = (REAL)rik[it];
            switch (c->src_kind) {
            case QO_SRC_FX: vx[is] += s; break;
            case QO_SRC_FY: vy[is] += s; break;
            case QO_SRC_FZ: vz[is] += s; break;
            default: pxx[is] += s; pyy[is] += s; pzz[is] += s; break;
            }
            const float *inj[3] = {c->inj_vx, c->inj_vy, c->inj_vz};
            for (int r = 0; r < c->ninj; r++) {
                long ij = (long)c->inj_x[r] * sx + (long)c->inj_y[r] * sy + c->inj_z[r];
                for (int f = 0; f < 3; f++)
                    if (inj[f]) F[QO_VX + f][ij] += (REAL)inj[f][(long)it * c->ninj + r];
            }
            for (int r = 0; r < c->nrec; r++) {
                long ir = (long)c->rx[r] * sx + (long)c->ry[r] * sy + c->rz[r];
                if (rec_vx) rec_vx[(long)it * c->nrec + r] = vx[ir];
                if (rec_vy) rec_vy[(long)it * c->nrec + r] = vy[ir];
                if (rec_vz) rec_vz[(long)it * c->nrec + r] = vz[ir];
            }
        }
        /* ===== wavefield history (reverse-time migration tests) */
        if (c->hist_field >= 0 && c->hist) {
            REAL *dst = (REAL *)c->hist + (long)it * N;
            if (c->hist_field == QO_PRESSURE) {
#pragma omp parallel for schedule(static)
                for (long ip = 0; ip < N; ip++) dst[ip] = (pxx[ip] + pyy[ip] + pzz[ip]) / (REAL)3;
            } else {
                memcpy(dst, F[c->hist_field], sizeof(REAL) * N);
            }
        }
        /* ===== dv <- V (get_dv_renewed, :1105-1134) */
        for (int f = 0; f < 3; f++) memcpy(dv[f], spec[6 + f], sizeof(REAL) * 2 * N);

        /* ===== snapshot: interior window, ix slowest (:1615-1627) */
        if (snaps && c->snap_every > 0 && it % c->snap_every == 0) {
            const REAL *src = F[c->snap_field];
            REAL *dst = snaps + snap_k * snap_n;
            long q = 0;
            for (int ix = pml; ix < nx - pml; ix++)
                for (int iy = pml; iy < ny - pml; iy++)
                    for (int iz = pml; iz < nz - pml; iz++)
                        dst[q++] = src[(long)ix * sx + (long)iy * sy + iz];
            snap_k++;
        }
    }
#undef SYM

    if (state)
        for (int f = 0; f < 9; f++) memcpy(state + (long)f * N, F[f], sizeof(REAL) * N);

    FN(plan3_free)(&plan);
    for (int f = 0; f < 9; f++) { free(F[f]); free(part[f]); free(spec[f]); }
    for (int f = 0; f < 3; f++) free(dv[f]);
    for (int f = 0; f < 45; f++) free(pv[f]);
    for (int f = 0; f < NPHI; f++) free(phi[f]);
    for (int j = 0; j < 5; j++) { free(cP[j]); free(cS[j]); free(cH[j]); free(cD[j]); }
    free(work); free(A1); free(A2); free(A3);
    free(kxf); free(kyf); free(kzf); free(dpx); free(dmx); free(dpy); free(dmy); free(dpz); free(dmz);
    free(ax); free(bx); free(ay); free(by); free(az); free(bz);
    return 0;
}

#undef FN
#undef CAT_
#undef CAT2_
