/*
 * CPU oracle for the DAW redatuming hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's Fortran, loop for loop, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  It is never part of the product path (the HIP library does not
 * link it and the Python package never imports it).
 *
 * Built by oracle/Makefile with -O2 -ffp-contract=off, OpenMP over grid
 * columns (each point written by one thread) and no -march, so the
 * single-precision functions perform the same IEEE operations in the same
 * order as the reference compiled for x86-64 (no FMA).  The propagators,
 * resampling and MINRES are instantiated twice from daw_oracle_prop.h:
 * *_f32 (reference arithmetic) and *_f64 (the FP64 plain method).
 *
 * Reference files followed (paths relative to the reference tree):
 *   src/kwsinc.f90        BESSI0 :612-635, kwsinc1d :94-179, kwsinc2d :212-378,
 *                         subsourcekws :582-606
 *   src/subacqui.f90      :12-62
 *   src/submodeling2.f90  submodeling :11-288, ..._esource :523-878,
 *                         ..._esource_data :880-1217, setup :1227-1393
 *   src/subwavefieldresample.f90 :1-41
 *   src/subresiduals.f90  :6-18
 *   src/sublinearminres.f90 :7-169, src/minresModule.f90 MINRES :10-360,
 *                         blurring :363-403, filmartin2d :486-614, sphanning :619-632
 *   src/libfourier.f90    REALFT2S :199-276, four2 :282-331
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Grid loops run in parallel (OpenMP) only where every grid point is written
 * by one thread from values no other thread writes in the same loop, so the
 * results are those of the serial loops; small grids stay serial. */
#define OMP_BIG(S) ((long)(S)->n1e * (S)->n2e > 20000)

/* ------------------------------------------------------------------------ */
/* Hicks (Kaiser-windowed sinc) tables, kwsinc.f90                          */
/* ------------------------------------------------------------------------ */

/* BESSI0 (kwsinc.f90:612-635): argument work in REAL*4, polynomial in REAL*8,
 * result rounded to REAL*4. */
float oracle_bessi0(float x)
{
    if (fabsf(x) < 3.75f) {
        float q = x / 3.75f;
        double y = (double)(q * q);
        double r = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
                   y * (0.360768e-1 + y * 0.45813e-2)))));
        return (float)r;
    } else {
        float ax = fabsf(x);
        double y = (double)(3.75f / ax);
        float e = expf(ax) / sqrtf(ax);
        double r = (double)e * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 +
                   y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 +
                   y * 0.392377e-2))))))));
        return (float)r;
    }
}

/* kwsinc1d monopole (kwsinc.f90:94-179): out[(n+10) + 21*(ia-1)], n=-10..10,
 * ia=1..100, the Fortran layout of coef_hicks1D_monopole(-10:10,1:100). */
void oracle_hicks1d(float *out)
{
    const float PI = 3.14159265359f;
    const float b = 14.18f;                 /* bhicks0(irad=10) (:43,:113) */
    float i0d = oracle_bessi0(b);
    for (int ia = 1; ia <= 100; ia++) {
        float alpha = 0.5f - (float)(ia - 1) / 100.0f;
        for (int n = -10; n <= 10; n++) {
            float xn = (float)n + alpha;
            float c = 0.0f;
            if (fabsf(xn) <= 10.0f) {
                if (fabsf(xn) > 1e-15f) {
                    float i0n = oracle_bessi0(b * sqrtf(1.0f - xn * xn / 100.0f));
                    float kaiser = i0n / i0d;
                    c = kaiser * sinf(PI * xn) / (PI * xn);
                    if (fabsf(c) < 1e-15f) c = 0.0f;
                } else {
                    c = 1.0f;
                }
            }
            out[(n + 10) + 21 * (ia - 1)] = c;
        }
    }
}

/*
 * subhicks1d of subresample_hicks.f90:87-157 with irad = 4, nalpha = 10
 * (:40-41): the same Kaiser-windowed sinc with Kaiser beta bhicks0(4) = 6.31
 * (:103-106).  out[(n+4) + 9*(ia-1)], n = -4..4, ia = 1..10.  Entries with
 * |n + alpha| > irad are never written there (:124); here they are 0, the
 * window's value outside its support.
 */
void oracle_subhicks1d(float *out)
{
    const float PI = 3.14159265359f;
    const float b = 6.31f;
    float i0d = oracle_bessi0(b);
    for (int ia = 1; ia <= 10; ia++) {
        float alpha = 0.5f - (float)(ia - 1) / 10.0f;
        for (int n = -4; n <= 4; n++) {
            float xn = (float)n + alpha;
            float c = 0.0f;
            if (fabsf(xn) <= 4.0f) {
                if (fabsf(xn) > 1e-15f) {
                    float i0n = oracle_bessi0(b * sqrtf(1.0f - xn * xn / 16.0f));
                    float kaiser = i0n / i0d;
                    c = kaiser * sinf(PI * xn) / (PI * xn);
                    if (fabsf(c) < 1e-15f) c = 0.0f;
                } else {
                    c = 1.0f;
                }
            }
            out[(n + 4) + 9 * (ia - 1)] = c;
        }
    }
}

/*
 * subresample (subresample_hicks.f90:25-83): xin(n1in,n2) sampled at d1in
 * to xout(n1out,n2) at d1out, each output sample the 9-point windowed-sinc
 * sum around the nearest input sample (:57-75), in REAL arithmetic.
 */
void oracle_subresample(const float *xin, float *xout, int n1in, int n2, int n1out, float d1in, float d1out)
{
    float c[9 * 10];
    oracle_subhicks1d(c);
    for (int i2 = 0; i2 < n2; i2++) {
        const float *x = xin + (size_t)i2 * n1in;
        float *y = xout + (size_t)i2 * n1out;
        for (int i1 = 1; i1 <= n1out; i1++) {
            float tout = (float)(i1 - 1) * d1out;
            int it0 = (int)lroundf(tout / d1in) + 1;
            float t0 = (float)(it0 - 1) * d1in;
            float alpha = (t0 - tout) / d1in;
            int ia = (int)lroundf((0.5f - alpha) * 10.0f) + 1;
            if (ia > 10) ia = 10;
            if (ia < 1) ia = 1;   /* alpha > 0.55: unreachable by rounding; the reference has no guard */
            int jmin = 1 - it0 > -4 ? 1 - it0 : -4;
            int jmax = n1in - it0 < 4 ? n1in - it0 : 4;
            float acc = 0.0f;
            for (int j = jmin; j <= jmax; j++) acc = acc + c[(j + 4) + 9 * (ia - 1)] * x[it0 + j - 1];
            y[i1 - 1] = acc;
        }
    }
}

static float g_h1d[21 * 100];
static int g_h1d_ready = 0;

static const float *hicks1d_cached(void)
{
    if (!g_h1d_ready) {
        oracle_hicks1d(g_h1d);
        g_h1d_ready = 1;
    }
    return g_h1d;
}

/* One 21x21 tile of coef_hicks2D_monopole(:,:,ia1,ia2,ip) as kwsinc2d(ofs)
 * leaves it (kwsinc.f90:262-278 tensor product, :316-364 free-surface
 * anti-mirror).  out[(n1+10) + 21*(n2+10)]. */
void oracle_hicks2d_tile(int ia1, int ia2, int ip, int ofs, float *out)
{
    const float *c1 = hicks1d_cached();
    float temp[441];
    for (int n2 = -10; n2 <= 10; n2++)
        for (int n1 = -10; n1 <= 10; n1++)
            temp[(n1 + 10) + 21 * (n2 + 10)] = c1[(n2 + 10) + 21 * (ia2 - 1)] * c1[(n1 + 10) + 21 * (ia1 - 1)];
    if (ip == 0) {
        memcpy(out, temp, sizeof temp);
        return;
    }
    if (abs(ofs) != 1) {            /* table slice left at its zero initialisation (:68) */
        memset(out, 0, sizeof temp);
        return;
    }
    memcpy(out, temp, sizeof temp);
    int isym = -10 + ip;
    for (int n2 = -10; n2 <= 10; n2++) {
        out[(isym + 10) + 21 * (n2 + 10)] = 0.0f;
        for (int k = 1; k <= ip; k++)
            out[(isym + k + 10) + 21 * (n2 + 10)] =
                temp[(isym + k + 10) + 21 * (n2 + 10)] - temp[(isym - k + 10) + 21 * (n2 + 10)];
    }
}

/* Fortran NINT: nearest integer, halves away from zero */
static int f_nint(float x) { return (int)lroundf(x); }

/* subsourcekws (kwsinc.f90:582-606) for ndim=2, xs=(z,x) */
void oracle_subsourcekws(float z, float x, float h, int ofs, int *ix, int *ialpha, int *nfs)
{
    float xs[2] = {z, x};
    for (int i = 0; i < 2; i++) {
        ix[i] = f_nint(xs[i] / h) + 1;
        float x0 = (float)(ix[i] - 1) * h;
        float alpha = (x0 - xs[i]) / h;
        int ia = f_nint((0.5f - alpha) * 100.0f) + 1;
        ialpha[i] = ia < 100 ? ia : 100;
    }
    int nf = 10 - f_nint(xs[0] / h);
    nf = nf > 0 ? nf : 0;
    if (ofs == 0) nf = 0;
    *nfs = nf;
}

/* subacqui (subacqui.f90:12-62) on coordinate lists: tab is (n,5) Fortran
 * order: ix1+npml, ix2+npml, ialpha1, ialpha2, nalpha_frees. */
void oracle_acqui_table(const float *z, const float *x, int n, float h, int npml, int ofs, int *tab)
{
    for (int i = 0; i < n; i++) {
        int ix[2], ia[2], nf;
        oracle_subsourcekws(z[i], x[i], h, ofs, ix, ia, &nf);
        tab[i] = ix[0] + npml;
        tab[i + n] = ix[1] + npml;
        tab[i + 2 * n] = ia[0];
        tab[i + 3 * n] = ia[1];
        tab[i + 4 * n] = ofs == 0 ? 0 : nf;
    }
}

/* subresiduals (subresiduals.f90:6-18): res(nt-it+1,r) = dobs(it,r)-dcal(it,r) */
void oracle_residuals(const float *dobs, const float *dcal, float *res, int nt, int nr)
{
    for (int r = 0; r < nr; r++)
        for (int it = 0; it < nt; it++)
            res[(size_t)(nt - 1 - it) + (size_t)r * nt] = dobs[(size_t)it + (size_t)r * nt] - dcal[(size_t)it + (size_t)r * nt];
}

/* ------------------------------------------------------------------------ */
/* Numerical-Recipes real FFT as libfourier.f90 writes it                   */
/* ------------------------------------------------------------------------ */

/* four2 (libfourier.f90:282-331): complex FFT of nn points, data 1-based
 * interleaved re/im, REAL*8 throughout. */
static void four2(double *data, int nn, int isign)
{
    double *d = data - 1;   /* 1-based */
    int n = 2 * nn, j = 1;
    for (int i = 1; i <= n; i += 2) {
        if (j > i) {
            double tr = d[j], ti = d[j + 1];
            d[j] = d[i]; d[j + 1] = d[i + 1];
            d[i] = tr; d[i + 1] = ti;
        }
        int m = n / 2;
        while (m >= 2 && j > m) { j -= m; m /= 2; }
        j += m;
    }
    int mmax = 2;
    while (n > mmax) {
        int istep = 2 * mmax;
        double theta = 6.28318530717959 / (double)(isign * mmax);
        double s = sin(0.5 * theta);
        double wpr = -2.0 * (s * s);
        double wpi = sin(theta);
        double wr = 1.0, wi = 0.0;
        for (int m = 1; m <= mmax; m += 2) {
            for (int i = m; i <= n; i += istep) {
                int jj = i + mmax;
                double tempr = wr * d[jj] - wi * d[jj + 1];
                double tempi = wr * d[jj + 1] + wi * d[jj];
                d[jj] = d[i] - tempr;
                d[jj + 1] = d[i + 1] - tempi;
                d[i] = d[i] + tempr;
                d[i + 1] = d[i + 1] + tempi;
            }
            double wtemp = wr;
            wr = wr * wpr - wi * wpi + wr;
            wi = wi * wpr + wtemp * wpi + wi;
        }
        mmax = istep;
    }
}

/* REALFT2S (libfourier.f90:199-276).  C1, C2, WRS, WIS, H1R.. are implicitly
 * typed REAL*4 in the reference, so the unpacking arithmetic is single
 * precision between REAL*8 loads and stores; restated as such. */
void oracle_realft2s(int n1, int isign1, double *dre, double *dim)
{
    double *D = dre - 1, *DI = dim - 1;   /* 1-based */
    int N = n1 / 2;
    int isign = isign1 == 2 ? 1 : -1;
    double theta = 3.141592653589793 / (double)N;
    const float C1 = 0.5f;
    float C2;
    if (isign == -1) {
        double *buf = malloc((size_t)N * sizeof(double));
        for (int i = 0; i < N; i++) buf[i] = D[i + 1];
        int k = 0;
        for (int i = 1; i <= 2 * N; i += 2) {
            k++;
            D[i] = buf[k - 1];
            D[i + 1] = DI[k];
        }
        free(buf);
    }
    if (isign == 1) {
        C2 = -0.5f;
        four2(dre, N, +1);
    } else {
        C2 = 0.5f;
        theta = -theta;
    }
    double s = sin(0.5 * theta);
    double wpr = -2.0 * (s * s);
    double wpi = sin(theta);
    double wr = 1.0 + wpr;
    double wi = wpi;
    int n2p3 = 2 * N + 3;
    for (int i = 2; i <= N / 2; i++) {
        int i1 = 2 * i - 1, i2 = i1 + 1, i3 = n2p3 - i2, i4 = i3 + 1;
        float wrs = (float)wr, wis = (float)wi;
        float h1r = (float)((double)C1 * (D[i1] + D[i3]));
        float h1i = (float)((double)C1 * (D[i2] - D[i4]));
        float h2r = (float)(-((double)C2 * (D[i2] + D[i4])));
        float h2i = (float)((double)C2 * (D[i1] - D[i3]));
        D[i1] = (double)(h1r + wrs * h2r - wis * h2i);
        D[i2] = (double)(h1i + wrs * h2i + wis * h2r);
        D[i3] = (double)(h1r - wrs * h2r + wis * h2i);
        D[i4] = (double)(-h1i + wrs * h2i + wis * h2r);
        double wtemp = wr;
        wr = wr * wpr - wi * wpi + wr;
        wi = wi * wpr + wtemp * wpi + wi;
    }
    if (isign == 1) {
        float h1r = (float)D[1];
        D[1] = (double)h1r + D[2];
        D[2] = (double)h1r - D[2];
    } else {
        float h1r = (float)D[1];
        D[1] = (double)C1 * ((double)h1r + D[2]);
        D[2] = (double)C1 * ((double)h1r - D[2]);
        four2(dre, N, -1);
    }
    if (isign == 1) {
        int k = 0;
        for (int i = 1; i <= 2 * N; i += 2) {
            k++;
            DI[k] = D[i + 1];
            D[k] = D[i];
        }
    }
    if (isign == -1)
        for (int i = 1; i <= 2 * N; i++) D[i] = D[i] / (double)(float)N;
}

/* sphanning (minresModule.f90:619-632), han 1-based of length nf+1 */
static void sphanning(float *han, int nf, int nord)
{
    const float pi = 3.14159265f;
    for (int j = 1; j <= nf + 1; j++) {
        float h1 = 2.0f * pi * (float)(j - 1) / (float)(2 * nf);
        float h2 = (cosf(h1) + 1.0f) / 2.0f + 0.000001f;
        float h3 = logf(h2) * (float)nord;
        han[j] = expf(h3);
    }
}

/* Filter design half of filmartin2d (minresModule.f90:501-586).  dmartin
 * receives nto+1 REAL*8 weights (1-based: dmartin[1..nto+1]); returns nto,
 * or -1 when nt exceeds the reference's length list. */
int oracle_filmartin_design(int nt, float dt, float flow, float fhig, float llow, float lhig, double *dmartin,
                            int *iout)
{
    static const int inum[14] = {8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536};
    int nto = -1;
    for (int i = 0; i < 14; i++)
        if (nt <= inum[i]) { nto = inum[i]; break; }
    if (nto < 0) return -1;
    int nord = 2, nf = nto;
    float fe = 1.0f / dt;
    float fmax = fe / 2.0f;
    if (fhig > fmax) { fhig = fmax; lhig = 0.0f; }
    int conv = (int)((float)nto / fe);
    int iflow = (int)(flow * (float)conv) + 1;
    int ifhig = (int)(fhig * (float)conv) + 1;
    int illow = (int)(llow * (float)conv) + 1;
    int ilhig = (int)(lhig * (float)conv) + 1;
    if (iflow - illow < 1 && iflow > 0) illow = iflow - 1;
    if (ifhig + ilhig > nf) ilhig = nf - ifhig;
    float *fm = calloc((size_t)nf + 2, sizeof(float));
    float *han = calloc((size_t)(illow > ilhig ? illow : ilhig) + 2, sizeof(float));
    if (illow > 0) {
        sphanning(han, illow, nord);
        for (int j = 1; j <= illow; j++) fm[iflow - j + 1] = han[j];
    }
    if (ilhig > 0) {
        sphanning(han, ilhig, nord);
        for (int j = 1; j <= ilhig; j++) fm[ifhig + j - 1] = han[j];
    }
    for (int j = iflow; j <= ifhig; j++) fm[j] = 1.0f;
    for (int j = 1; j <= nf + 1; j++) dmartin[j] = (double)(fm[j] + 0.0001f);
    if (iout) { iout[0] = iflow; iout[1] = ifhig; iout[2] = illow; iout[3] = ilhig; iout[4] = conv; }
    free(fm);
    free(han);
    return nto;
}

/* filmartin2d (minresModule.f90:486-614) on din(nt,nr) -> dout(nt,nr) */
int oracle_filmartin2d(const float *din, int nr, int nt, float dt, float flow, float fhig, float llow,
                       float lhig, float *dout)
{
    double *dm = malloc((size_t)65540 * sizeof(double));
    int nto = oracle_filmartin_design(nt, dt, flow, fhig, llow, lhig, dm, NULL);
    if (nto < 0) { free(dm); return -1; }
    double *pre = malloc((size_t)nto * sizeof(double)), *pim = malloc((size_t)nto * sizeof(double));
    for (int r = 0; r < nr; r++) {
        for (int j = 0; j < nto; j++) { pim[j] = 0.0; pre[j] = 0.0; }
        for (int j = 0; j < nt; j++) pre[j] = (double)din[(size_t)j + (size_t)r * nt];
        oracle_realft2s(nto, 2, pre, pim);
        for (int j = 0; j < nto; j++) {
            pre[j] = pre[j] * dm[j + 1];
            pim[j] = pim[j] * dm[j + 1];
        }
        oracle_realft2s(nto, 1, pre, pim);
        for (int j = 0; j < nt; j++) dout[(size_t)j + (size_t)r * nt] = (float)pre[j];
    }
    free(pre); free(pim); free(dm);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 16-bit snapshot codec (no reference counterpart: the reference keeps its  */
/* snapshots sp, svx, svz in REAL*4, submodeling2.f90:1060-1087).  A value   */
/* v of a 64-column segment with maximum |v| = amax is stored as            */
/* round16(v / 2^e), 2^e the smallest power of two above amax (e >= -100),  */
/* and read back as round16(...) * 2^e.                                     */
/* ------------------------------------------------------------------------ */
/* 1: the long-running solvers print one line per iteration to stderr */
static int oracle_progress = 0;
void oracle_set_progress(int on) { oracle_progress = on; }

float oracle_snap_scale(float amax)
{
    if (!(amax > 0.0f)) return 1.0f;
    int e;
    (void)frexpf(amax, &e);
    return ldexpf(1.0f, e < -100 ? -100 : e);
}

/* Round x (|x| <= 1) to the nearest BF16 (fmt 1) or FP16 (fmt 2) value,
 * ties to even, returned as a float. */
float oracle_snap_round(float x, int fmt)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if (fmt == 1) {                          /* keep 8 significant bits */
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    } else if (fabsf(x) >= 0x1p-14f) {       /* FP16 normal: keep 11 significant bits */
        u += 0x0fffu + ((u >> 13) & 1u);
        u &= 0xffffe000u;
    } else {                                 /* FP16 subnormal: multiples of 2^-24 */
        volatile float t = x + 0.75f;        /* ulp(0.75) = 2^-24 */
        float y = t - 0.75f;
        memcpy(&u, &y, 4);
        if (y == 0.0f) u = signbit(x) ? 0x80000000u : 0u;   /* keep the sign of zero */
    }
    float y;
    memcpy(&y, &u, 4);
    return y;
}

void oracle_snap_round_array(float *x, int n, int fmt)
{
    for (int i = 0; i < n; i++) x[i] = oracle_snap_round(x[i], fmt);
}

/* ------------------------------------------------------------------------ */
/* Precision-generic propagators, resampling, MINRES                         */
/* ------------------------------------------------------------------------ */

#define R float
#define SFX(name) name##_f32
#define EXPR(x) expf(x)
#define SQRTR(x) sqrtf(x)
#include "daw_oracle_prop.h"
#undef R
#undef SFX
#undef EXPR
#undef SQRTR

#define R double
#define SFX(name) name##_f64
#define EXPR(x) exp(x)
#define SQRTR(x) sqrt(x)
#include "daw_oracle_prop.h"
#undef R
#undef SFX
#undef EXPR
#undef SQRTR
