/*
 * Precision-generic part of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Included twice by daw_oracle.c: once with R=float (the reference's own
 * single-precision arithmetic, op-for-op, built with -ffp-contract=off so the
 * x86-64 result matches the reference built without FMA), once with R=double
 * (the FP64 "plain method" against which every FP32 result's error is
 * stated).  Arrays use the reference's Fortran layout: first index fastest,
 * 1-based index helpers below.
 *
 * Required macros: R (scalar type), SFX(name) (name suffix), EXPR(x),
 * SQRTR(x) (exp/sqrt of type R).
 */

/* ---- setup: cspongea / cspongeb (submodeling2.f90:1331-1393) ---------- */
static void SFX(cspongea)(int npml, R fac, R *damp, R *damp1, int n)
{
    /* damp, damp1 are 1-based arrays of length n + 2*npml (index 0 unused) */
    R spg[1024], spg1[1024];
    for (int i = 1; i <= npml; i++) {
        R x = (R)(npml - i);
        R t = fac * x;
        spg[i] = EXPR(-(t * t));
        spg1[i] = EXPR(-((R)0.5 * (t * t)));
    }
    for (int i = 1; i <= npml; i++) {
        damp[i] = spg[i];
        damp1[i] = spg1[i];
        damp[npml + n + i] = spg[npml - i + 1];
        damp1[npml + n + i] = spg1[npml - i + 1];
    }
    for (int i = 1; i <= n; i++) {
        damp[i + npml] = 1;
        damp1[i + npml] = 1;
    }
}

static void SFX(cspongeb)(int npml, R fac, R *damp, R *damp1, int n)
{
    R spg[1024], spg1[1024];
    for (int i = 1; i <= npml - 1; i++) {
        R x = (R)(npml - 1 - i) + (R)0.5;
        R t = fac * x;
        spg[i] = EXPR(-(t * t));
        spg1[i] = EXPR(-((R)0.5 * (t * t)));
    }
    for (int i = 1; i <= npml - 1; i++) {
        damp[i] = spg[i];
        damp1[i] = spg1[i];
    }
    for (int i = 1; i <= n + 1; i++) {
        damp[i + npml - 1] = 1;
        damp1[i + npml - 1] = 1;
    }
    for (int i = 1; i <= npml; i++) {
        R x = (R)(i - 1) + (R)0.5;
        R t = fac * x;
        spg[i] = EXPR(-(t * t));
        spg1[i] = EXPR(-((R)0.5 * (t * t)));
    }
    for (int i = 1; i <= npml; i++) {
        damp[npml + n + i] = spg[i];
        damp1[npml + n + i] = spg1[i];
    }
}

/* Per-call setup of submodeling2.f90:55-110 with rho == 1 (:93).  Only the
 * arrays the time loops read are kept: kappa, kappaw1/2, buw1/2 and bue.
 * bu1 = bu2 = bue = 1 wherever the loops read them (substagger
 * submodeling2.f90:1227-1270 with rho = 1; its uninitialised last row/column
 * is never read by the bounded loops). */
typedef struct {
    int n1, n2, npml, n1e, n2e;
    R dt, dtdx;
    R *damp1a, *damp1a1, *damp1b, *damp1b1;   /* 1-based, n1e */
    R *damp2a, *damp2a1, *damp2b, *damp2b1;   /* 1-based, n2e */
    R *kappa, *kappaw1, *kappaw2, *buw1, *buw2; /* n1e*n2e, Fortran order */
} SFX(setup_t);

static void SFX(setup_free)(SFX(setup_t) *S)
{
    free(S->damp1a); free(S->damp1a1); free(S->damp1b); free(S->damp1b1);
    free(S->damp2a); free(S->damp2a1); free(S->damp2b); free(S->damp2b1);
    free(S->kappa); free(S->kappaw1); free(S->kappaw2); free(S->buw1); free(S->buw2);
}

static void SFX(setup_make)(SFX(setup_t) *S, const float *vp, int n1, int n2, int npml,
                            float h, float dt)
{
    int n1e = n1 + 2 * npml, n2e = n2 + 2 * npml;
    S->n1 = n1; S->n2 = n2; S->npml = npml; S->n1e = n1e; S->n2e = n2e;
    S->dt = (R)dt;
    S->dtdx = (R)dt / (R)h;
    const R xxfac = (R)0.12;    /* real, parameter :: xxfac=0.12 (:35) */
    S->damp1a = calloc(n1e + 1, sizeof(R)); S->damp1a1 = calloc(n1e + 1, sizeof(R));
    S->damp1b = calloc(n1e + 1, sizeof(R)); S->damp1b1 = calloc(n1e + 1, sizeof(R));
    S->damp2a = calloc(n2e + 1, sizeof(R)); S->damp2a1 = calloc(n2e + 1, sizeof(R));
    S->damp2b = calloc(n2e + 1, sizeof(R)); S->damp2b1 = calloc(n2e + 1, sizeof(R));
    SFX(cspongea)(npml, xxfac, S->damp2a, S->damp2a1, n2);
    SFX(cspongeb)(npml, xxfac, S->damp2b, S->damp2b1, n2);
    SFX(cspongea)(npml, xxfac, S->damp1a, S->damp1a1, n1);
    SFX(cspongeb)(npml, xxfac, S->damp1b, S->damp1b1, n1);
    size_t ne = (size_t)n1e * n2e;
    S->kappa = malloc(ne * sizeof(R)); S->kappaw1 = malloc(ne * sizeof(R));
    S->kappaw2 = malloc(ne * sizeof(R)); S->buw1 = malloc(ne * sizeof(R));
    S->buw2 = malloc(ne * sizeof(R));
    for (int i2 = 1; i2 <= n2e; i2++) {
        int j2 = i2 - npml; if (j2 < 1) j2 = 1; if (j2 > n2) j2 = n2;      /* submodext :1281-1321 */
        for (int i1 = 1; i1 <= n1e; i1++) {
            int j1 = i1 - npml; if (j1 < 1) j1 = 1; if (j1 > n1) j1 = n1;
            double v = (double)vp[(size_t)(j1 - 1) + (size_t)(j2 - 1) * n1];
            R kap = (R)(1.0 * v * v);  /* kappa=dble(rho)*dble(vp)**2 (:1235), then real(4) */
            size_t k = (size_t)(i1 - 1) + (size_t)(i2 - 1) * n1e;
            S->kappa[k] = kap;
            S->kappaw2[k] = S->dtdx * kap * S->damp2a1[i2];   /* :105 */
            S->kappaw1[k] = S->dtdx * kap * S->damp1a1[i1];   /* :106 */
            S->buw2[k] = S->dtdx * (R)1 * S->damp2b1[i2];     /* :107, bu2 = 1 */
            S->buw1[k] = S->dtdx * (R)1 * S->damp1b1[i1];     /* :108, bu1 = 1 */
        }
    }
}

#define E(a, i1, i2) (a)[(size_t)((i1) - 1) + (size_t)((i2) - 1) * n1e]
#define I3(a, i1, i2, it) (a)[(size_t)((i1) - 1) + (size_t)((i2) - 1) * n1 + (size_t)((it) - 1) * n1 * n2]

/* This thread's share c0..c1 of the columns lo..hi (all of them outside a
 * parallel region). */
static void SFX(my_cols)(int lo, int hi, int *c0, int *c1)
{
#ifdef _OPENMP
    const long nth = omp_get_num_threads(), th = omp_get_thread_num();
#else
    const long nth = 1, th = 0;
#endif
    const long n = hi - lo + 1;
    *c0 = lo + (int)(n * th / nth);
    *c1 = lo + (int)(n * (th + 1) / nth) - 1;
}

/* pressure update, split PML; sgn=+1 forward (:148-155), -1 back-prop (:657-664) */
static void SFX(p_update_cols)(const SFX(setup_t) *S, R *px, R *pz, const R *vx, const R *vz, int lo, int hi1,
                               int c0, int c1, int sgn)
{
    const int n1e = S->n1e;
    const R a0 = (R)1.125, a1 = -((R)1 / (R)24);   /* :48-49 */
    for (int i2 = c0; i2 <= c1; i2++)
        for (int i1 = lo; i1 <= hi1; i1++) {
            R der2 = a0 * (E(vx, i1, i2) - E(vx, i1, i2 - 1)) + a1 * (E(vx, i1, i2 + 1) - E(vx, i1, i2 - 2));
            R der1 = a0 * (E(vz, i1, i2) - E(vz, i1 - 1, i2)) + a1 * (E(vz, i1 + 1, i2) - E(vz, i1 - 2, i2));
            if (sgn > 0) {
                E(px, i1, i2) = S->damp2a[i2] * E(px, i1, i2) + E(S->kappaw2, i1, i2) * der2;
                E(pz, i1, i2) = S->damp1a[i1] * E(pz, i1, i2) + E(S->kappaw1, i1, i2) * der1;
            } else {
                E(px, i1, i2) = S->damp2a[i2] * E(px, i1, i2) - E(S->kappaw2, i1, i2) * der2;
                E(pz, i1, i2) = S->damp1a[i1] * E(pz, i1, i2) - E(S->kappaw1, i1, i2) * der1;
            }
        }
}

static void SFX(p_update)(SFX(setup_t) *S, R *px, R *pz, const R *vx, const R *vz, int lo, int hi1, int hi2, int sgn)
{
#pragma omp parallel if (OMP_BIG(S))
    {
        int c0, c1;
        SFX(my_cols)(lo, hi2, &c0, &c1);
        SFX(p_update_cols)(S, px, pz, vx, vz, lo, hi1, c0, c1, sgn);
    }
}

/* velocity update; sgn=+1 forward (:229-236), -1 back-prop (:710-717) */
static void SFX(v_update_cols)(const SFX(setup_t) *S, R *vx, R *vz, const R *p, int lo, int hi1, int c0, int c1,
                               int sgn)
{
    const int n1e = S->n1e;
    const R a0 = (R)1.125, a1 = -((R)1 / (R)24);   /* :48-49 */
    for (int i2 = c0; i2 <= c1; i2++)
        for (int i1 = lo; i1 <= hi1; i1++) {
            R der2 = a0 * (E(p, i1, i2 + 1) - E(p, i1, i2)) + a1 * (E(p, i1, i2 + 2) - E(p, i1, i2 - 1));
            R der1 = a0 * (E(p, i1 + 1, i2) - E(p, i1, i2)) + a1 * (E(p, i1 + 2, i2) - E(p, i1 - 1, i2));
            if (sgn > 0) {
                E(vx, i1, i2) = S->damp2b[i2] * E(vx, i1, i2) + E(S->buw2, i1, i2) * der2;
                E(vz, i1, i2) = S->damp1b[i1] * E(vz, i1, i2) + E(S->buw1, i1, i2) * der1;
            } else {
                E(vx, i1, i2) = S->damp2b[i2] * E(vx, i1, i2) - E(S->buw2, i1, i2) * der2;
                E(vz, i1, i2) = S->damp1b[i1] * E(vz, i1, i2) - E(S->buw1, i1, i2) * der1;
            }
        }
}

static void SFX(v_update)(SFX(setup_t) *S, R *vx, R *vz, const R *p, int lo, int hi1, int hi2, int sgn)
{
#pragma omp parallel if (OMP_BIG(S))
    {
        int c0, c1;
        SFX(my_cols)(lo, hi2, &c0, &c1);
        SFX(v_update_cols)(S, vx, vz, p, lo, hi1, c0, c1, sgn);
    }
}

/* dst(i1,i2) = src(i1+npml, i2+npml) over the interior (a snapshot or
 * wavefield slice of n1*n2 values from the extended grid), columns c0..c1 */
static void SFX(grab_cols)(R *dst, const R *src, int n1, int npml, int n1e, int c0, int c1)
{
    for (int i2 = c0; i2 <= c1; i2++)
        for (int i1 = 1; i1 <= n1; i1++)
            dst[(size_t)(i1 - 1) + (size_t)(i2 - 1) * n1] = src[(size_t)(i1 + npml - 1) + (size_t)(i2 + npml - 1) * n1e];
}

static void SFX(grab_interior)(const SFX(setup_t) *S, R *dst, const R *src)
{
#pragma omp parallel if (OMP_BIG(S))
    {
        int c0, c1;
        SFX(my_cols)(1, S->n2, &c0, &c1);
        SFX(grab_cols)(dst, src, S->n1, S->npml, S->n1e, c0, c1);
    }
}

/* Extended source of the re-injection pass (:1121-1125, :1175-1186):
 * f(i1+npml, i2+npml) = f + ((ss*dt)*snap(i1,i2))*w(i1+npml,i2+npml) over
 * i1 = lo..hi1, i2 = lo..hi2, with w = kappa, or w = 1 (bue) when kap is
 * NULL; snap is the n1*n2 slice of the step. */
static void SFX(reinject_cols)(R *f, const R *snap, const R *kap, R ssdt, int n1, int npml, int n1e, int lo, int hi1,
                               int c0, int c1)
{
    for (int i2 = c0; i2 <= c1; i2++)
        for (int i1 = lo; i1 <= hi1; i1++) {
            const size_t e = (size_t)(i1 + npml - 1) + (size_t)(i2 + npml - 1) * n1e;
            const R sv = snap[(size_t)(i1 - 1) + (size_t)(i2 - 1) * n1];
            f[e] = f[e] + ssdt * sv * (kap ? kap[e] : (R)1);
        }
}

static void SFX(reinject)(const SFX(setup_t) *S, R *f, const R *snap, const R *kap, R ssdt, int lo, int hi1, int hi2)
{
#pragma omp parallel if (OMP_BIG(S))
    {
        int c0, c1;
        SFX(my_cols)(lo, hi2, &c0, &c1);
        SFX(reinject_cols)(f, snap, kap, ssdt, S->n1, S->npml, S->n1e, lo, hi1, c0, c1);
    }
}

/* p = px + pz; free surface p(1:npml+1,:) = 0 (:183-189) */
static void SFX(p_sum)(SFX(setup_t) *S, R *p, const R *px, const R *pz, int ofs)
{
    const int n1e = S->n1e, n2e = S->n2e;
    long ne = (long)n1e * n2e;
#pragma omp parallel for schedule(static) if (OMP_BIG(S))
    for (long k = 0; k < ne; k++) p[k] = px[k] + pz[k];
    if (ofs == 1)
        for (int i2 = 1; i2 <= n2e; i2++)
            for (int i1 = 1; i1 <= S->npml + 1; i1++) E(p, i1, i2) = 0;
}

/* vz(npml,:) = vz(npml+1,:) (:260-264) */
static void SFX(vz_mirror)(SFX(setup_t) *S, R *vz, int ofs)
{
    const int n1e = S->n1e;
    if (ofs == 1)
        for (int i2 = 1; i2 <= S->n2e; i2++) E(vz, S->npml, i2) = E(vz, S->npml + 1, i2);
}

/* Hicks window clipped to the extended grid (:167-170).  mode 0: add
 * ((c*dt)*amp)*kappa (source, :175); mode 1: subtract it (receiver
 * residual, :685); mode 2: add ((c*dt)*amp)*bue with bue = 1 (:252). */
static void SFX(hicks_inject)(SFX(setup_t) *S, R *f, const int *tab5, int stride, const float *coef441,
                              R amp, int mode)
{
    const int n1e = S->n1e, n2e = S->n2e;
    int ix1 = tab5[0], ix2 = tab5[stride];
    int r1min = 1 - ix1 > -10 ? 1 - ix1 : -10, r1max = n1e - ix1 < 10 ? n1e - ix1 : 10;
    int r2min = 1 - ix2 > -10 ? 1 - ix2 : -10, r2max = n2e - ix2 < 10 ? n2e - ix2 : 10;
    for (int i1 = r1min; i1 <= r1max; i1++)
        for (int i2 = r2min; i2 <= r2max; i2++) {
            R c = (R)coef441[(i1 + 10) + 21 * (i2 + 10)];
            R t = c * S->dt * amp;
            if (mode == 0) E(f, ix1 + i1, ix2 + i2) = E(f, ix1 + i1, ix2 + i2) + t * E(S->kappa, ix1 + i1, ix2 + i2);
            else if (mode == 1) E(f, ix1 + i1, ix2 + i2) = E(f, ix1 + i1, ix2 + i2) - t * E(S->kappa, ix1 + i1, ix2 + i2);
            else E(f, ix1 + i1, ix2 + i2) = E(f, ix1 + i1, ix2 + i2) + t * (R)1;
        }
}

/* hicks_inject mode 1 for every receiver r = 0..nr-1 of (nr,5) table ir with
 * amplitudes amp[r*amp_stride].  Windows of neighbouring receivers overlap,
 * so each grid point must take its adds in receiver order: threads own
 * disjoint column ranges and walk all receivers in order, each applying the
 * part of the window in its columns. */
static void SFX(hicks_inject_receivers)(SFX(setup_t) *S, R *f, const int *ir, int nr, const float *rt,
                                        const R *amp, size_t amp_stride)
{
#pragma omp parallel if (OMP_BIG(S))
    {
#ifdef _OPENMP
        const int nth = omp_get_num_threads(), th = omp_get_thread_num();
#else
        const int nth = 1, th = 0;
#endif
        const int n1e = S->n1e, n2e = S->n2e;
        const int c0 = 1 + (int)((long)n2e * th / nth), c1 = (int)((long)n2e * (th + 1) / nth);
        for (int r = 0; r < nr; r++) {
            const int *tab5 = ir + r;
            const float *coef441 = rt + (size_t)r * 441;
            const R amp_r = amp[(size_t)r * amp_stride];
            int ix1 = tab5[0], ix2 = tab5[nr];
            int r1min = 1 - ix1 > -10 ? 1 - ix1 : -10, r1max = n1e - ix1 < 10 ? n1e - ix1 : 10;
            int r2min = 1 - ix2 > -10 ? 1 - ix2 : -10, r2max = n2e - ix2 < 10 ? n2e - ix2 : 10;
            if (ix2 + r2min < c0) r2min = c0 - ix2;
            if (ix2 + r2max > c1) r2max = c1 - ix2;
            for (int i1 = r1min; i1 <= r1max; i1++)
                for (int i2 = r2min; i2 <= r2max; i2++) {
                    R c = (R)coef441[(i1 + 10) + 21 * (i2 + 10)];
                    R t = c * S->dt * amp_r;
                    E(f, ix1 + i1, ix2 + i2) = E(f, ix1 + i1, ix2 + i2) - t * E(S->kappa, ix1 + i1, ix2 + i2);
                }
        }
    }
}

/* Hicks record sis(it,r) += c*p over the extended-grid-clipped window (:199-224) */
static R SFX(hicks_record)(SFX(setup_t) *S, const R *p, const int *tab5, int stride, const float *coef441)
{
    const int n1e = S->n1e, n2e = S->n2e;
    int ix1 = tab5[0], ix2 = tab5[stride];
    int r1min = 1 - ix1 > -10 ? 1 - ix1 : -10, r1max = n1e - ix1 < 10 ? n1e - ix1 : 10;
    int r2min = 1 - ix2 > -10 ? 1 - ix2 : -10, r2max = n2e - ix2 < 10 ? n2e - ix2 : 10;
    R acc = 0;
    for (int i1 = r1min; i1 <= r1max; i1++)
        for (int i2 = r2min; i2 <= r2max; i2++)
            acc = acc + (R)coef441[(i1 + 10) + 21 * (i2 + 10)] * E(p, ix1 + i1, ix2 + i2);
    return acc;
}

typedef struct { R *p, *px, *pz, *vx, *vz; } SFX(fields_t);

static void SFX(fields_alloc)(SFX(fields_t) *F, size_t ne)
{
    F->p = calloc(ne, sizeof(R)); F->px = calloc(ne, sizeof(R)); F->pz = calloc(ne, sizeof(R));
    F->vx = calloc(ne, sizeof(R)); F->vz = calloc(ne, sizeof(R));
}
static void SFX(fields_zero)(SFX(fields_t) *F, size_t ne)
{
    memset(F->p, 0, ne * sizeof(R)); memset(F->px, 0, ne * sizeof(R)); memset(F->pz, 0, ne * sizeof(R));
    memset(F->vx, 0, ne * sizeof(R)); memset(F->vz, 0, ne * sizeof(R));
}
static void SFX(fields_free)(SFX(fields_t) *F)
{
    free(F->p); free(F->px); free(F->pz); free(F->vx); free(F->vz);
}

/* Coefficient tiles for a (nr,5) index table: 441 floats per entry. */
static float *SFX(tiles_for)(const int *tab, int nr, int ofs)
{
    float *t = malloc((size_t)nr * 441 * sizeof(float));
    for (int r = 0; r < nr; r++)
        oracle_hicks2d_tile(tab[r + 2 * nr], tab[r + 3 * nr], tab[r + 4 * nr], ofs, t + (size_t)r * 441);
    return t;
}

/*
 * submodeling (submodeling2.f90:11-288).  sis(nt,nr) out; wavefield(n1,n2,nt)
 * out when non-NULL (the reference always writes it).  is5 is the 5-entry
 * source row, ir the (nr,5) receiver table.
 */
int SFX(oracle_forward)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                        const float *s, int type_src, int type_rec, const int *is5, const int *ir, int nr,
                        R *sis, R *wavefield)
{
    SFX(setup_t) S;
    SFX(setup_make)(&S, vp, n1, n2, npml, h, dt);
    const int n1e = S.n1e, n2e = S.n2e;
    size_t ne = (size_t)n1e * n2e;
    SFX(fields_t) F;
    SFX(fields_alloc)(&F, ne);
    float stile[441];
    oracle_hicks2d_tile(is5[2], is5[3], is5[4], ofs, stile);
    float *rt = SFX(tiles_for)(ir, nr, ofs);
    memset(sis, 0, (size_t)nt * nr * sizeof(R));
    if (wavefield) memset(wavefield, 0, (size_t)n1 * n2 * nt * sizeof(R));
    for (int it = 1; it <= nt; it++) {
        SFX(p_update)(&S, F.px, F.pz, F.vx, F.vz, 3, n1e - 2, n2e - 2, +1);
        if (type_src == 0) SFX(hicks_inject)(&S, F.px, is5, 1, stile, (R)s[it - 1], 0);
        SFX(p_sum)(&S, F.p, F.px, F.pz, ofs);
        if (type_rec == 0) {
            if (wavefield) SFX(grab_interior)(&S, wavefield + (size_t)(it - 1) * n1 * n2, F.p);
#pragma omp parallel for schedule(static) if (OMP_BIG(&S))
            for (int r = 0; r < nr; r++)
                sis[(size_t)(it - 1) + (size_t)r * nt] = SFX(hicks_record)(&S, F.p, ir + r, nr, rt + (size_t)r * 441);
        }
        SFX(v_update)(&S, F.vx, F.vz, F.p, 3, n1e - 2, n2e - 2, +1);
        if (type_src == 1) SFX(hicks_inject)(&S, F.vz, is5, 1, stile, (R)s[it - 1], 2);
        SFX(vz_mirror)(&S, F.vz, ofs);
    }
    free(rt);
    SFX(fields_free)(&F);
    SFX(setup_free)(&S);
    return 0;
}

/*
 * The two-pass Hessian-type operator shared by submodeling_adjoint_esource
 * (submodeling2.f90:523-878, variant=1) and ..._esource_data (:880-1217,
 * variant=0).  Pass 1 back-propagates res with the sign-flipped scheme
 * (bounds 3..n-3) and keeps sp, svx, svz; pass 2 re-propagates forward with
 * their time reverse as an extended source and records sis.  Variant 1
 * restricts the pressure extended source to 2..n-1 (:777-778) and stores
 * wavefieldp (:793-797).  The snapshot arrays are owned here.  If
 * snap_cost is non-NULL it receives sum(0.5*sp**2) (:734-742).
 */
static void SFX(exact_pass1)(SFX(setup_t) *S, int ofs, int nt, const R *res, int nr, const int *ir,
                             const float *rt, int variant, R *sp, R *svx, R *svz);

/* Replaces every snapshot value by its 16-bit round trip (fmt 1 BF16, 2 FP16):
 * per time step and row, one scale per 64 columns of the extended grid
 * (extended 0-based column npml + i2 - 1 in segment (npml + i2 - 1) / 64). */
void SFX(snap_quantize)(R *a, int n1, int n2, int npml, int nt, int fmt)
{
    for (int it = 1; it <= nt; it++)
        for (int i1 = 1; i1 <= n1; i1++)
            for (int k = npml / 64; k <= (npml + n2 - 1) / 64; k++) {
                int lo = 64 * k - npml + 1, hi = 64 * k + 63 - npml + 1;
                if (lo < 1) lo = 1;
                if (hi > n2) hi = n2;
                float amax = 0.0f;
                for (int i2 = lo; i2 <= hi; i2++) amax = fmaxf(amax, fabsf((float)I3(a, i1, i2, it)));
                const float sc = oracle_snap_scale(amax), inv = 1.0f / sc;
                for (int i2 = lo; i2 <= hi; i2++)
                    I3(a, i1, i2, it) = (R)(oracle_snap_round((float)I3(a, i1, i2, it) * inv, fmt) * sc);
            }
}

static _Thread_local R *SFX(snap_buf);
static _Thread_local size_t SFX(snap_cap);

static int SFX(oracle_esource_impl)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt,
                                    float dt, const R *res, int type_rec, int nr, const int *ir, R *sis,
                                    R *wavefieldp, int variant, R *snap_cost, int exact, int snap_fmt)
{
    SFX(setup_t) S;
    SFX(setup_make)(&S, vp, n1, n2, npml, h, dt);
    const int n1e = S.n1e, n2e = S.n2e;
    size_t ne = (size_t)n1e * n2e;
    size_t nint3 = (size_t)n1 * n2 * nt;
    /* The snapshot store is kept per calling thread between calls (pass 1
     * writes every element before pass 2 reads it; the exact variant clears
     * it first): first-touch page faults on ~1.5 GB per call otherwise
     * dominate the operator's time at the 121x401, nt = 2500 size. */
    if (3 * nint3 > SFX(snap_cap)) {
        free(SFX(snap_buf));
        SFX(snap_buf) = malloc(3 * nint3 * sizeof(R));
        SFX(snap_cap) = SFX(snap_buf) ? 3 * nint3 : 0;
    }
    if (!SFX(snap_buf)) { SFX(setup_free)(&S); return -1; }
    R *sp = SFX(snap_buf), *svx = sp + nint3, *svz = svx + nint3;
    if (exact) memset(sp, 0, 3 * nint3 * sizeof(R));
    SFX(fields_t) F;
    SFX(fields_alloc)(&F, ne);
    float *rt = SFX(tiles_for)(ir, nr, ofs);
    /* pass 1 */
    if (exact) SFX(exact_pass1)(&S, ofs, nt, res, nr, ir, rt, variant, sp, svx, svz);
    for (int it = 1; it <= (exact ? 0 : nt); it++) {
        SFX(p_update)(&S, F.px, F.pz, F.vx, F.vz, 3, n1e - 3, n2e - 3, -1);
        SFX(hicks_inject_receivers)(&S, F.px, ir, nr, rt, res + (it - 1), (size_t)nt);
        SFX(p_sum)(&S, F.p, F.px, F.pz, ofs);
        SFX(grab_interior)(&S, sp + (size_t)(it - 1) * n1 * n2, F.p);
        SFX(v_update)(&S, F.vx, F.vz, F.p, 3, n1e - 3, n2e - 3, -1);
        SFX(vz_mirror)(&S, F.vz, ofs);
        SFX(grab_interior)(&S, svx + (size_t)(it - 1) * n1 * n2, F.vx);
        SFX(grab_interior)(&S, svz + (size_t)(it - 1) * n1 * n2, F.vz);
    }
    if (snap_fmt == 1 || snap_fmt == 2) {
        SFX(snap_quantize)(sp, n1, n2, npml, nt, snap_fmt);
        SFX(snap_quantize)(svx, n1, n2, npml, nt, snap_fmt);
        SFX(snap_quantize)(svz, n1, n2, npml, nt, snap_fmt);
    }
    if (snap_cost) {
        R c = 0;
        for (size_t k = 0; k < nint3; k++) c = c + (R)0.5 * (sp[k] * sp[k]);
        *snap_cost = c;
    }
    /* pass 2 */
    SFX(fields_zero)(&F, ne);
    memset(sis, 0, (size_t)nt * nr * sizeof(R));
    if (wavefieldp) memset(wavefieldp, 0, nint3 * sizeof(R));
    const R ss = (R)-1;
    const int plo = variant == 1 ? 2 : 1, phi1 = variant == 1 ? n1 - 1 : n1, phi2 = variant == 1 ? n2 - 1 : n2;
    for (int it = 1; it <= nt; it++) {
        SFX(p_update)(&S, F.px, F.pz, F.vx, F.vz, 3, n1e - 3, n2e - 3, +1);
        int ic = nt - it + 1;
        const R ssdt = ss * S.dt;
        const size_t slice = (size_t)(ic - 1) * n1 * n2;
        SFX(reinject)(&S, F.px, sp + slice, S.kappa, ssdt, plo, phi1, phi2);
        SFX(p_sum)(&S, F.p, F.px, F.pz, ofs);
        if (type_rec == 0) {
            if (wavefieldp) SFX(grab_interior)(&S, wavefieldp + (size_t)(it - 1) * n1 * n2, F.p);
#pragma omp parallel for schedule(static) if (OMP_BIG(&S))
            for (int r = 0; r < nr; r++)
                sis[(size_t)(it - 1) + (size_t)r * nt] = SFX(hicks_record)(&S, F.p, ir + r, nr, rt + (size_t)r * 441);
        }
        SFX(v_update)(&S, F.vx, F.vz, F.p, 3, n1e - 3, n2e - 3, +1);
        SFX(reinject)(&S, F.vx, svx + slice, NULL, ssdt, 1, n1, n2);
        SFX(reinject)(&S, F.vz, svz + slice, NULL, ssdt, 1, n1, n2);
        SFX(vz_mirror)(&S, F.vz, ofs);
    }
    free(rt);
    SFX(fields_free)(&F);
    SFX(setup_free)(&S);
    return 0;
}

int SFX(oracle_esource_data)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                             const R *res, int type_rec, int nr, const int *ir, R *sis)
{
    return SFX(oracle_esource_impl)(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis, NULL, 0, NULL, 0, 0);
}

/* esource_data with the snapshots kept in 16-bit (snap_fmt 1 BF16, 2 FP16). */
int SFX(oracle_esource_data_snap)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                                  const R *res, int type_rec, int nr, const int *ir, R *sis, int snap_fmt)
{
    return SFX(oracle_esource_impl)(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis, NULL, 0, NULL, 0,
                                    snap_fmt);
}

/* Exact-Hessian variant (no reference counterpart): pass 1 is the transpose
 * of pass 2 applied to the time-reversed input, so that
 * Rev o esource_data_exact is symmetric. */
int SFX(oracle_esource_data_exact)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                                   const R *res, int nr, const int *ir, R *sis)
{
    return SFX(oracle_esource_impl)(vp, n1, n2, npml, h, ofs, nt, dt, res, 0, nr, ir, sis, NULL, 0, NULL, 1, 0);
}

int SFX(oracle_esource)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                        const R *res, int type_rec, int nr, const int *ir, R *sis, R *wavefieldp, R *cost)
{
    return SFX(oracle_esource_impl)(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis, wavefieldp, 1, cost, 0, 0);
}

/* subwavefieldresample (subwavefieldresample.f90:1-41): window clipped to the
 * interior grid, indices ir - npml. */
int SFX(oracle_resample)(int n1, int n2, int npml, int nt, const int *ir, int nr, int ofs, R *sis,
                         const R *wavefield)
{
    float *rt = SFX(tiles_for)(ir, nr, ofs);
#pragma omp parallel for schedule(static) if ((long)nt * nr > 4096)
    for (int it = 1; it <= nt; it++)
        for (int r = 0; r < nr; r++) {
            int ix1 = ir[r] - npml, ix2 = ir[r + nr] - npml;
            int r1min = 1 - ix1 > -10 ? 1 - ix1 : -10, r1max = n1 - ix1 < 10 ? n1 - ix1 : 10;
            int r2min = 1 - ix2 > -10 ? 1 - ix2 : -10, r2max = n2 - ix2 < 10 ? n2 - ix2 : 10;
            R acc = 0;
            const float *c = rt + (size_t)r * 441;
            for (int i1 = r1min; i1 <= r1max; i1++)
                for (int i2 = r2min; i2 <= r2max; i2++)
                    acc = acc + (R)c[(i1 + 10) + 21 * (i2 + 10)] * I3(wavefield, ix1 + i1, ix2 + i2, it);
            sis[(size_t)(it - 1) + (size_t)r * nt] = acc;
        }
    free(rt);
    return 0;
}

/* datanorm (sublinearminres.f90:117-133): t outer, r inner */
static R SFX(datanorm)(const R *d, int nt, int nr)
{
    R acc = 0;
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < nr; j++) acc = acc + d[(size_t)i + (size_t)j * nt] * d[(size_t)i + (size_t)j * nt];
    return SQRTR(acc);
}

static R SFX(dot)(const R *a, const R *b, size_t n)
{
    R acc = 0;
    for (size_t i = 0; i < n; i++) acc = acc + a[i] * b[i];
    return acc;
}

/* blurring (minresModule.f90:363-403): esource_data then time reversal */
static int SFX(blurring)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                         int type_rec, int nr, const int *ir, const R *x, R *y, R *tmp)
{
    int rc = SFX(oracle_esource_data)(vp, n1, n2, npml, h, ofs, nt, dt, x, type_rec, nr, ir, tmp);
    for (int r = 0; r < nr; r++)
        for (int j = 0; j < nt; j++) y[(size_t)j + (size_t)r * nt] = tmp[(size_t)(nt - 1 - j) + (size_t)r * nt];
    return rc;
}

/*
 * sublinearminres (sublinearminres.f90:7-86) with MINRES
 * (minresModule.f90:10-360) as it runs there: checkA=.false., precon=.false.
 * (the `normal` flag sits in the precon slot, :74), x0 ignored (:76).  The
 * Arnorml-initialising blurring (:132-133) does not influence any output
 * and is skipped.  hist[k] = qrnorm/beta1 after iteration k+1 (the value the
 * log prints, :323).  out[0..6] = fac, Anorm, Acond, rnorm, Arnorm, ynorm,
 * beta1; iout[0..1] = istop, itn.
 */
/* Optional per-iteration record of what MINRES's log shows (:270-285,
 * :308-325): trace[5*(itn-1) + 0..4] = x(1), qrnorm/beta1, Anorm, Acond and 1
 * if the line is printed.  The print rule uses the reference's REAL epsilon
 * in both precisions.  NULL (the default) records nothing. */
static R *SFX(minres_trace);
void SFX(oracle_set_minres_trace)(R *trace) { SFX(minres_trace) = trace; }

int SFX(oracle_linearminres)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                             int type_rec, int nr, const int *ir, float mu, int itnlim, float rtol,
                             const R *deltadr, R *x, R *hist, R *out, int *iout)
{
    size_t n = (size_t)nt * nr;
    R *b = malloc(n * sizeof(R)), *de = malloc(n * sizeof(R));
    R *r1 = malloc(n * sizeof(R)), *r2 = malloc(n * sizeof(R)), *v = malloc(n * sizeof(R));
    R *w = malloc(n * sizeof(R)), *w1 = malloc(n * sizeof(R)), *w2 = malloc(n * sizeof(R));
    R *y = malloc(n * sizeof(R)), *tmp = malloc(n * sizeof(R));
    memcpy(b, deltadr, n * sizeof(R));
    int rc = SFX(oracle_esource_data)(vp, n1, n2, npml, h, ofs, nt, dt, b, type_rec, nr, ir, de);
    R ndr = SFX(datanorm)(b, nt, nr), nde = SFX(datanorm)(de, nt, nr);
    R fac = ndr / nde * (R)mu;
    R shift = -fac;
    const R eps = (R)(sizeof(R) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16);
    int istop = 0, itn = 0;
    R Anorm = 0, Acond = 0, rnorm = 0, ynorm = 0, Arnorml = 0;
    for (size_t i = 0; i < n; i++) x[i] = 0;
    memcpy(y, b, n * sizeof(R));
    memcpy(r1, b, n * sizeof(R));
    R beta1 = SFX(dot)(b, y, n);
    if (beta1 < 0) { istop = 8; goto done; }
    if (beta1 == 0) { istop = 0; goto done; }
    beta1 = SQRTR(beta1);
    {
        R oldb = 0, beta = beta1, dbar = 0, epsln = 0, qrnorm = beta1, phibar = beta1, rhs1 = beta1, rhs2 = 0;
        R tnorm2 = 0, ynorm2 = 0, cs = -1, sn = 0, gmax = 0, gmin = 0;
        for (size_t i = 0; i < n; i++) { w[i] = 0; w2[i] = 0; r2[i] = r1[i]; }
        for (;;) {
            itn++;
            R s = (R)1 / beta;
            for (size_t i = 0; i < n; i++) v[i] = s * y[i];
            rc |= SFX(blurring)(vp, n1, n2, npml, h, ofs, nt, dt, type_rec, nr, ir, v, y, tmp);
            for (size_t i = 0; i < n; i++) y[i] = y[i] - shift * v[i];
            if (itn >= 2) {
                R c = beta / oldb;
                for (size_t i = 0; i < n; i++) y[i] = y[i] - c * r1[i];
            }
            R alfa = SFX(dot)(v, y, n);
            R c2 = alfa / beta;
            for (size_t i = 0; i < n; i++) y[i] = y[i] - c2 * r2[i];
            memcpy(r1, r2, n * sizeof(R));
            memcpy(r2, y, n * sizeof(R));
            oldb = beta;
            beta = SFX(dot)(r2, y, n);
            if (beta < 0) { istop = 6; break; }
            beta = SQRTR(beta);
            tnorm2 = tnorm2 + alfa * alfa + oldb * oldb + beta * beta;
            if (itn == 1) {
                if (beta / beta1 <= (R)10 * eps) istop = -1;
                gmax = alfa < 0 ? -alfa : alfa;
                gmin = gmax;
            }
            R oldeps = epsln;
            R delta = cs * dbar + sn * alfa;
            R gbar = sn * dbar - cs * alfa;
            epsln = sn * beta;
            dbar = -cs * beta;
            R gamma = SQRTR(gbar * gbar + beta * beta);
            cs = gbar / gamma;
            sn = beta / gamma;
            R phi = cs * phibar;
            phibar = sn * phibar;
            R denom = (R)1 / gamma;
            for (size_t i = 0; i < n; i++) {
                w1[i] = w2[i];
                w2[i] = w[i];
                w[i] = (v[i] - oldeps * w1[i] - delta * w2[i]) * denom;
                x[i] = x[i] + phi * w[i];
            }
            gmax = gmax > gamma ? gmax : gamma;
            gmin = gmin < gamma ? gmin : gamma;
            R z = rhs1 / gamma;
            ynorm2 = z * z + ynorm2;
            rhs1 = rhs2 - delta * z;
            rhs2 = -epsln * z;
            Anorm = SQRTR(tnorm2);
            ynorm = SQRTR(ynorm2);
            qrnorm = phibar;
            R rnorml = rnorm;
            rnorm = qrnorm;
            R rootl = SQRTR(gbar * gbar + dbar * dbar);
            Arnorml = rnorml * rootl;
            Acond = gmax / gmin;
            if (hist) hist[itn - 1] = qrnorm / beta1;
            if (oracle_progress) fprintf(stderr, "[oracle] MINRES itn %d qrnorm/beta1 %.6e\n", itn, (double)(qrnorm / beta1));
            if (istop == 0) {
                if (itn >= itnlim) istop = 5;
                if (qrnorm / beta1 <= (R)rtol) istop = 1;
            }
            if (SFX(minres_trace) && itn <= itnlim) {
                const R feps = (R)1.1920928955078125e-07;
                R epsx = Anorm * ynorm * feps, epsr = Anorm * ynorm * (R)rtol, rel = rootl / Anorm;
                int prnt = n <= 40 || itn <= 10 || itn >= itnlim - 10 || itn % 10 == 0 ||
                           qrnorm <= 10 * epsx || qrnorm <= 10 * epsr || rel <= 10 * epsx || rel <= 10 * epsr ||
                           Acond >= (R)1.0e-2 / feps || istop != 0;
                R *t = SFX(minres_trace) + (size_t)5 * (itn - 1);
                t[0] = x[0]; t[1] = qrnorm / beta1; t[2] = Anorm; t[3] = Acond; t[4] = (R)prnt;
            }
            if (istop != 0) break;
        }
    }
done:
    if (out) {
        out[0] = fac; out[1] = Anorm; out[2] = Acond; out[3] = rnorm; out[4] = Arnorml; out[5] = ynorm;
        out[6] = beta1;
    }
    if (iout) { iout[0] = istop; iout[1] = itn; }
    free(b); free(de); free(r1); free(r2); free(v); free(w); free(w1); free(w2); free(y); free(tmp);
    return rc;
}


/*
 * Exact discrete adjoint of submodeling (no reference counterpart: the
 * reference back-propagates with a sign-flipped forward scheme,
 * submodeling2.f90:657-717, which is not the transpose).  For the linear map
 * F: s(nt) -> sis(nt,nr) of the forward scheme with type_src = type_rec = 0,
 * this returns g = F^T d, the transpose applied step by step in reverse time:
 * mirror, velocity update, record, free surface, source injection, pressure
 * update.  The dot test <F s, d> = <s, F^T d> pins it.
 */
int SFX(oracle_forward_adjoint)(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                                const R *d, const int *is5, const int *ir, int nr, R *g)
{
    SFX(setup_t) S;
    SFX(setup_make)(&S, vp, n1, n2, npml, h, dt);
    const int n1e = S.n1e, n2e = S.n2e;
    const size_t ne = (size_t)n1e * n2e;
    const R a0 = (R)1.125, a1 = -((R)1 / (R)24);
    R *qpx = calloc(ne, sizeof(R)), *qpz = calloc(ne, sizeof(R)), *qvx = calloc(ne, sizeof(R));
    R *qvz = calloc(ne, sizeof(R)), *qp = calloc(ne, sizeof(R)), *ux = calloc(ne, sizeof(R)), *uz = calloc(ne, sizeof(R));
    float stile[441];
    oracle_hicks2d_tile(is5[2], is5[3], is5[4], ofs, stile);
    float *rt = SFX(tiles_for)(ir, nr, ofs);
    const int lo = 3, hi1 = n1e - 2, hi2 = n2e - 2;     /* forward stencil box */
    for (int it = nt; it >= 1; it--) {
        /* mirror vz(npml,:) = vz(npml+1,:) */
        if (ofs == 1)
            for (int i2 = 1; i2 <= n2e; i2++) {
                E(qvz, npml + 1, i2) = E(qvz, npml + 1, i2) + E(qvz, npml, i2);
                E(qvz, npml, i2) = 0;
            }
        /* velocity update: v' = damp_b v + buw D+ p on the box */
        memset(qp, 0, ne * sizeof(R));
        memset(ux, 0, ne * sizeof(R));
        memset(uz, 0, ne * sizeof(R));
        for (int i2 = lo; i2 <= hi2; i2++)
            for (int i1 = lo; i1 <= hi1; i1++) {
                E(ux, i1, i2) = E(S.buw2, i1, i2) * E(qvx, i1, i2);
                E(uz, i1, i2) = E(S.buw1, i1, i2) * E(qvz, i1, i2);
                E(qvx, i1, i2) = S.damp2b[i2] * E(qvx, i1, i2);
                E(qvz, i1, i2) = S.damp1b[i1] * E(qvz, i1, i2);
            }
        for (int i2 = lo; i2 <= hi2; i2++)
            for (int i1 = lo; i1 <= hi1; i1++) {
                R u = E(ux, i1, i2), w = E(uz, i1, i2);
                E(qp, i1, i2 + 1) += a0 * u; E(qp, i1, i2) -= a0 * u; E(qp, i1, i2 + 2) += a1 * u; E(qp, i1, i2 - 1) -= a1 * u;
                E(qp, i1 + 1, i2) += a0 * w; E(qp, i1, i2) -= a0 * w; E(qp, i1 + 2, i2) += a1 * w; E(qp, i1 - 1, i2) -= a1 * w;
            }
        /* record sis(it,r) = sum c p  ->  qp += c d(it,r) */
        for (int r = 0; r < nr; r++) {
            const int ix1 = ir[r], ix2 = ir[r + nr];
            const int r1min = 1 - ix1 > -10 ? 1 - ix1 : -10, r1max = n1e - ix1 < 10 ? n1e - ix1 : 10;
            const int r2min = 1 - ix2 > -10 ? 1 - ix2 : -10, r2max = n2e - ix2 < 10 ? n2e - ix2 : 10;
            const R dv = d[(size_t)(it - 1) + (size_t)r * nt];
            for (int i1 = r1min; i1 <= r1max; i1++)
                for (int i2 = r2min; i2 <= r2max; i2++)
                    E(qp, ix1 + i1, ix2 + i2) += (R)rt[(size_t)r * 441 + (i1 + 10) + 21 * (i2 + 10)] * dv;
        }
        /* p = M (px + pz) */
        for (int i2 = 1; i2 <= n2e; i2++)
            for (int i1 = 1; i1 <= n1e; i1++) {
                R q = (ofs == 1 && i1 <= npml + 1) ? (R)0 : E(qp, i1, i2);
                E(qpx, i1, i2) += q;
                E(qpz, i1, i2) += q;
            }
        /* source injection px += c dt s(it) kappa  ->  g(it) */
        {
            const int ix1 = is5[0], ix2 = is5[1];
            const int r1min = 1 - ix1 > -10 ? 1 - ix1 : -10, r1max = n1e - ix1 < 10 ? n1e - ix1 : 10;
            const int r2min = 1 - ix2 > -10 ? 1 - ix2 : -10, r2max = n2e - ix2 < 10 ? n2e - ix2 : 10;
            R acc = 0;
            for (int i1 = r1min; i1 <= r1max; i1++)
                for (int i2 = r2min; i2 <= r2max; i2++)
                    acc += (R)stile[(i1 + 10) + 21 * (i2 + 10)] * S.dt * E(S.kappa, ix1 + i1, ix2 + i2) *
                           E(qpx, ix1 + i1, ix2 + i2);
            g[it - 1] = acc;
        }
        /* pressure update: p' = damp_a p + kappaw D- v on the box */
        memset(ux, 0, ne * sizeof(R));
        memset(uz, 0, ne * sizeof(R));
        for (int i2 = lo; i2 <= hi2; i2++)
            for (int i1 = lo; i1 <= hi1; i1++) {
                E(ux, i1, i2) = E(S.kappaw2, i1, i2) * E(qpx, i1, i2);
                E(uz, i1, i2) = E(S.kappaw1, i1, i2) * E(qpz, i1, i2);
                E(qpx, i1, i2) = S.damp2a[i2] * E(qpx, i1, i2);
                E(qpz, i1, i2) = S.damp1a[i1] * E(qpz, i1, i2);
            }
        for (int i2 = lo; i2 <= hi2; i2++)
            for (int i1 = lo; i1 <= hi1; i1++) {
                R u = E(ux, i1, i2), w = E(uz, i1, i2);
                E(qvx, i1, i2) += a0 * u; E(qvx, i1, i2 - 1) -= a0 * u; E(qvx, i1, i2 + 1) += a1 * u; E(qvx, i1, i2 - 2) -= a1 * u;
                E(qvz, i1, i2) += a0 * w; E(qvz, i1 - 1, i2) -= a0 * w; E(qvz, i1 + 1, i2) += a1 * w; E(qvz, i1 - 2, i2) -= a1 * w;
            }
    }
    free(rt);
    free(qpx); free(qpz); free(qvx); free(qvz); free(qp); free(ux); free(uz);
    SFX(setup_free)(&S);
    return 0;
}

/* Transpose of the re-injection pass (box 3..n-3), reverse time, with the
 * adjoint data d(it) = res(nt-it+1) spread by the Hicks tiles; the adjoints
 * of its extended sources become the snapshots sp, svx, svz of index
 * icount = nt-it+1. */
static void SFX(exact_pass1)(SFX(setup_t) *S, int ofs, int nt, const R *res, int nr, const int *ir,
                             const float *rt, int variant, R *sp, R *svx, R *svz)
{
    const int n1e = S->n1e, n2e = S->n2e, n1 = S->n1, n2 = S->n2, npml = S->npml;
    const size_t ne = (size_t)n1e * n2e;
    const R a0 = (R)1.125, a1 = -((R)1 / (R)24);
    R *qpx = calloc(ne, sizeof(R)), *qpz = calloc(ne, sizeof(R)), *qvx = calloc(ne, sizeof(R));
    R *qvz = calloc(ne, sizeof(R)), *qp = calloc(ne, sizeof(R)), *ux = calloc(ne, sizeof(R)), *uz = calloc(ne, sizeof(R));
    const int lo = 3, hi1 = n1e - 3, hi2 = n2e - 3;
    const int plo = variant == 1 ? 2 : 1, phi1 = variant == 1 ? n1 - 1 : n1, phi2 = variant == 1 ? n2 - 1 : n2;
    for (int it = nt; it >= 1; it--) {
        const int ic = nt - it + 1;
        if (ofs == 1)
            for (int i2 = 1; i2 <= n2e; i2++) {
                E(qvz, npml + 1, i2) = E(qvz, npml + 1, i2) + E(qvz, npml, i2);
                E(qvz, npml, i2) = 0;
            }
        /* vx += -dt svx(ic), vz += -dt svz(ic) on the interior */
        for (int i2 = 1; i2 <= n2; i2++)
            for (int i1 = 1; i1 <= n1; i1++) {
                I3(svx, i1, i2, ic) = -S->dt * E(qvx, i1 + npml, i2 + npml);
                I3(svz, i1, i2, ic) = -S->dt * E(qvz, i1 + npml, i2 + npml);
            }
        memset(qp, 0, ne * sizeof(R));
        memset(ux, 0, ne * sizeof(R));
        memset(uz, 0, ne * sizeof(R));
        for (int i2 = lo; i2 <= hi2; i2++)
            for (int i1 = lo; i1 <= hi1; i1++) {
                E(ux, i1, i2) = E(S->buw2, i1, i2) * E(qvx, i1, i2);
                E(uz, i1, i2) = E(S->buw1, i1, i2) * E(qvz, i1, i2);
                E(qvx, i1, i2) = S->damp2b[i2] * E(qvx, i1, i2);
                E(qvz, i1, i2) = S->damp1b[i1] * E(qvz, i1, i2);
            }
        for (int i2 = lo; i2 <= hi2; i2++)
            for (int i1 = lo; i1 <= hi1; i1++) {
                R u = E(ux, i1, i2), w = E(uz, i1, i2);
                E(qp, i1, i2 + 1) += a0 * u; E(qp, i1, i2) -= a0 * u; E(qp, i1, i2 + 2) += a1 * u; E(qp, i1, i2 - 1) -= a1 * u;
                E(qp, i1 + 1, i2) += a0 * w; E(qp, i1, i2) -= a0 * w; E(qp, i1 + 2, i2) += a1 * w; E(qp, i1 - 1, i2) -= a1 * w;
            }
        for (int r = 0; r < nr; r++) {
            const int ix1 = ir[r], ix2 = ir[r + nr];
            const int r1min = 1 - ix1 > -10 ? 1 - ix1 : -10, r1max = n1e - ix1 < 10 ? n1e - ix1 : 10;
            const int r2min = 1 - ix2 > -10 ? 1 - ix2 : -10, r2max = n2e - ix2 < 10 ? n2e - ix2 : 10;
            const R dv = res[(size_t)(nt - it) + (size_t)r * nt];
            for (int i1 = r1min; i1 <= r1max; i1++)
                for (int i2 = r2min; i2 <= r2max; i2++)
                    E(qp, ix1 + i1, ix2 + i2) += (R)rt[(size_t)r * 441 + (i1 + 10) + 21 * (i2 + 10)] * dv;
        }
        for (int i2 = 1; i2 <= n2e; i2++)
            for (int i1 = 1; i1 <= n1e; i1++) {
                R q = (ofs == 1 && i1 <= npml + 1) ? (R)0 : E(qp, i1, i2);
                E(qpx, i1, i2) += q;
                E(qpz, i1, i2) += q;
            }
        /* px += -dt sp(ic) kappa on the pressure range */
        for (int i2 = 1; i2 <= n2; i2++)
            for (int i1 = 1; i1 <= n1; i1++) {
                const int in = i1 >= plo && i1 <= phi1 && i2 >= plo && i2 <= phi2;
                I3(sp, i1, i2, ic) = in ? -S->dt * E(S->kappa, i1 + npml, i2 + npml) * E(qpx, i1 + npml, i2 + npml) : (R)0;
            }
        memset(ux, 0, ne * sizeof(R));
        memset(uz, 0, ne * sizeof(R));
        for (int i2 = lo; i2 <= hi2; i2++)
            for (int i1 = lo; i1 <= hi1; i1++) {
                E(ux, i1, i2) = E(S->kappaw2, i1, i2) * E(qpx, i1, i2);
                E(uz, i1, i2) = E(S->kappaw1, i1, i2) * E(qpz, i1, i2);
                E(qpx, i1, i2) = S->damp2a[i2] * E(qpx, i1, i2);
                E(qpz, i1, i2) = S->damp1a[i1] * E(qpz, i1, i2);
            }
        for (int i2 = lo; i2 <= hi2; i2++)
            for (int i1 = lo; i1 <= hi1; i1++) {
                R u = E(ux, i1, i2), w = E(uz, i1, i2);
                E(qvx, i1, i2) += a0 * u; E(qvx, i1, i2 - 1) -= a0 * u; E(qvx, i1, i2 + 1) += a1 * u; E(qvx, i1, i2 - 2) -= a1 * u;
                E(qvz, i1, i2) += a0 * w; E(qvz, i1 - 1, i2) -= a0 * w; E(qvz, i1 + 1, i2) += a1 * w; E(qvz, i1 - 2, i2) -= a1 * w;
            }
    }
    free(qpx); free(qpz); free(qvx); free(qvz); free(qp); free(ux); free(uz);
}

#undef E
#undef I3
