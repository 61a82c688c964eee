/*
 * daw.h -- C-ABI of the MI355X DAW redatuming library (libdaw.so).
 *
 * Every entry point replaces one routine of the reference's Fortran hot path
 * (paths relative to the reference tree) and keeps its argument meaning:
 * arrays are the reference's column-major arrays passed as plain pointers
 * (first index fastest), integer index tables are the (n,5) src_tab /
 * rec_tab of subacqui.f90:33-58 in Fortran order.  All functions are
 * synchronous, return 0 on success and a negative DAW_ERR_* code otherwise
 * (message in daw_last_error()); the reference has no error returns and
 * aborts instead, which the Fortran shim reproduces with `error stop`.
 *
 * Host-pointer entry points copy to and from HBM internally.  The batched
 * entry point daw_redatum_batch takes device pointers for its large arrays.
 */
#ifndef DAW_H
#define DAW_H

#ifdef __cplusplus
extern "C" {
#endif

#define DAW_ABI_VERSION 2

#define DAW_OK 0
#define DAW_ERR_ARG (-1)     /* invalid argument */
#define DAW_ERR_STATE (-2)   /* kwsinc2d not initialised */
#define DAW_ERR_NOMEM (-3)   /* device allocation failed */
#define DAW_ERR_HIP (-4)     /* HIP runtime error */

const char *daw_last_error(void);
int daw_version(void);

/* ---- module kwsinc (kwsinc.f90) ------------------------------------- */

/* kwsinc2d(ifreesurf), kwsinc.f90:212-378 (with allocatekws2d :64-71).
 * Selects the free-surface flag the Hicks tables are built for; must precede
 * the propagators, as main_redatum_DAW.f90:92-93 orders it.  Coefficient
 * tiles are formed per point on demand instead of as the 194 MB table. */
int daw_kwsinc2d(int ofs);
/* deallocatekws2d, kwsinc.f90:73-77: drops the tables and cached model. */
int daw_deallocatekws2d(void);
/* kwsinc1d monopole table coef_hicks1D_monopole(-10:10,1:100),
 * kwsinc.f90:94-179; out holds 2100 floats, Fortran order. */
int daw_hicks1d(float *out);
/* coef_hicks2D_monopole(:,:,ia1,ia2,ip) as kwsinc2d(ofs) builds it,
 * kwsinc.f90:262-364; out holds 21x21 floats, first index fastest. */
int daw_hicks2d_tile(int ia1, int ia2, int ip, int ofs, float *out);
/* subsourcekws(2, (/z,x/), h, ofs, ix, ialpha, nalpha_frees),
 * kwsinc.f90:582-606. */
int daw_subsourcekws(float z, float x, float h, int ofs, int *ix, int *ialpha, int *nfs);

/* ---- acquisition (subacqui.f90:12-62) -------------------------------- */

/* src_tab/rec_tab rows for n points (z[i], x[i]) -> tab(n,5), Fortran order. */
int daw_subacqui(const float *z, const float *x, int n, float h, int npml, int ofs, int *tab);

/* ---- propagators (submodeling2.f90) ---------------------------------- */

/* submodeling(is0,vp,n1,n2,npml,h,ofs,nt,dt,s,type_src,type_rec,is,ir,nr,sis,wavefield),
 * submodeling2.f90:11-288.  vp(n1,n2); s(nt) already divided by h**2;
 * is5 = src_tab(isrc,1:5); ir(nr,5); out sis(nt,nr); wavefield(n1,n2,nt) may
 * be NULL (the reference always fills it; the batched path never needs it). */
int daw_submodeling(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt, const float *s,
                    int type_src, int type_rec, const int *is5, const int *ir, int nr, float *sis, float *wavefield);

/* submodeling_adjoint_esource(is,vp,n1,n2,npml,h,ofs,nt,dt,res,type_rec,nr,ir,sis,wavefieldp),
 * submodeling2.f90:523-878.  wavefieldp may be NULL. */
int daw_submodeling_adjoint_esource(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                                    const float *res, int type_rec, int nr, const int *ir, float *sis,
                                    float *wavefieldp);

/* submodeling_adjoint_esource_data(is,vp,n1,n2,npml,h,ofs,nt,dt,res,type_rec,nr,ir,sis),
 * submodeling2.f90:880-1217 -- the Hessian-type operator of the MINRES solve. */
int daw_submodeling_adjoint_esource_data(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt,
                                         float dt, const float *res, int type_rec, int nr, const int *ir, float *sis);

/* Exact discrete adjoint of daw_submodeling (type_src = type_rec = 0) with
 * respect to the source wavelet: g(nt) = F^T d for data d(nt,nr).  No
 * reference counterpart -- the reference back-propagates with a
 * sign-flipped forward scheme (submodeling2.f90:657-717), not the transpose;
 * <F s, d> = <s, F^T d> holds here to rounding. */
int daw_submodeling_transpose(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                              const float *d, const int *is5, const int *ir, int nr, float *g);

/* Hessian of the MINRES solve: 0 (default) = the reference's operator
 * (sign-flipped back-propagation then re-injection, submodeling2.f90:880-1217);
 * 1 = exact: the back-propagation pass is replaced by the exact transpose of
 * the re-injection pass, which makes the MINRES operator symmetric. */
int daw_set_hessian(int exact);

/* Arithmetic of the MINRES solve (daw_sublinearminres, daw_redatum_batch,
 * daw_redatum_uneven): 0 (default) = FP32 vectors with FP64 dot products,
 * norms and Lanczos/Givens scalars; 1 = the reference's own arithmetic,
 * REAL throughout as minresModule.f90:10-360 and sublinearminres.f90:64-72,
 * 117-133 run it: sequential FP32 dot products in dot_product's order,
 * datanorm in its t-outer/r-inner order, FP32 scalars and vector updates
 * rounded as written.  Mode 1 reproduces the reference's solve (one lane
 * per shot adds each dot product in order, ~2 ms per 1.2 M-sample product);
 * mode 0 is the more accurate.  No reference counterpart (a knob). */
int daw_set_minres_arith(int arith);

/* Storage of the wavefield snapshots sp, svx, svz that the reference
 * Hessian-type operator keeps between its two passes (submodeling2.f90:1060-1087
 * write, :1164-1185 read): 0 = FP32 as the reference (default), 1 = BF16,
 * 2 = FP16, each 16-bit value scaled by a power of two per 64-column segment
 * of a snapshot row (the segment maximum maps below 1).  16-bit snapshots
 * halve the snapshot store and its traffic; the exact Hessian
 * (daw_set_hessian(1)) always keeps FP32.  No reference counterpart. */
int daw_set_snapshot_format(int fmt);

/* Storage of the MINRES Krylov vectors r1, r2 and w, w1, w2 (the reference's
 * work vectors, minresModule.f90:29) in the default (FP64) MINRES arithmetic:
 * 0 = FP32 (default), 1 = BF16, 2 = FP16, each 256-element chunk scaled by a
 * power of two at or above its max |value|.  Dot products, norms and scalars
 * stay FP64, v, y and x FP32; the 16-bit vectors are decoded where they are
 * read.  Ignored by the reference arithmetic (daw_set_minres_arith(1)).  No
 * reference counterpart. */
int daw_set_krylov_format(int fmt);

/* subwavefieldresample(n1,n2,npml,nt,dt,ir,nr,sis,wavefield),
 * subwavefieldresample.f90:1-41. */
int daw_subwavefieldresample(int n1, int n2, int npml, int nt, float dt, const int *ir, int nr, float *sis,
                             const float *wavefield);

/* subresiduals(is,dataobs,datacal,res,nt,nr), subresiduals.f90:6-18. */
int daw_subresiduals(const float *dobs, const float *dcal, float *res, int nt, int nr);

/* sublinearminres(istot,grid,inv,ir), sublinearminres.f90:7-86, with the
 * grid%/inv% fields it reads passed explicitly: c0(n1,n2), ir(nrec,5),
 * deltadr(nt,nrec) in, deltade(nt,nrec) out.  maxiter is REAL as in
 * grid_type (main_redatum_DAW.f90:14).  hist (may be NULL) receives
 * qrnorm/beta1 per iteration (the column the MINRES log prints,
 * minresModule.f90:323); info (may be NULL) receives 9 doubles: fac, Anorm,
 * Acond, rnorm, Arnorm, ynorm, beta1, istop, itn. */
int daw_sublinearminres(const float *c0, int n1, int n2, int npml, float h, int ofs, int nt, float dt, int type_rec,
                        int nrec, const int *ir, float mu, float maxiter, float error, const float *deltadr,
                        float *deltade, double *hist, double *info);

/* sublinearminres with its log open (grid%oio = 1, sublinearminres.f90:51-61,
 * 78-80): the same solve, and MINRES writes its heading, iteration lines and
 * final status to log_path (MINRES_<shot>.txt in the reference) in the
 * formats of minresModule.f90:346-358.  log_path NULL: no log. */
int daw_sublinearminres_log(const float *c0, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                            int type_rec, int nrec, const int *ir, float mu, float maxiter, float error,
                            const float *deltadr, float *deltade, const char *log_path);

/* The MINRES log of one solve from a log block that the batched entry points
 * return (daw_redatum_args.minres_log_host): 9 doubles (fac, Anorm, Acond,
 * rnorm, Arnorm, ynorm, beta1, istop, itn), then 5 per iteration up to itnlim
 * (x(1), qrnorm/beta1, Anorm, Acond, 1 if MINRES prints the line,
 * minresModule.f90:308-325).  n is the system size, nrec*nt. */
int daw_write_minres_log(const char *path, int n, int itnlim, float rtol, const double *block);

/* subresample(xin,xout,n1in,n2,n1out,d1in,d1out), subresample_hicks.f90:25-83:
 * each trace of xin(n1in,n2), sampled at d1in, resampled to n1out samples at
 * d1out with the 9-point Kaiser-windowed sinc (irad = 4, nalpha = 10).  The
 * reference ships it but does not link it (src/Makefile:24-25); it is the
 * tool for the dt -> dto decimation that sub_redatum_DAW.f90:24 sizes. */
int daw_subresample(const float *xin, float *xout, int n1in, int n2, int n1out, float d1in, float d1out);

/* The same on device traces (xin, xout device pointers, [n2][n1in] and
 * [n2][n1out], not overlapping). */
int daw_subresample_dev(const float *xin, float *xout, int n1in, int n2, int n1out, float d1in, float d1out);

/* daw_subresample_dev that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The resampling plan and table go to a fixed device array of
 * the library (one per GPU, held until the process ends) instead of
 * buffers allocated per call, so n1out is limited to 65536 output samples
 * (DAW_ERR_ARG beyond).  The library's lock serializes calls within a
 * process; the array is the last call's, so calls that overlapped on one
 * GPU would race.  Results are identical to daw_subresample_dev. */
int daw_subresample_dev_noalloc(const float *xin, float *xout, int n1in, int n2, int n1out, float d1in, float d1out);

/* filmartin2d(deltadeo,nr,nt,dt,flow,fhig,llow,lhig,deltade),
 * minresModule.f90:486-614, on host traces: copied to the GPU, filtered as
 * daw_filmartin2d_dev does, copied back. */
int daw_filmartin2d(const float *din, int nr, int nt, float dt, float flow, float fhig, float llow, float lhig,
                    float *dout);

/* filmartin2d on device traces din/dout [nr][nt] (may alias).  Every FFT
 * length of the reference's list (up to 65536) runs on the GPU in FP64:
 * in LDS up to 8192, as a four-step transform through an HBM work area of up
 * to 256 MB for 16384-65536. */
int daw_filmartin2d_dev(const float *din, int nr, int nt, float dt, float flow, float fhig, float llow, float lhig,
                        float *dout);

/* ---- batched redatuming (no reference counterpart: it replaces the
 * para_range shot loop of sub_redatum_DAW.f90:30-155 plus mode 2
 * :230-238 for a block of shots advanced together on one GPU) ---------- */
typedef struct daw_redatum_args {
    const float *c0;          /* host, background model c0(n1,n2) */
    int n1, n2, npml;
    float h, dt;
    int ofs, nt, type_src;
    const float *s_dev;       /* device, wavelet s(nt), already / h**2 */
    int nshots;               /* shots redatumed by this call */
    int nsrc_total;           /* leading dimension of src_tab */
    const int *src_tab;       /* host, src_tab(nsrc_total,5); rows 1..nshots used */
    int nrec;
    const int *rec_tab;       /* host, rec_tab(nrec,5) acquisition receivers */
    int ndat;
    const int *dat_tab;       /* host, datum receivers (ndat,5), may be NULL if ndat = 0 */
    const float *dobs_dev;    /* device, observed data [shot][nrec][nt] */
    float mu;
    int maxiter;
    float rtol;               /* grid%error */
    float *datum_dev;         /* device out, redatumed gathers [shot][ndat][nt] (band-passed when filter = 1) */
    float *datae_dev;         /* device out or NULL: datar + dataesc [shot][nrec][nt] */
    float *deltade_dev;       /* device out or NULL: MINRES solutions [shot][nrec][nt] */
    double *hist_host;        /* host out or NULL: qrnorm/beta1 [shot][maxiter] */
    int max_batch;            /* shots advanced together; <= 0: admit by the HBM budget, at most
                                 as many as keep a launch's wave state within ~192 MB, in
                                 multiples of 8 (the XCD count) where that fits */
    double hbm_budget_bytes;  /* <= 0: 85 % of free HBM */
    int filter;               /* 1: band-pass the gathers as mode 2 does (filmartin2d) */
    float flow, fhig, llow, lhig;
    double *minres_log_host;  /* host out or NULL: MINRES log blocks [shot][9 + 5*maxiter]
                                 (layout at daw_write_minres_log) */
    int method;               /* 0 or 1: DAW (mode 1, :73-155); 3: Symes (mode 3, :157-228):
                                 no background forward, the residual is dobs time-reversed,
                                 datum and datae carry u^e / dataesc alone; s_dev may be NULL */
} daw_redatum_args;

int daw_redatum_batch(const daw_redatum_args *args);

/* Uneven batch (no reference counterpart; the reference runs every shot on
 * the whole model with one receiver table and one nt): each shot s is
 * redatumed on its own subdomain, columns x0[s] .. x0[s]+n2s[s]-1 of c0, with
 * its own receivers and nt[s] -- exactly what the reference computes when
 * run on that cropped model, acquisition and nt -- and all shots of a
 * sub-batch advance together in one launch per time step.  Tables and trace
 * blocks of consecutive shots follow one another. */
typedef struct daw_uneven_args {
    const float *c0;          /* host, full model c0(n1,n2) */
    int n1, n2, npml;
    float h, dt;
    int ofs, type_src;
    const float *s_dev;       /* device, wavelet of max(nts) samples, already / h**2 */
    int nshots;
    const int *x0;            /* host [nshots]: first column (0-based) of each subdomain */
    const int *n2s;           /* host [nshots]: subdomain widths */
    const int *nts;           /* host [nshots]: time steps per shot */
    const int *src_tab;       /* host (nshots,5), subdomain coordinates (subacqui on the subdomain) */
    const int *nrecs;         /* host [nshots] */
    const int *rec_tabs;      /* host: shot s's (nrecs[s],5) table, one after another */
    const int *ndats;         /* host [nshots] or NULL: datum receivers per shot */
    const int *dat_tabs;      /* host or NULL: shot s's (ndats[s],5) table, one after another */
    const float *dobs_dev;    /* device: shot s's [nrecs[s]][nts[s]] traces, one after another */
    float mu;
    int maxiter;
    float rtol;
    float *datum_dev;         /* device out or NULL: [ndats[s]][nts[s]] per shot, one after another */
    float *datae_dev;         /* device out or NULL: as dobs_dev */
    float *deltade_dev;       /* device out or NULL: as dobs_dev */
    double *hist_host;        /* host out or NULL: [shot][maxiter] */
    int max_batch;            /* shots per sub-batch; <= 0: by the HBM budget and the ~192 MB state cap */
    double hbm_budget_bytes;  /* <= 0: 85 % of free HBM */
    int filter;
    float flow, fhig, llow, lhig;
    double *minres_log_host;  /* host out or NULL: [shot][9 + 5*maxiter], as daw_redatum_args */
    int method;               /* as daw_redatum_args */
} daw_uneven_args;

int daw_redatum_uneven(const daw_uneven_args *args);

/* Mode-0 forward modelling of sub_redatum_DAW.f90:55-65 (submodeling per
 * shot, submodeling2.f90:11-288) for a block of shots advanced together:
 * dobs_dev [shot][nrec][nt]; with datum receivers also the datum gather of
 * each shot's wavefield (what mode 2 resamples, subwavefieldresample.f90)
 * without storing the wavefield.  Device pointers for s/dobs/datum. */
typedef struct daw_modeling_args {
    const float *vp;          /* host, vp(n1,n2) */
    int n1, n2, npml;
    float h, dt;
    int ofs, nt, type_src, type_rec;
    const float *s_dev;
    int nshots, nsrc_total;
    const int *src_tab;
    int nrec;
    const int *rec_tab;
    int ndat;
    const int *dat_tab;
    float *dobs_dev;          /* device out [shot][nrec][nt] */
    float *datum_dev;         /* device out [shot][ndat][nt] or NULL */
    int max_batch;            /* <= 0: all shots at once (up to 256) */
} daw_modeling_args;

int daw_modeling_batch(const daw_modeling_args *args);

/* ---- the reference driver's files (host only, no GPU work) ------------- */

/* redatum.par as subreadpar reads it (subreadpar.f90:9-21: one list-directed
 * READ per line).  maxiter is REAL as in grid_type (main_redatum_DAW.f90:14);
 * names are NUL-terminated. */
typedef struct daw_par {
    int mode;
    char name_true[160], name_init[160], name_acqui[160], name_src[160];
    float flow, fhig, llow, lhig;
    int n1, n2, npml;
    float h;
    int nt;
    float dt, dto;
    int type_src, type_rec, ofs;
    float mu;
    int oio;
    float error, maxiter;
} daw_par;

int daw_readpar(const char *path, daw_par *par);

/* The acquisition file: "nsrc nrec" (main_redatum_DAW.f90:120-123), then one
 * "z x" record per source and per receiver (subacqui.f90:29-57).  Counts go
 * to *nsrc, *nrec; src_zx / rec_zx (NULL to skip) receive z,x pairs and must
 * hold src_cap / rec_cap points.  daw_subacqui turns them into tables. */
int daw_read_acqui(const char *path, int *nsrc, int *nrec, float *src_zx, int src_cap, float *rec_zx, int rec_cap);

/* ---- measurement ------------------------------------------------------ */
typedef struct daw_stats {
    double sweep_ms[3];           /* time in forward / back-propagation / re-injection sweeps (timing on) */
    long long sweep_steps[3];     /* time steps in those sweeps */
    double batch_ms;              /* device time of daw_redatum_batch calls */
    long long point_updates;      /* sum over executed sweeps of n1e*n2e per shot per step */
    long long step_launches;
    int max_batch_used;
    double record_ms[3];          /* of sweep_ms: the batched record launches (daw_set_record_batch(1)) */
} daw_stats;

int daw_set_timing(int on);
/* 1: capture each full sweep of a batch (state reset + nt step launches) as a
 * hipGraph once and replay it for the later sweeps of the same kind, e.g. the
 * 2 x itn Hessian sweeps of the MINRES solve; 0 (default): launch step by
 * step.  Results are identical.  No reference counterpart. */
int daw_set_graph(int on);
/* Step-kernel tile (rows x columns): 8x128, 16x64, 16x128, 20x64, 24x64, 32x64, 32x128
 * or 64x64, or R x 58 (R = 8, 12, 16, 24, 32): the column march, one wavefront
 * per 58 owned columns walking R rows (forward sweeps, and the Hessian sweeps
 * with FP32 snapshots; with 16-bit snapshots those take 16x64), or R x 52
 * (R = 16, 24, 32, 40, 48): the two-step column march, two time steps per
 * launch (re-injection sweeps with FP32 snapshots and batched records; the
 * other sweeps take 16x64, and an odd last step the R x 58 march, R <= 32);
 * rows = 0 lets daw_redatum_batch time every shape on a short probe sweep per
 * batch shape and sweep mode and keep the fastest (the default).  Results are
 * identical.  No reference counterpart: the launch shape of the autotuner. */
int daw_set_tile(int rows, int cols);
/* 1 (default): the Hicks records of a sweep (data and datum gathers) run as
 * batched launches every few steps, each workgroup staging the band of up to
 * 64 neighbouring points in LDS; 0: as record blocks inside the step
 * launches.  Results are identical.  No reference counterpart. */
int daw_set_record_batch(int on);
int daw_get_tile(int *rows, int *cols);
/* Time steps advanced per launch of the sweep kernels: 1 (one launch per
 * step), 2 (the two-step kernel: the wave state crosses HBM once per two
 * steps; tiles 16x64, 32x64, 16x128) or 0 (default) to let the
 * probe of daw_redatum_batch pick it with the tile.  Results are identical.
 * No reference counterpart. */
int daw_set_steps(int steps);
int daw_get_steps(void);
/* Launch configuration the autotuner chose for sweep mode 0 (forward),
 * 1 (back-propagation) or 2 (re-injection) in the last batch. */
int daw_get_launch(int mode, int *steps, int *rows, int *cols);
int daw_reset_stats(void);
int daw_get_stats(daw_stats *st);

#ifdef __cplusplus
}
#endif
#endif
