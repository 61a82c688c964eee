// Launch wrappers of the gfx950 kernels (daw_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include "daw_device.h"

namespace daw {

// snap: SnapFormat of the snapshot store written by kBack and read by kReinject.
void launch_step(int mode, int snap, const StepParams &P, int tz, int txw, hipStream_t st);
// Tile shape (rows tz, 64*txw columns) has compiled kernels.
bool tile_shape_ok(int tz, int txw);
// Two steps (it, it+1) per launch: tile of tz rows and 64*nc columns.
void launch_step2(int mode, int snap, const StepParams &P, int tz, int nc, hipStream_t st);
bool tile2_shape_ok(int tz, int nc);
// Column march (one step per launch): wavefronts of 58 owned columns walking
// h rows each; FP32 snapshots only.  march_rows_ok: h has a compiled kernel.
void launch_march(int mode, const StepParams &P, int h, hipStream_t st);
bool march_rows_ok(int h);
constexpr int kMarchCols = 58;
// Rows of the column march's short segments (rows with many Hicks injection terms).
constexpr int kMarchBandRows = 2;
// Two-step column march of the re-injection (time steps it and it+1 in one
// launch): wavefronts of 52 owned columns walking the rows of each tile
// entry's segment; FP32 snapshots.  A shot with it+1 == nt takes step it alone.
void launch_march2(const StepParams &P, hipStream_t st);
constexpr int kMarch2Cols = 52;
// Hicks record of the items of ngroups record groups at steps it0 ..
// it0+nsteps-1 (nsteps <= kBandSlots) from the band's time slots, after
// those steps ran.
void launch_rec_batch(const ShotDev *shots, const RecItem *items, const RecGroup *groups, int ngroups, int it0,
                      int nsteps, hipStream_t st);
// out[t][r] = in[r][t], r < nr, t < nt
void launch_transpose(const float *in, float *out, int nr, int nt, hipStream_t st);
// Exact adjoint of one step (reverse time); hess selects the transpose of
// the re-injection pass (snapshots out) over that of the forward sweep.
void launch_adjoint(bool hess, const StepParams &P, int tz, int txw, hipStream_t st);
void launch_fill(float *p, size_t n, float v, hipStream_t st);
void launch_residuals(const float *dobs, const float *dcal, float *res, int nt, size_t ntr, hipStream_t st);
void launch_add(const float *a, const float *b, float *c, size_t n, hipStream_t st);
void launch_resample(const float *wf, int n1, int n2, int nt, const int *ix1, const int *ix2, const float *coef,
                     int nr, float *sis, hipStream_t st);

// Band-pass (filmartin2d) of ntr device traces of nt samples with FFT length
// nto <= 8192; wgt = dmartin(1..nto/2+1), tw[m] = exp(-2 pi i m / nto).
int launch_filter(const float *din, float *dout, int ntr, int nt, int nto, const double *wgt, const double2 *tw,
                  hipStream_t st);
// The same for FFT lengths nto = 16384, 32768, 65536: a four-step FFT in
// three passes through work (pairs_per_group * nto complex values), trace
// pairs pairs_per_group at a time; tw[m] = exp(-2 pi i m / nto), m < nto.
// Returns -1 for other lengths.  din and dout may be the same traces.
int launch_filter_long(const float *din, float *dout, int ntr, int nt, int nto, const double *wgt, const double2 *tw,
                       double2 *work, int pairs_per_group, hipStream_t st);
// nto = n1 * n2 of the long band-pass, or -1 when nto is not 16384..65536
int filter_long_split(int nto, int &n1, int &n2);

// MINRES over a batch: vectors of shot s start at off[s] and hold len[s]
// floats; part holds minres_vec_blocks() FP64 partial sums per shot; scal
// holds minres_scal_stride() FP64 scalars per shot.
int minres_scal_stride();
int minres_vec_blocks();
void launch_dot(const float *a, const float *b, const size_t *off, const int *len, double *part, int nshots,
                hipStream_t st);
void launch_minres_init(const double *part0, const double *part1, double *scal, float mu, int nshots,
                        hipStream_t st);
void launch_minres_v(const float *y, float *v, const size_t *off, const int *len, const double *scal, int nshots,
                     hipStream_t st);
void launch_minres_l1(float *y, const float *v, const float *r1, const size_t *off, const int *len,
                      const double *scal, double *part, int nshots, hipStream_t st);
void launch_minres_l2(float *y, const float *r2, const size_t *off, const int *len, const double *scal,
                      double *part, int nshots, hipStream_t st);
// subresample of ntr traces [ntr][n1in] -> [ntr][n1out]; plan holds (it0,
// ialpha) per output sample (2 ints each), coef the 9x10 table
void launch_time_resample(const float *xin, float *xout, int n1in, int n1out, size_t ntr, const int *plan,
                          const float *coef, hipStream_t st);

// trace (may be null): kTraceStride doubles per shot and iteration,
// [shot][itn-1] = x(1), qrnorm/beta1, Anorm, Acond, printed (1/0) -- the
// columns of the MINRES log line of that iteration and whether it is written
void launch_minres_scalars(const double *part, double *scal, double *hist, double *trace, const int *len, int itnlim,
                           double rtol, int nshots, hipStream_t st);
void launch_minres_x(float *x, float *wn, const float *w2, const float *w1, const float *v, const size_t *off,
                     const int *len, double *scal, double *trace, int itnlim, int nshots, hipStream_t st);
constexpr int kTraceStride = 5;

// MINRES in the reference's FP32 arithmetic.  launch_seqdot32 writes one
// sequential FP32 sum per shot into its scalar row: which = 0, 1, 2 (the
// initialisation's datanorm(b)^2, datanorm(Hb)^2, b.b) or 3 (alfa);
// order 0 sums a[i]*b[i] for ascending i (dot_product), order 1 in datanorm's
// t-outer, r-inner order over [r][t] traces of nts[s] samples.
void launch_seqdot32(const float *a, const float *b, const size_t *off, const int *len, const int *nts, double *scal,
                     int which, int order, bool skip_inactive, int nshots, hipStream_t st);
void launch_minres32_init(double *scal, float mu, int nshots, hipStream_t st);
void launch_minres32_v(const float *y, float *v, const size_t *off, const int *len, const double *scal, int nshots,
                       hipStream_t st);
void launch_minres32_l1(float *y, const float *v, const float *r1, const size_t *off, const int *len,
                        const double *scal, int nshots, hipStream_t st);
void launch_minres32_l2(float *y, const float *r2, const size_t *off, const int *len, const double *scal,
                        int nshots, hipStream_t st);
// r2.y of the iteration is read from slot which = 0
void launch_minres32_scalars(double *scal, double *hist, double *trace, const int *len, int itnlim, float rtol,
                             int nshots, hipStream_t st);
void launch_minres32_x(float *x, float *wn, const float *w2, const float *w1, const float *v, const size_t *off,
                       const int *len, double *scal, double *trace, int itnlim, int nshots, hipStream_t st);

// MINRES of the FP64 path with r1, r2, w, w1, w2 stored in 16 bits (fmt =
// kSnapBF16 or kSnapF16): element i of shot s at q[off[s] + i], the scale of
// its 256-element chunk at sc[coff[s] + i/256].  The scalar kernel
// (launch_minres_scalars) is shared with the FP32-vector path.
void launch_mk_init(int fmt, const float *b, unsigned short *qa, float *sa, unsigned short *qb, float *sb,
                    const size_t *off, const size_t *coff, const int *len, int nshots, hipStream_t st);
void launch_mk_zero(unsigned short *q0, float *s0, unsigned short *q1, float *s1, unsigned short *q2, float *s2,
                    size_t nq, size_t nsc, hipStream_t st);
void launch_mk_v(int fmt, const unsigned short *q, const float *sc, float *v, const size_t *off, const size_t *coff,
                 const int *len, const double *scal, int nshots, hipStream_t st);
void launch_mk_l1(int fmt, float *y, const float *v, const unsigned short *q1, const float *s1, const size_t *off,
                  const size_t *coff, const int *len, const double *scal, double *part, int nshots, hipStream_t st);
void launch_mk_l2(int fmt, float *y, const unsigned short *q2, const float *s2, unsigned short *qo, float *so,
                  const size_t *off, const size_t *coff, const int *len, const double *scal, double *part, int nshots,
                  hipStream_t st);
void launch_mk_x(int fmt, float *x, unsigned short *qn, float *sn, const unsigned short *q2, const float *s2,
                 const unsigned short *q1, const float *s1, const float *v, const size_t *off, const size_t *coff,
                 const int *len, double *scal, double *trace, int itnlim, int nshots, hipStream_t st);

// indices into a shot's scalar row (see k_minres_scalars)
enum MinresScalIndex {
    kSBeta = 0, kSOldb, kSBeta1, kSDbar, kSEpsln, kSQrnorm, kSPhibar, kSRhs1, kSRhs2, kSTnorm2, kSYnorm2, kSCs,
    kSSn, kSGmax, kSGmin, kSAnorm, kSYnorm, kSRnorm, kSAcond, kSAlfa, kSShift, kSOldeps, kSDelta, kSPhi,
    kSDenom, kSIstop, kSItn, kSArnorm, kSFac
};

}  // namespace daw
