// gfx950 kernels of the DAW redatuming path.
//
// All arithmetic that the reference performs in single precision is done
// here in single precision with FP contraction disabled for the whole file
// (-ffp-contract=off), operation for operation in the reference's order, so
// the propagators, the Hicks gathers and the resampling produce bit-identical
// results to the reference's CPU path.  MINRES scalars and dot products are
// FP64.
#include "daw_device.h"
#include "daw_kernels.h"

namespace daw {

namespace {

__device__ __forceinline__ bool in_box(int z, int x, int lo, int hiz, int hix)
{
    return z >= lo && z <= hiz && x >= lo && x <= hix;
}

// ---------------------------------------------------------------------------
// Trailing record blocks: sis(it, r) = sum over the Hicks window of c * p,
// i1 outer / i2 inner as submodeling2.f90:214-219 (extended-grid clip) and
// subwavefieldresample.f90:32-37 (interior clip) accumulate it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void record_item(const StepParams &P, int k)
{
    const RecItem R = P.rec[k];
    const ShotDev &S = P.shots[R.shot];
    if (P.rec_it >= S.nt) return;
    int lo1, hi1, lo2, hi2;
    if (R.clip == 0) {
        lo1 = -R.ix1; hi1 = S.n1e - 1 - R.ix1;
        lo2 = -R.ix2; hi2 = S.n2e - 1 - R.ix2;
    } else {
        lo1 = S.npml - R.ix1; hi1 = S.npml + S.n1 - 1 - R.ix1;
        lo2 = S.npml - R.ix2; hi2 = S.npml + S.n2 - 1 - R.ix2;
    }
    lo1 = lo1 > -10 ? lo1 : -10; hi1 = hi1 < 10 ? hi1 : 10;
    lo2 = lo2 > -10 ? lo2 : -10; hi2 = hi2 < 10 ? hi2 : 10;
    const float *band = S.pband[P.rec_slot];
    float acc = 0.0f;
    for (int i1 = lo1; i1 <= hi1; ++i1) {
        const float *row = band + (size_t)S.band_row[R.ix1 + i1] * S.ldx + R.ix2;
        const float *c = R.coef + (i1 + 10);
        for (int i2 = lo2; i2 <= hi2; ++i2) acc = acc + c[21 * (i2 + 10)] * row[i2];
    }
    R.out[R.reverse ? (S.nt - 1 - P.rec_it) : P.rec_it] = acc;
}

}  // namespace

// ---------------------------------------------------------------------------
// One time step of one sweep for every active shot of a batch.
//
// Per owned tile (TZ x TX points):
//   1. stage old vx (x halo 3/3) and vz (z halo 3/3) in LDS;
//   2. over the tile plus a one/two-cell pressure ring: split-PML pressure
//      update, Hicks injection gather (kFwd source, kBack residuals) or the
//      time-reversed snapshot re-injection (kReinject), p = px + pz, free
//      surface; p of the ring stays in LDS, owned px/pz/p go to HBM;
//   3. velocity update of the owned points from LDS p, vz source or
//      snapshot re-injection, vz(npml) = vz(npml+1) mirror; write vx/vz
//      (and the kBack velocity snapshots).
// Blocks past ntiles run the Hicks record of the previous step.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_step(StepParams P)
{
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    if (b >= P.ntiles) {
        const int k = (b - P.ntiles) * kBlock + tid;
        if (k < P.nrec) record_item(P, k);
        return;
    }
    __shared__ float s_vx[VXH * VXW];
    __shared__ float s_vz[VZH * VZW];
    __shared__ float s_p[PH * PW];

    const int2 t = P.tiles[b];
    const ShotDev &S = P.shots[t.x];
    const int it = P.it;
    if (it >= S.nt) return;
    const int n1e = S.n1e, n2e = S.n2e, ldx = S.ldx, npml = S.npml;
    const int ntx = (n2e + TX - 1) / TX;
    const int z0 = (t.y / ntx) * TZ;
    const int x0 = (t.y % ntx) * TX;
    // stencil bounds: 3..n-2 forward (submodeling2.f90:148-149), 3..n-3 for
    // both passes of the Hessian-type sweeps (:1014-1015, :1108-1109)
    const int lo = 2;
    const int hiz = (MODE == kFwd) ? n1e - 3 : n1e - 4;
    const int hix = (MODE == kFwd) ? n2e - 3 : n2e - 4;
    const int cur = P.cur, nxt = P.cur ^ 1;
    const float *__restrict__ vxo = S.vx[cur];
    const float *__restrict__ vzo = S.vz[cur];
    const float *__restrict__ pxo = S.px[cur];
    const float *__restrict__ pzo = S.pz[cur];
    float *__restrict__ pxn = S.px[nxt];
    float *__restrict__ pzn = S.pz[nxt];
    float *__restrict__ vxn = S.vx[nxt];
    float *__restrict__ vzn = S.vz[nxt];
    const float a0 = 1.125f;
    const float a1 = -(1.0f / 24.0f);

    for (int i = tid; i < VXH * VXW; i += kBlock) {
        const int r = i / VXW, c = i - r * VXW;
        const int gz = z0 - 1 + r, gx = x0 - 3 + c;
        s_vx[i] = (gz >= 0 && gz < n1e && gx >= 0 && gx < n2e) ? vxo[(size_t)gz * ldx + gx] : 0.0f;
    }
    for (int i = tid; i < VZH * VZW; i += kBlock) {
        const int r = i / VZW, c = i - r * VZW;
        const int gz = z0 - 3 + r, gx = x0 - 1 + c;
        s_vz[i] = (gz >= 0 && gz < n1e && gx >= 0 && gx < n2e) ? vzo[(size_t)gz * ldx + gx] : 0.0f;
    }
    __syncthreads();

    const int n1 = S.n1, n2 = S.n2;
    const size_t plane = (size_t)n1 * n2;
    const int ic = S.nt - 1 - it;   // time-reversed snapshot index (icount = nt-it+1)

    for (int i = tid; i < PH * PW; i += kBlock) {
        const int r = i / PW, c = i - r * PW;
        const int gz = z0 - 1 + r, gx = x0 - 1 + c;
        float p = 0.0f;
        if (gz >= 0 && gz < n1e && gx >= 0 && gx < n2e) {
            const size_t g = (size_t)gz * ldx + gx;
            float pxv = pxo[g], pzv = pzo[g];
            const float kap = S.kappa[g];
            if (in_box(gz, gx, lo, hiz, hix)) {
                const float *vr = &s_vx[r * VXW + c + 2];          // vx(gz, gx)
                const float *vc = &s_vz[(r + 2) * VZW + c];        // vz(gz, gx)
                const float der2 = a0 * (vr[0] - vr[-1]) + a1 * (vr[1] - vr[-2]);
                const float der1 = a0 * (vc[0] - vc[-VZW]) + a1 * (vc[VZW] - vc[-2 * VZW]);
                const float tk = S.dtdx * kap;                      // kappaw = dtdx*kappa*damp1 (:105-106)
                const float kw2 = tk * S.d2a1[gx];
                const float kw1 = tk * S.d1a1[gz];
                if (MODE == kBack) {
                    pxv = S.d2a[gx] * pxv - kw2 * der2;
                    pzv = S.d1a[gz] * pzv - kw1 * der1;
                } else {
                    pxv = S.d2a[gx] * pxv + kw2 * der2;
                    pzv = S.d1a[gz] * pzv + kw1 * der1;
                }
            }
            if (MODE != kReinject && S.inj_row != nullptr && !S.vinj) {
                const int q = S.inj_row[gz];
                if (q >= 0) {
                    const int *ptr = S.inj_ptr + (size_t)q * (n2e + 1);
                    const int kb = ptr[gx], ke = ptr[gx + 1];
                    for (int k = kb; k < ke; ++k) {
                        const float term = (S.inj_cdt[k] * S.inj_amp[(size_t)S.inj_r[k] * S.amp_stride + it]) * kap;
                        if (MODE == kBack) pxv = pxv - term;   // :1042
                        else pxv = pxv + term;                 // :175
                    }
                }
            }
            if (MODE == kReinject) {
                if (gz >= S.xlo1 && gz <= S.xhi1 && gx >= S.xlo2 && gx <= S.xhi2) {
                    const float sv = S.sp[(size_t)ic * plane + (size_t)(gz - npml) * n2 + (gx - npml)];
                    pxv = pxv + ((-S.dt) * sv) * kap;          // ss*dt*sp*kappa, ss = -1 (:1123)
                }
            }
            p = pxv + pzv;
            if (S.ofs == 1 && gz <= npml) p = 0.0f;           // p(1:npml+1,:) = 0
            if (r >= 1 && r <= TZ && c >= 1 && c <= TX) {
                pxn[g] = pxv;
                pzn[g] = pzv;
                const bool interior = gz >= npml && gz < npml + n1 && gx >= npml && gx < npml + n2;
                if (MODE == kBack && interior)
                    S.sp[(size_t)it * plane + (size_t)(gz - npml) * n2 + (gx - npml)] = p;
                if (MODE != kBack) {
                    if (S.band_row != nullptr) {
                        const int slot = S.band_row[gz];
                        if (slot >= 0) S.pband[it & 1][(size_t)slot * ldx + gx] = p;
                    }
                    if (S.wf != nullptr && interior)
                        S.wf[(size_t)it * plane + (size_t)(gx - npml) * n1 + (gz - npml)] = p;
                }
            }
        }
        s_p[i] = p;
    }
    __syncthreads();

    for (int i = tid; i < TZ * TX; i += kBlock) {
        const int r = i / TX, c = i - r * TX;
        const int gz = z0 + r, gx = x0 + c;
        if (gz >= n1e || gx >= n2e) continue;
        const size_t g = (size_t)gz * ldx + gx;
        float vxv = s_vx[(r + 1) * VXW + c + 3];
        float vzv = s_vz[(r + 3) * VZW + c + 1];
        if (in_box(gz, gx, lo, hiz, hix)) {
            const float *pp = &s_p[(r + 1) * PW + c + 1];         // p(gz, gx)
            const float der2 = a0 * (pp[1] - pp[0]) + a1 * (pp[2] - pp[-1]);
            const float der1 = a0 * (pp[PW] - pp[0]) + a1 * (pp[2 * PW] - pp[-PW]);
            if (MODE == kBack) {
                vxv = S.d2b[gx] * vxv - S.bw2[gx] * der2;
                vzv = S.d1b[gz] * vzv - S.bw1[gz] * der1;
            } else {
                vxv = S.d2b[gx] * vxv + S.bw2[gx] * der2;
                vzv = S.d1b[gz] * vzv + S.bw1[gz] * der1;
            }
        }
        if (MODE == kFwd && S.vinj && S.inj_row != nullptr) {
            const int q = S.inj_row[gz];
            if (q >= 0) {
                const int *ptr = S.inj_ptr + (size_t)q * (n2e + 1);
                for (int k = ptr[gx]; k < ptr[gx + 1]; ++k)
                    vzv = vzv + S.inj_cdt[k] * S.inj_amp[(size_t)S.inj_r[k] * S.amp_stride + it];   // * bue (= 1), :252
            }
        }
        const bool interior = gz >= npml && gz < npml + n1 && gx >= npml && gx < npml + n2;
        if (MODE == kReinject && interior) {
            const size_t s = (size_t)ic * plane + (size_t)(gz - npml) * n2 + (gx - npml);
            vxv = vxv + (-S.dt) * S.svx[s];                        // * bue (= 1), :1177
            vzv = vzv + (-S.dt) * S.svz[s];                        // :1184
        }
        vxn[g] = vxv;
        if (S.ofs == 1 && gz == npml - 1) {
            // vz(npml,:) is overwritten by the mirror below; its owner writes nothing
        } else {
            vzn[g] = vzv;
            if (S.ofs == 1 && gz == npml) vzn[g - ldx] = vzv;     // vz(npml,:) = vz(npml+1,:)
        }
        if (MODE == kBack && interior) {
            const size_t s = (size_t)it * plane + (size_t)(gz - npml) * n2 + (gx - npml);
            S.svx[s] = vxv;
            S.svz[s] = vzv;
        }
    }
}

template __global__ void k_step<kFwd>(StepParams);
template __global__ void k_step<kBack>(StepParams);
template __global__ void k_step<kReinject>(StepParams);

void launch_step(int mode, const StepParams &P, hipStream_t st)
{
    const int grid = P.ntiles + P.nrec_blocks;
    if (grid == 0) return;
    if (mode == kFwd) hipLaunchKernelGGL(k_step<kFwd>, dim3(grid), dim3(kBlock), 0, st, P);
    else if (mode == kBack) hipLaunchKernelGGL(k_step<kBack>, dim3(grid), dim3(kBlock), 0, st, P);
    else hipLaunchKernelGGL(k_step<kReinject>, dim3(grid), dim3(kBlock), 0, st, P);
}

// ---------------------------------------------------------------------------
// Small elementwise kernels
// ---------------------------------------------------------------------------
__global__ void k_fill(float *p, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void launch_fill(float *p, size_t n, float v, hipStream_t st)
{
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
}

// res(nt-it+1, r) = dobs(it, r) - dcal(it, r)   (subresiduals.f90:12-16); traces [r][t]
__global__ void k_residuals(const float *dobs, const float *dcal, float *res, int nt, size_t ntr)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ntr * nt; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / nt;
        const int t = (int)(i - r * nt);
        res[r * nt + (nt - 1 - t)] = dobs[i] - dcal[i];
    }
}

void launch_residuals(const float *dobs, const float *dcal, float *res, int nt, size_t ntr, hipStream_t st)
{
    size_t n = ntr * nt;
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_residuals, dim3((unsigned)blocks), dim3(256), 0, st, dobs, dcal, res, nt, ntr);
}

__global__ void k_add(const float *a, const float *b, float *c, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c[i] = a[i] + b[i];
}

void launch_add(const float *a, const float *b, float *c, size_t n, hipStream_t st)
{
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_add, dim3((unsigned)blocks), dim3(256), 0, st, a, b, c, n);
}

// subwavefieldresample (subwavefieldresample.f90:18-39) from a stored
// interior wavefield in the reference's (n1,n2,nt) layout; one thread per
// (receiver, step), window clipped to the interior.
__global__ void k_resample(const float *wf, int n1, int n2, int nt, const int *ix1, const int *ix2,
                           const float *coef, int nr, float *sis)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nr * nt) return;
    const int r = (int)(i / nt), t = (int)(i - (size_t)r * nt);
    const int c1 = ix1[r], c2 = ix2[r];       // 0-based interior centre
    int lo1 = -c1 > -10 ? -c1 : -10, hi1 = n1 - 1 - c1 < 10 ? n1 - 1 - c1 : 10;
    int lo2 = -c2 > -10 ? -c2 : -10, hi2 = n2 - 1 - c2 < 10 ? n2 - 1 - c2 : 10;
    const float *w = wf + (size_t)t * n1 * n2;
    const float *c = coef + (size_t)r * 441;
    float acc = 0.0f;
    for (int i1 = lo1; i1 <= hi1; ++i1)
        for (int i2 = lo2; i2 <= hi2; ++i2)
            acc = acc + c[(i1 + 10) + 21 * (i2 + 10)] * w[(size_t)(c2 + i2) * n1 + (c1 + i1)];
    sis[(size_t)r * nt + t] = acc;
}

void launch_resample(const float *wf, int n1, int n2, int nt, const int *ix1, const int *ix2, const float *coef,
                     int nr, float *sis, hipStream_t st)
{
    const size_t n = (size_t)nr * nt;
    if (n == 0) return;
    hipLaunchKernelGGL(k_resample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, wf, n1, n2, nt, ix1, ix2,
                       coef, nr, sis);
}

// ---------------------------------------------------------------------------
// MINRES (minresModule.f90:10-360) over a batch of shots, vectors FP32,
// dot products and the Lanczos / Givens scalars FP64 on the device.
// Each shot owns vectors of length n_s = nt*nr laid out [r][t] exactly as
// the reference's 1-D vector (sub2d21d, sublinearminres.f90:138-151).
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-level deterministic FP64 sum (wavefront shuffle tree, then LDS).
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double s_w[kBlock / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s_w[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) r += s_w[i];
    __syncthreads();
    return r;
}

// Launch shape of the MINRES vector kernels: grid (kVecBlocks, nshots);
// block (b, s) handles elements b*256 + tid + k*kVecBlocks*256 of shot s.
constexpr int kVecBlocks = 64;

// s.scal layout, one row of kScal doubles per shot
enum {
    kBeta = 0, kOldb, kBeta1, kDbar, kEpsln, kQrnorm, kPhibar, kRhs1, kRhs2, kTnorm2, kYnorm2, kCs, kSn,
    kGmax, kGmin, kAnorm, kYnorm, kRnorm, kAcond, kAlfa, kShift, kOldeps, kDelta, kPhi, kDenom, kIstop,
    kItn, kArnorm, kFac, kStopNext, kScal
};

__global__ void k_dot(const float *a, const float *b, const size_t *off, const int *len, double *part)
{
    const int s = blockIdx.y;
    const float *x = a + off[s], *y = b + off[s];
    const size_t n = len[s];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock)
        acc += (double)x[i] * (double)y[i];
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[(size_t)s * kVecBlocks + blockIdx.x] = acc;
}

void launch_dot(const float *a, const float *b, const size_t *off, const int *len, double *part, int nshots,
                hipStream_t st)
{
    hipLaunchKernelGGL(k_dot, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, a, b, off, len, part);
}

__device__ __forceinline__ double reduce_parts(const double *part, int s)
{
    double r = 0.0;
    for (int i = 0; i < kVecBlocks; ++i) r += part[(size_t)s * kVecBlocks + i];
    return r;
}

// fac = ||deltadr|| / ||H deltadr|| * mu, shift = -fac (sublinearminres.f90:64-72);
// then MINRES initialisation with y = r1 = b (:83-153).  part0 = b.b,
// part1 = (H b).(H b).
__global__ void k_minres_init(const double *part0, const double *part1, double *scal, float mu, int nshots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double *q = scal + (size_t)s * kScal;
    const double bb = reduce_parts(part0, s), hh = reduce_parts(part1, s);
    const double fac = (hh > 0.0) ? sqrt(bb) / sqrt(hh) * (double)mu : 0.0;
    for (int i = 0; i < kScal; ++i) q[i] = 0.0;
    q[kFac] = fac;
    q[kShift] = -fac;
    if (bb < 0.0) { q[kIstop] = 8; return; }
    if (bb == 0.0) { q[kIstop] = 0; q[kItn] = -1; return; }   // x = 0, nothing to iterate
    const double beta1 = sqrt(bb);
    q[kBeta1] = beta1;
    q[kBeta] = beta1;
    q[kQrnorm] = beta1;
    q[kPhibar] = beta1;
    q[kRhs1] = beta1;
    q[kCs] = -1.0;
}

// active(s): the shot still iterates
__device__ __forceinline__ bool active(const double *q) { return q[kIstop] == 0.0 && q[kItn] >= 0.0; }

// v = y / beta (:181-182)
__global__ void k_minres_v(const float *y, float *v, const size_t *off, const int *len, const double *scal)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    if (!active(q)) return;
    const double sc = 1.0 / q[kBeta];
    const size_t n = len[s];
    const float *ys = y + off[s];
    float *vs = v + off[s];
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock)
        vs[i] = (float)(sc * (double)ys[i]);
}

// y = Hv - shift*v - (beta/oldb)*r1 (itn >= 2); partial v.y   (:185-190)
__global__ void k_minres_l1(float *y, const float *v, const float *r1, const size_t *off, const int *len,
                            const double *scal, double *part)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    double acc = 0.0;
    if (active(q)) {
        const double shift = q[kShift];
        const double c = (q[kItn] >= 1.0) ? q[kBeta] / q[kOldb] : 0.0;   // kItn counts finished iterations
        const size_t n = len[s];
        float *ys = y + off[s];
        const float *vs = v + off[s], *rs = r1 + off[s];
        for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock) {
            const double vi = vs[i];
            const float yi = (float)((double)ys[i] - shift * vi - c * (double)rs[i]);
            ys[i] = yi;
            acc += vi * (double)yi;
        }
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[(size_t)s * kVecBlocks + blockIdx.x] = acc;
}

__global__ void k_minres_alfa(const double *part, double *scal, int nshots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double *q = scal + (size_t)s * kScal;
    if (!active(q)) return;
    q[kAlfa] = reduce_parts(part, s);
}

// y = y - (alfa/beta)*r2; partial y.y  (:191-197)
__global__ void k_minres_l2(float *y, const float *r2, const size_t *off, const int *len, const double *scal,
                            double *part)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    double acc = 0.0;
    if (active(q)) {
        const double c = q[kAlfa] / q[kBeta];
        const size_t n = len[s];
        float *ys = y + off[s];
        const float *rs = r2 + off[s];
        for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock) {
            const float yi = (float)((double)ys[i] - c * (double)rs[i]);
            ys[i] = yi;
            acc += (double)yi * (double)yi;
        }
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[(size_t)s * kVecBlocks + blockIdx.x] = acc;
}

// Scalar part of one MINRES iteration (:196-303) for every shot.
__global__ void k_minres_scalars(const double *part, double *scal, double *hist, int itnlim, double rtol,
                                 int nshots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double *q = scal + (size_t)s * kScal;
    if (!active(q)) return;
    const double eps = 1.1920928955078125e-07;   // epsilon of the reference's REAL
    const int itn = (int)q[kItn] + 1;
    const double alfa = q[kAlfa];
    const double oldb = q[kBeta];
    double beta = reduce_parts(part, s);
    q[kOldb] = oldb;
    if (beta < 0.0) { q[kIstop] = 6; q[kItn] = itn; return; }
    beta = sqrt(beta);
    q[kTnorm2] = q[kTnorm2] + alfa * alfa + oldb * oldb + beta * beta;
    int istop = 0;
    if (itn == 1) {
        if (beta / q[kBeta1] <= 10.0 * eps) istop = -1;
        q[kGmax] = fabs(alfa);
        q[kGmin] = q[kGmax];
    }
    const double cs = q[kCs], sn = q[kSn], dbar = q[kDbar];
    const double oldeps = q[kEpsln];
    const double delta = cs * dbar + sn * alfa;
    const double gbar = sn * dbar - cs * alfa;
    const double epsln = sn * beta;
    const double dbar2 = -cs * beta;
    const double gamma = sqrt(gbar * gbar + beta * beta);
    const double cs2 = gbar / gamma, sn2 = beta / gamma;
    const double phi = cs2 * q[kPhibar];
    const double phibar = sn2 * q[kPhibar];
    q[kOldeps] = oldeps;
    q[kDelta] = delta;
    q[kEpsln] = epsln;
    q[kDbar] = dbar2;
    q[kCs] = cs2;
    q[kSn] = sn2;
    q[kPhi] = phi;
    q[kPhibar] = phibar;
    q[kDenom] = 1.0 / gamma;
    q[kBeta] = beta;
    q[kGmax] = fmax(q[kGmax], gamma);
    q[kGmin] = fmin(q[kGmin], gamma);
    const double z = q[kRhs1] / gamma;
    q[kYnorm2] = z * z + q[kYnorm2];
    q[kRhs1] = q[kRhs2] - delta * z;
    q[kRhs2] = -epsln * z;
    q[kAnorm] = sqrt(q[kTnorm2]);
    q[kYnorm] = sqrt(q[kYnorm2]);
    const double rnorml = q[kRnorm];
    q[kQrnorm] = phibar;
    q[kRnorm] = phibar;
    q[kArnorm] = rnorml * sqrt(gbar * gbar + dbar2 * dbar2);
    q[kAcond] = q[kGmax] / q[kGmin];
    if (hist) hist[(size_t)s * itnlim + (itn - 1)] = phibar / q[kBeta1];
    if (istop == 0) {
        if (itn >= itnlim) istop = 5;
        if (phibar / q[kBeta1] <= rtol) istop = 1;
    }
    q[kItn] = itn;
    // the x update of this iteration still runs (:250-257 precede the exit);
    // k_minres_close applies the stop afterwards
    q[kStopNext] = istop;
}

// w1 <- w2, w2 <- w, w <- (v - oldeps*w1 - delta*w2) * denom, x += phi*w (:252-257).
// Buffer roles rotate on the host; here w_new is written over w1_old.
__global__ void k_minres_x(float *x, float *w_new_over_w1, const float *w2_is_old_w, const float *w1_is_old_w2,
                           const float *v, const size_t *off, const int *len, double *scal)
{
    const int s = blockIdx.y;
    double *q = scal + (size_t)s * kScal;
    if (!(q[kIstop] == 0.0 && q[kItn] >= 1.0)) return;
    const double oldeps = q[kOldeps], delta = q[kDelta], denom = q[kDenom], phi = q[kPhi];
    const size_t n = len[s];
    float *xs = x + off[s], *wn = w_new_over_w1 + off[s];
    const float *w2 = w2_is_old_w + off[s], *w1 = w1_is_old_w2 + off[s], *vs = v + off[s];
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock) {
        const float wi = (float)(((double)vs[i] - oldeps * (double)w1[i] - delta * (double)w2[i]) * denom);
        wn[i] = wi;
        xs[i] = (float)((double)xs[i] + phi * (double)wi);
    }
}

// Close the iteration: apply the parked stop flag.
__global__ void k_minres_close(double *scal, int nshots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double *q = scal + (size_t)s * kScal;
    if (q[kIstop] != 0.0 || q[kItn] < 0.0) return;
    q[kIstop] = q[kStopNext];
}

void launch_minres_init(const double *part0, const double *part1, double *scal, float mu, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres_init, dim3((nshots + 63) / 64), dim3(64), 0, st, part0, part1, scal, mu, nshots);
}

void launch_minres_v(const float *y, float *v, const size_t *off, const int *len, const double *scal, int nshots,
                     hipStream_t st)
{
    hipLaunchKernelGGL(k_minres_v, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, y, v, off, len, scal);
}

void launch_minres_l1(float *y, const float *v, const float *r1, const size_t *off, const int *len,
                      const double *scal, double *part, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres_l1, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, y, v, r1, off, len, scal, part);
    hipLaunchKernelGGL(k_minres_alfa, dim3((nshots + 63) / 64), dim3(64), 0, st, part, const_cast<double *>(scal),
                       nshots);
}

void launch_minres_l2(float *y, const float *r2, const size_t *off, const int *len, const double *scal,
                      double *part, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres_l2, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, y, r2, off, len, scal, part);
}

void launch_minres_scalars(const double *part, double *scal, double *hist, int itnlim, double rtol, int nshots,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_minres_scalars, dim3((nshots + 63) / 64), dim3(64), 0, st, part, scal, hist, itnlim, rtol,
                       nshots);
}

void launch_minres_x(float *x, float *wn, const float *w2, const float *w1, const float *v, const size_t *off,
                     const int *len, double *scal, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres_x, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, x, wn, w2, w1, v, off, len, scal);
    hipLaunchKernelGGL(k_minres_close, dim3((nshots + 63) / 64), dim3(64), 0, st, scal, nshots);
}

int minres_scal_stride() { return kScal; }
int minres_vec_blocks() { return kVecBlocks; }

}  // namespace daw
