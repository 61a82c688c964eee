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

// Pointers the step kernels dereference are global memory: typed so, the
// compiler emits global (not flat) loads and stores with 32-bit offsets, and
// LDS waits do not wait on them.
#define GLB __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GLB T *glb(T *p)
{
    return (GLB T *)p;
}

// Element idx of a global array addressed as a uniform base plus a 32-bit
// byte offset (the scalar-base form of the global load and store).
template <class T>
__device__ __forceinline__ T ld(const GLB T *base, unsigned idx)
{
    return *(const GLB T *)((const GLB char *)base + idx * (unsigned)sizeof(T));
}

template <class T>
__device__ __forceinline__ void st(GLB T *base, unsigned idx, T v)
{
    *(GLB T *)((GLB char *)base + idx * (unsigned)sizeof(T)) = v;
}

// A field of a shot's state or profile block: one uniform base (px[0] of the
// shot's 8 state planes, or the model's profile block) and the 32-bit element
// offset of the field in it, so the kernels keep one base pointer per block.
struct Fld {
    GLB float *base;
    unsigned off;
    __device__ __forceinline__ float operator[](int i) const { return ld(base, off + (unsigned)i); }
};
__device__ __forceinline__ float ld(const Fld &f, unsigned idx) { return ld(f.base, f.off + idx); }
__device__ __forceinline__ void st(const Fld &f, unsigned idx, float v) { st(f.base, f.off + idx, v); }

// Non-temporal store, for the snapshot store: a snapshot is written once per
// back-propagation and read thousands of steps later, so its writes should
// not displace the wave state (read again at the next step) from the caches.
template <class T>
__device__ __forceinline__ void st_nt(GLB T *base, unsigned idx, T v)
{
    __builtin_nontemporal_store(v, (GLB T *)((GLB char *)base + idx * (unsigned)sizeof(T)));
}

__device__ __forceinline__ int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ bool in_box(int z, int x, int lo, int hiz, int hix)
{
    return z >= lo && z <= hiz && x >= lo && x <= hix;
}

// ---------------------------------------------------------------------------
// Record blocks: the Hicks record of (shot, receiver) items at step rec_it,
// kRecPerBlock items per workgroup.  sis(it, r) = sum over the 21x21 window
// of c(i1,i2) * p(i1,i2), one sequential chain i1 outer / i2 inner exactly as
// submodeling2.f90:214-219 (window clipped to the extended grid) and
// subwavefieldresample.f90:32-37 (clipped to the interior) accumulate it.
// The workgroup first forms every product of its items (all loads in flight
// at once, product k of an item at (i1+10)*21 + (i2+10), coefficients stored
// in that order) in LDS; then one thread per item runs the chain over its
// products in k order, so the additions happen in the reference's order.  A
// clipped-away entry contributes +0, which leaves the sum unchanged: a sum
// that starts at +0 is never -0 under round-to-nearest.
// ---------------------------------------------------------------------------
struct RecSlot {
    const GLB float *band;
    const GLB int *brow;
    const GLB float *coef;
    GLB float *out;
    int ix1, ix2, lo1, hi1, lo2, hi2, ldx, live;
};

__device__ __forceinline__ void record_block(const StepParams &P, int blk, float *s_lds)
{
    __shared__ RecSlot s_slot[kRecPerBlock];
    const int tid = threadIdx.x;
    const int first = blk * kRecPerBlock;
    if (tid < kRecPerBlock) {
        RecSlot q;
        q.live = 0;
        const int q_all = first + tid;            // (step, item) of the launch's record work
        const int sub = q_all / max(P.nrec, 1), item = q_all - sub * P.nrec;
        const int rit = P.rec_it + sub;
        if (sub < P.rec_nsub) {
            const RecItem R = P.rec[item];
            const ShotDev *__restrict__ S = P.shots + R.shot;
            const int nt = S->nt;
            if (rit < nt) {
                q.live = 1;
                if (R.clip == 0) {
                    q.lo1 = -R.ix1; q.hi1 = S->n1e - 1 - R.ix1;
                    q.lo2 = -R.ix2; q.hi2 = S->n2e - 1 - R.ix2;
                } else {
                    q.lo1 = S->npml - R.ix1; q.hi1 = S->npml + S->n1 - 1 - R.ix1;
                    q.lo2 = S->npml - R.ix2; q.hi2 = S->npml + S->n2 - 1 - R.ix2;
                }
                q.band = glb(S->pband + (size_t)(rit & (kBandSlots - 1)) * S->band_stride);
                q.brow = glb(S->band_row);
                q.coef = glb(R.coef);
                q.out = glb(R.out) + (R.reverse ? (nt - 1 - rit) : rit);
                q.ix1 = R.ix1;
                q.ix2 = R.ix2;
                q.ldx = S->ldx;
            }
        }
        s_slot[tid] = q;
    }
    __syncthreads();
    constexpr int NPROD = kRecPerBlock * kRecStride;
    constexpr int NJ = (NPROD + kBlock - 1) / kBlock;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int f = tid + kBlock * j;
        if (f < NPROD) {
            const int it = f / kRecStride, k = f - kRecStride * it;
            const int i1 = k / 21 - 10, i2 = k - 21 * (k / 21) - 10;
            const RecSlot &q = s_slot[it];
            float v = 0.0f;
            if (q.live && k < 441 && i1 >= q.lo1 && i1 <= q.hi1 && i2 >= q.lo2 && i2 <= q.hi2)
                v = q.coef[k] * q.band[q.brow[q.ix1 + i1] * q.ldx + (q.ix2 + i2)];
            s_lds[f] = v;
        }
    }
    __syncthreads();
    if (tid < kRecPerBlock && s_slot[tid].live) {
        const float4 *pr = reinterpret_cast<const float4 *>(s_lds + tid * kRecStride);
        float acc = 0.0f;
#pragma unroll 8
        for (int m = 0; m < kRecStride / 4; ++m) {
            const float4 w = pr[m];
            acc = acc + w.x;
            acc = acc + w.y;
            acc = acc + w.z;
            acc = acc + w.w;
        }
        *s_slot[tid].out = acc;
    }
}

// Pressure half-step at one point of the pressure rectangle (the reference's
// statement order: split-PML update on the stencil box, Hicks injection,
// snapshot re-injection, p = px + pz, free surface).  vr points at vx(gz,gx)
// in LDS, vc at vz(gz,gx).  Returns p; writes px/pz when `owned`.
// Flat per-launch view of one shot: every field a scalar or a pointer, the
// double-buffer halves already selected, so it lives in registers.
struct View {
    int n1, n2, npml, n1e, n2e, ldx, nt, ofs, par;
    float dt, dtdx;
    const GLB float *kappa, *d1a, *d1a1, *d1b, *bw1, *d2a, *d2a1, *d2b, *bw2;
    const GLB float *pxo, *pzo, *vxo, *vzo;
    GLB float *pxn, *pzn, *vxn, *vzn;
    GLB float *sp, *svx, *svz;
    GLB unsigned short *sp16, *svx16, *svz16;
    GLB float *ssc;
    int nseg, spitch, sxb;
    const ShotDev *__restrict__ sd;   // the rest (injection, record band, ranges), read where used
    // the same fields as base + offset (one base per block)
    Fld fpxo, fpzo, fvxo, fvzo, fpxn, fpzn, fvxn, fvzn;
    Fld fd1a, fd1a1, fd1b, fbw1, fd2a, fd2a1, fd2b, fbw2;
};

__device__ __forceinline__ View make_view(const ShotDev *__restrict__ S, int cur, int it)
{
    View v;
    v.n1 = S->n1; v.n2 = S->n2; v.npml = S->npml; v.n1e = S->n1e; v.n2e = S->n2e; v.ldx = S->ldx;
    v.nt = S->nt; v.ofs = S->ofs; v.par = it & (kBandSlots - 1); v.sd = S;
    v.dt = S->dt; v.dtdx = S->dtdx;
    v.kappa = glb(S->kappa); v.d1a = glb(S->d1a); v.d1a1 = glb(S->d1a1); v.d1b = glb(S->d1b); v.bw1 = glb(S->bw1);
    v.d2a = glb(S->d2a); v.d2a1 = glb(S->d2a1); v.d2b = glb(S->d2b); v.bw2 = glb(S->bw2);
    const bool c1 = cur != 0;
    v.pxo = glb(c1 ? S->px[1] : S->px[0]); v.pzo = glb(c1 ? S->pz[1] : S->pz[0]);
    v.vxo = glb(c1 ? S->vx[1] : S->vx[0]); v.vzo = glb(c1 ? S->vz[1] : S->vz[0]);
    v.pxn = glb(c1 ? S->px[0] : S->px[1]); v.pzn = glb(c1 ? S->pz[0] : S->pz[1]);
    v.vxn = glb(c1 ? S->vx[0] : S->vx[1]); v.vzn = glb(c1 ? S->vz[0] : S->vz[1]);
    v.sp = glb(S->sp); v.svx = glb(S->svx); v.svz = glb(S->svz);
    v.sp16 = glb(S->sp16); v.svx16 = glb(S->svx16); v.svz16 = glb(S->svz16); v.ssc = glb(S->ssc); v.nseg = S->nseg;
    v.spitch = S->spitch; v.sxb = S->sxb;
    {
        GLB float *st0 = glb(S->px[0]);
        const unsigned pe = (unsigned)(S->n1e * S->ldx), ro = c1 ? 4 * pe : 0u, wo = c1 ? 0u : 4 * pe;
        v.fpxo = {st0, ro}; v.fpzo = {st0, ro + pe}; v.fvxo = {st0, ro + 2 * pe}; v.fvzo = {st0, ro + 3 * pe};
        v.fpxn = {st0, wo}; v.fpzn = {st0, wo + pe}; v.fvxn = {st0, wo + 2 * pe}; v.fvzn = {st0, wo + 3 * pe};
        GLB float *pr = (GLB float *)glb(S->d1a);
        const unsigned a = (unsigned)S->n1e, b = (unsigned)S->n2e;
        v.fd1a = {pr, 0}; v.fd1a1 = {pr, a}; v.fd1b = {pr, 2 * a}; v.fbw1 = {pr, 3 * a};
        v.fd2a = {pr, 4 * a}; v.fd2a1 = {pr, 4 * a + b}; v.fd2b = {pr, 4 * a + 2 * b}; v.fbw2 = {pr, 4 * a + 3 * b};
    }
    return v;
}


// ---- 16-bit snapshot codec.  Each 64-column row segment of a snapshot plane
// is written by one wavefront, which finds its max |value| with a DPP
// reduction and stores the values divided by the power of two at or above
// it (an exact rescaling) as BF16 or FP16, round to nearest even.
//
// Max of |v| over the wavefront: non-negative floats order like their bit
// patterns, so the integer maximum is taken (NaN counts as 0, as fmaxf
// ignores it).  Prefix maxima within each 16-lane row (row_shr 1, 2, 4, 8),
// then row 0 into row 1 and row 2 into row 3 (row_bcast:15), then row 1 into
// rows 2-3 (row_bcast:31): lane 63 holds the maximum.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_max(int x)
{
    const int y = __builtin_amdgcn_update_dpp(0, x, CTRL, ROW_MASK, 0xf, false);
    return x > y ? x : y;
}

__device__ __forceinline__ float wave_max_abs(float v)
{
    int x = __float_as_int(fabsf(v));
    x = x > 0x7f800000 ? 0 : x;
    x = dpp_max<0x111, 0xf>(x);   // row_shr:1
    x = dpp_max<0x112, 0xf>(x);   // row_shr:2
    x = dpp_max<0x114, 0xf>(x);   // row_shr:4
    x = dpp_max<0x118, 0xf>(x);   // row_shr:8
    x = dpp_max<0x142, 0xa>(x);   // row_bcast:15
    x = dpp_max<0x143, 0xc>(x);   // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

__device__ __forceinline__ float pow2_scale(float amax)
{
    if (!(amax > 0.0f)) return 1.0f;
    int e;
    (void)frexpf(amax, &e);          // amax = m * 2^e, m in [0.5, 1)
    return ldexpf(1.0f, e < -100 ? -100 : e);   // > amax; 1/scale stays finite
}

template <int F>
__device__ __forceinline__ unsigned short snap_encode(float v)
{
    if (F == kSnapBF16) {
        unsigned int u = __float_as_uint(v);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (unsigned short)(u >> 16);
    } else {
        return __builtin_bit_cast(unsigned short, (_Float16)v);   // round to nearest even
    }
}

template <int F>
__device__ __forceinline__ float snap_decode(unsigned short q)
{
    if (F == kSnapBF16) return __uint_as_float((unsigned int)q << 16);
    return (float)__builtin_bit_cast(_Float16, q);
}

// Interior snapshot element: plane t of an (n1*n2)-point snapshot array
// (a uniform 64-bit base) plus a 32-bit offset; segment scales likewise.
template <class T>
__device__ __forceinline__ T *plane_at(T *base, const View &S, int t)
{
    return base + (size_t)t * ((size_t)S.n1 * S.n2);
}

// snapshot plane t (rows of spitch elements) and the element (gz, gx) in it
template <class T>
__device__ __forceinline__ T *splane_at(T *base, const View &S, int t)
{
    return base + (size_t)t * ((size_t)S.n1 * S.spitch);
}

__device__ __forceinline__ unsigned snap_off(const View &S, int gz, int gx)
{
    return (unsigned)((gz - S.npml) * S.spitch + (gx - S.sxb));
}

__device__ __forceinline__ GLB float *scale_at(const View &S, int t, int field)
{
    return S.ssc + ((size_t)t * 3 + field) * ((size_t)S.n1 * S.nseg);
}

__device__ __forceinline__ unsigned scale_off(const View &S, int gz, int gx)
{
    return (unsigned)((gz - S.npml) * S.nseg + (gx >> 6));
}

template <int F>
__device__ __forceinline__ float snap_load(const View &S, const GLB float *f32, const GLB unsigned short *f16, int field,
                                           int t, int gz, int gx)
{
    if (F == kSnapF32) return ld(splane_at(f32, S, t), snap_off(S, gz, gx));
    return snap_decode<F>(ld(splane_at(f16, S, t), snap_off(S, gz, gx))) * ld(scale_at(S, t, field), scale_off(S, gz, gx));
}

// The same for a whole wavefront on one snapshot row gz (uniform) of its
// 64-column segment wseg (uniform): the scale is one scalar load.
template <int F>
__device__ __forceinline__ float snap_load_w(const View &S, const GLB float *f32, const GLB unsigned short *f16, int field,
                                             int t, int gz, int gx, int wseg)
{
    if (F == kSnapF32) return ld(splane_at(f32, S, t), snap_off(S, gz, gx));
    const unsigned so = (unsigned)__builtin_amdgcn_readfirstlane((gz - S.npml) * S.nseg + wseg);
    return snap_decode<F>(ld(splane_at(f16, S, t), snap_off(S, gz, gx))) * ld(scale_at(S, t, field), so);
}

// Whole wavefront (all 64 lanes, converged, one 64-aligned column segment):
// store v where `interior`.  A wavefront of a tile that reaches past the grid
// (segment >= nseg) has no interior lane and writes no scale.
template <int F>
__device__ __forceinline__ void snap_store_wave(const View &S, GLB float *f32, GLB unsigned short *f16, int field, int t,
                                                int gz, int gx, float v, bool interior, bool row_interior)
{
    if (F == kSnapF32) {
        if (interior) st_nt(splane_at(f32, S, t), snap_off(S, gz, gx), v);
        return;
    }
    const float sc = pow2_scale(wave_max_abs(interior ? v : 0.0f));
    if (interior) st_nt(splane_at(f16, S, t), snap_off(S, gz, gx), snap_encode<F>(v * (1.0f / sc)));
    if (row_interior && (threadIdx.x & 63) == 0 && (gx >> 6) < S.nseg) st(scale_at(S, t, field), scale_off(S, gz, gx), sc);
}

// Hicks injection into one point (the reference's receiver loop, one term at a
// time in receiver order): entries kb .. ke-1 of the row's CSR, term =
// (c*dt * amp(r, it)) * kappa added (SIGN > 0) or subtracted.  Entries go four
// at a time, all their loads issued before the first dependent one, so a row
// costs two load round trips per four entries instead of two per entry; the
// additions stay one rounded operation each, in order.
template <int SIGN>
__device__ __forceinline__ float inject_terms(float acc, float kap, int kb, int ke, const GLB int *rid,
                                              const GLB float *cdt, const GLB float *amp, int stride, int toff)
{
    int q = kb;
    for (; q + 4 <= ke; q += 4) {
        const int r0 = rid[q], r1 = rid[q + 1], r2 = rid[q + 2], r3 = rid[q + 3];
        const float c0 = cdt[q], c1 = cdt[q + 1], c2 = cdt[q + 2], c3 = cdt[q + 3];
        const float a0 = amp[r0 * stride + toff], a1 = amp[r1 * stride + toff];
        const float a2 = amp[r2 * stride + toff], a3 = amp[r3 * stride + toff];
        const float t0 = (c0 * a0) * kap, t1 = (c1 * a1) * kap, t2 = (c2 * a2) * kap, t3 = (c3 * a3) * kap;
        if (SIGN > 0) {
            acc = acc + t0; acc = acc + t1; acc = acc + t2; acc = acc + t3;
        } else {
            acc = acc - t0; acc = acc - t1; acc = acc - t2; acc = acc - t3;
        }
    }
    for (; q < ke; ++q) {
        const float t = (cdt[q] * amp[rid[q] * stride + toff]) * kap;
        acc = SIGN > 0 ? acc + t : acc - t;
    }
    return acc;
}

// Pressure half-step of one point of the general tile body: the split update
// inside the stencil box (submodeling2.f90:152-153 forward, :1018-1019
// back-propagation, :1112-1113 re-injection), the Hicks injection terms of the
// point's CSR entries (injq = the row's slot, -1 for none), the re-injection's
// extended source, p = px + pz (:183, :1051, :1127) with the free-surface zero,
// and for owned points the stores of px, pz, the record band slot and the
// wavefield.
template <int MODE, int VZW_>
__device__ __forceinline__ float p_point(const View &S, int it, int ic, int gz, int gx, float pxv, float pzv,
                                         float kap, const float *vr, const float *vc, float d1a, float d1a1,
                                         float d2a, float d2a1, int injq, bool owned, int lo, int hiz, int hix,
                                         float spv, const Fld &pxn, const Fld &pzn)
{
    const float a0 = 1.125f;
    const float a1 = -(1.0f / 24.0f);
    if (in_box(gz, gx, lo, hiz, hix)) {
        const float der2 = a0 * (vr[0] - vr[-1]) + a1 * (vr[1] - vr[-2]);
        const float der1 = a0 * (vc[0] - vc[-VZW_]) + a1 * (vc[VZW_] - vc[-2 * VZW_]);
        const float tk = S.dtdx * kap;                      // kappaw = dtdx*kappa*damp1 (:105-106)
        const float kw2 = tk * d2a1;
        const float kw1 = tk * d1a1;
        if (MODE == kBack) {
            pxv = d2a * pxv - kw2 * der2;
            pzv = d1a * pzv - kw1 * der1;
        } else {
            pxv = d2a * pxv + kw2 * der2;
            pzv = d1a * pzv + kw1 * der1;
        }
    }
    if (MODE != kReinject && injq >= 0) {                // :1042 (subtracted), :175 (added)
        const GLB int *ptr = glb(S.sd->inj_ptr) + injq * (S.n2e + 1);
        pxv = inject_terms<MODE == kBack ? -1 : 1>(pxv, kap, ptr[gx], ptr[gx + 1], glb(S.sd->inj_r), glb(S.sd->inj_cdt),
                                                  glb(S.sd->inj_amp), S.sd->amp_stride, it * S.sd->amp_tstride);
    }
    if (MODE == kReinject) {
        if (gz >= S.sd->xlo1 && gz <= S.sd->xhi1 && gx >= S.sd->xlo2 && gx <= S.sd->xhi2)
            pxv = pxv + ((-S.dt) * spv) * kap;         // ss*dt*sp*kappa, ss = -1 (:1123)
    }
    float p = pxv + pzv;
    if (S.ofs == 1 && gz <= S.npml) p = 0.0f;           // p(1:npml+1,:) = 0
    if (owned) {
        const unsigned g = (unsigned)(gz * S.ldx + gx);
        st(pxn, g, pxv);
        st(pzn, g, pzv);
        const int npml = S.npml;
        const bool interior = gz >= npml && gz < npml + S.n1 && gx >= npml && gx < npml + S.n2;
        if (MODE != kBack) {
            if (S.sd->band_row != nullptr) {
                const int slot = glb(S.sd->band_row)[gz];
                if (slot >= 0) glb(S.sd->pband + (size_t)S.par * S.sd->band_stride)[slot * S.ldx + gx] = p;
            }
            if (S.sd->wf != nullptr && interior)
                plane_at(glb(S.sd->wf), S, it)[(gx - npml) * S.n1 + (gz - npml)] = p;
        }
    }
    return p;
}

// Interior tile of k_step: every point of the pressure rectangle (tile plus
// ring) lies where all four sponge profiles are exactly 1 (extended rows and
// columns npml .. npml+n-1), inside the stencil box, below the free surface,
// off the injection rows and inside the re-injection range.  There the
// reference's update multiplies by damp = damp1 = 1.0f and bw = dtdx*1*1 =
// dtdx, which are exact, so dropping those factors gives the same bits, and
// no load needs a clamp: the tile runs without per-point branches.  Layout,
// LDS use and statement order are those of the general body in k_step.
template <int MODE, int TZ, int TXW, int SNAP>
__device__ __forceinline__ void tile_interior(const View &S, float *s_vx, float *s_vz, float *s_p, int z0, int x0, int it,
                                              int ic, int wv, int lane, int wseg)
{
    constexpr int NW = kBlock / 64;
    constexpr int TXT = 64 * TXW;
    constexpr int NWZ = NW / TXW;
    constexpr int PW_ = TXT + 3, PH_ = TZ + 3;
    constexpr int VXW_ = TXT + 6, VXH_ = TZ + 3;
    constexpr int VZW_ = TXT + 3, VZH_ = TZ + 6;
    constexpr int PROWS = (PH_ + NWZ - 1) / NWZ;
    constexpr int VROWS = TZ / NWZ;
    constexpr int XR = (VXH_ + NWZ - 1) / NWZ, ZR = (VZH_ + NWZ - 1) / NWZ;
    constexpr int HXN = (VXH_ * 6 + kBlock - 1) / kBlock, HZN = (VZH_ * 3 + kBlock - 1) / kBlock;
    const float a0 = 1.125f;
    const float a1 = -(1.0f / 24.0f);
    const int tid = threadIdx.x, ldx = S.ldx;
    const int cl = (wv % TXW) * 64 + lane, rz = wv / TXW, gx = x0 + cl;
    const float dtdx = S.dtdx, mdt = -S.dt;

    // ---- 1. loads; a wavefront row past the end of a rectangle repeats the
    // last row (the same value lands twice in LDS, and that row is never owned)
    float vxl[XR], vzl[ZR];
#pragma unroll
    for (int j = 0; j < XR; ++j) {
        const int r = min(rz + NWZ * j, VXH_ - 1);
        vxl[j] = ld(S.fvxo, (unsigned)((z0 - 1 + r) * ldx + gx));
    }
#pragma unroll
    for (int j = 0; j < ZR; ++j) {
        const int r = min(rz + NWZ * j, VZH_ - 1);
        vzl[j] = ld(S.fvzo, (unsigned)((z0 - 3 + r) * ldx + gx));
    }
    float hx[HXN], hz[HZN];
    int hx_i[HXN], hz_i[HZN];
#pragma unroll
    for (int q = 0; q < HXN; ++q) {
        const int i = tid + q * kBlock;
        hx[q] = 0.0f;
        hx_i[q] = -1;
        if (i < VXH_ * 6) {
            const int r = i / 6, k = i - 6 * r;
            const int c = k < 3 ? k : TXT + k;
            hx_i[q] = r * VXW_ + c;
            hx[q] = ld(S.fvxo, (unsigned)((z0 - 1 + r) * ldx + x0 - 3 + c));
        }
    }
#pragma unroll
    for (int q = 0; q < HZN; ++q) {
        const int i = (kBlock - 1 - tid) + q * kBlock;
        hz[q] = 0.0f;
        hz_i[q] = -1;
        if (i < VZH_ * 3) {
            const int r = i / 3, k = i - 3 * r;
            const int c = k == 0 ? 0 : TXT + k;
            hz_i[q] = r * VZW_ + c;
            hz[q] = ld(S.fvzo, (unsigned)((z0 - 3 + r) * ldx + x0 - 1 + c));
        }
    }
    float pxr[PROWS], pzr[PROWS], kar[PROWS], spr[PROWS];
#pragma unroll
    for (int j = 0; j < PROWS; ++j) {
        const int r = min(rz + NWZ * j, PH_ - 1);
        const int gz = z0 - 1 + r;
        const unsigned g = (unsigned)(gz * ldx + gx);
        pxr[j] = ld(S.fpxo, g);
        pzr[j] = ld(S.fpzo, g);
        kar[j] = ld(S.kappa, g);
        spr[j] = MODE == kReinject ? snap_load_w<SNAP>(S, S.sp, S.sp16, 0, ic, gz, gx, wseg) : 0.0f;
    }
    const int sid = kBlock - 1 - tid;
    const bool side = sid < 3 * TZ;
    int s_r = 0, s_c = 0;
    float s_px = 0.0f, s_pz = 0.0f, s_ka = 0.0f, s_sp = 0.0f;
    if (side) {
        const int k = sid / TZ;
        s_r = 1 + (sid - k * TZ);
        s_c = k == 0 ? 0 : TXT + k;
        const int s_gz = z0 - 1 + s_r, s_gx = x0 - 1 + s_c;
        const unsigned g = (unsigned)(s_gz * ldx + s_gx);
        s_px = ld(S.fpxo, g);
        s_pz = ld(S.fpzo, g);
        s_ka = ld(S.kappa, g);
        if (MODE == kReinject) s_sp = snap_load<SNAP>(S, S.sp, S.sp16, 0, ic, s_gz, s_gx);
    }
    float svxr[VROWS], svzr[VROWS];
#pragma unroll
    for (int j = 0; j < VROWS; ++j) {
        const int gz = z0 + rz + NWZ * j;
        svxr[j] = MODE == kReinject ? snap_load_w<SNAP>(S, S.svx, S.svx16, 1, ic, gz, gx, wseg) : 0.0f;
        svzr[j] = MODE == kReinject ? snap_load_w<SNAP>(S, S.svz, S.svz16, 2, ic, gz, gx, wseg) : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < XR; ++j) s_vx[min(rz + NWZ * j, VXH_ - 1) * VXW_ + cl + 3] = vxl[j];
#pragma unroll
    for (int j = 0; j < ZR; ++j) s_vz[min(rz + NWZ * j, VZH_ - 1) * VZW_ + cl + 1] = vzl[j];
#pragma unroll
    for (int q = 0; q < HXN; ++q)
        if (hx_i[q] >= 0) s_vx[hx_i[q]] = hx[q];
#pragma unroll
    for (int q = 0; q < HZN; ++q)
        if (hz_i[q] >= 0) s_vz[hz_i[q]] = hz[q];
    __syncthreads();

    // ---- 2. pressure half-step: split update, extended source, p = px + pz
#pragma unroll
    for (int j = 0; j < PROWS; ++j) {
        const int r = min(rz + NWZ * j, PH_ - 1);
        const int gz = z0 - 1 + r;
        const float *vr = &s_vx[r * VXW_ + cl + 3];
        const float *vc = &s_vz[(r + 2) * VZW_ + cl + 1];
        const float der2 = a0 * (vr[0] - vr[-1]) + a1 * (vr[1] - vr[-2]);
        const float der1 = a0 * (vc[0] - vc[-VZW_]) + a1 * (vc[VZW_] - vc[-2 * VZW_]);
        const float tk = dtdx * kar[j];
        float pxv = MODE == kBack ? pxr[j] - tk * der2 : pxr[j] + tk * der2;
        const float pzv = MODE == kBack ? pzr[j] - tk * der1 : pzr[j] + tk * der1;
        if (MODE == kReinject) pxv = pxv + (mdt * spr[j]) * kar[j];
        const float p = pxv + pzv;
        if (r >= 1 && r <= TZ) {
            const unsigned g = (unsigned)(gz * ldx + gx);
            st(S.fpxn, g, pxv);
            st(S.fpzn, g, pzv);
            if (MODE != kBack && S.sd->band_row != nullptr) {
                const int slot = glb(S.sd->band_row)[gz];
                if (slot >= 0) glb(S.sd->pband + (size_t)S.par * S.sd->band_stride)[slot * ldx + gx] = p;
            }
            if (MODE == kBack) snap_store_wave<SNAP>(S, S.sp, S.sp16, 0, it, gz, gx, p, true, true);
        }
        s_p[r * PW_ + cl + 1] = p;
    }
    if (side) {
        const float *vr = &s_vx[s_r * VXW_ + s_c + 2];
        const float *vc = &s_vz[(s_r + 2) * VZW_ + s_c];
        const float der2 = a0 * (vr[0] - vr[-1]) + a1 * (vr[1] - vr[-2]);
        const float der1 = a0 * (vc[0] - vc[-VZW_]) + a1 * (vc[VZW_] - vc[-2 * VZW_]);
        const float tk = dtdx * s_ka;
        float pxv = MODE == kBack ? s_px - tk * der2 : s_px + tk * der2;
        const float pzv = MODE == kBack ? s_pz - tk * der1 : s_pz + tk * der1;
        if (MODE == kReinject) pxv = pxv + (mdt * s_sp) * s_ka;
        s_p[s_r * PW_ + s_c] = pxv + pzv;
    }
    __syncthreads();

    // ---- 3. velocity half-step, snapshot re-injection or store
#pragma unroll
    for (int j = 0; j < VROWS; ++j) {
        const int r = rz + NWZ * j;
        const int gz = z0 + r;
        const unsigned g = (unsigned)(gz * ldx + gx);
        float vxv = s_vx[(r + 1) * VXW_ + cl + 3];
        float vzv = s_vz[(r + 3) * VZW_ + cl + 1];
        const float *pp = &s_p[(r + 1) * PW_ + cl + 1];
        const float der2 = a0 * (pp[1] - pp[0]) + a1 * (pp[2] - pp[-1]);
        const float der1 = a0 * (pp[PW_] - pp[0]) + a1 * (pp[2 * PW_] - pp[-PW_]);
        vxv = MODE == kBack ? vxv - dtdx * der2 : vxv + dtdx * der2;
        vzv = MODE == kBack ? vzv - dtdx * der1 : vzv + dtdx * der1;
        if (MODE == kReinject) {
            vxv = vxv + mdt * svxr[j];
            vzv = vzv + mdt * svzr[j];
        }
        st(S.fvxn, g, vxv);
        st(S.fvzn, g, vzv);
        if (MODE == kBack) {
            snap_store_wave<SNAP>(S, S.svx, S.svx16, 1, it, gz, gx, vxv, true, true);
            snap_store_wave<SNAP>(S, S.svz, S.svz16, 2, it, gz, gx, vzv, true, true);
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// One time step of one sweep for every active shot of a batch.
//
// Block layout: the first nrec_blocks workgroups run the Hicks record of the
// previous step (so its latency-bound chains overlap the stencil work); the
// rest each own a TZ x 64 tile.  Wavefront w of a tile block owns rows
// w, w+4, ... with lane = column, so every global access is one aligned
// 256-byte row segment and per-row / per-column PML profiles stay in
// registers.  Per tile:
//   1. stage old vx (x halo 3/3) and vz (z halo 3/3) in LDS and load
//      px, pz, kappa of the pressure rectangle into registers;
//   2. pressure half-step on the tile plus the one/two-cell ring the
//      velocity stencil needs (the ring's p stays in LDS);
//   3. velocity half-step of the owned points, vz source or snapshot
//      re-injection, the vz(npml) = vz(npml+1) mirror, writes.
// ---------------------------------------------------------------------------
// Waves per SIMD the compiler must leave room for: 16x64 and 20x64 tiles fit
// eight (78 SGPRs, at most 64 VGPRs, no spills; 20x64 still fits eight
// workgroups' LDS per CU) where the default register budget gives seven (106
// SGPRs); measured at 201x801 with 16 shots per launch: forward 10 % faster,
// re-injection 1.5 %, back-propagation on 20x64 1.5 % faster than the best
// seven-wave tile.  The other shapes stay at the default: their LDS caps them
// at seven or fewer, or the cap would spill (re-injection on 20x64, and on
// 16x64 with 16-bit snapshots).
template <int MODE, int TZ, int TXW, int SNAP>
constexpr int step_min_waves()
{
    return TXW == 1 && ((TZ == 16 && (MODE != kReinject || SNAP == kSnapF32)) || (TZ == 20 && MODE != kReinject))
               ? 8 : 1;
}

template <int MODE, int TZ, int TXW, int SNAP>
__global__ __launch_bounds__(kBlock, (step_min_waves<MODE, TZ, TXW, SNAP>())) void k_step(StepParams P)
{
    constexpr int NW = kBlock / 64;
    constexpr int TXT = 64 * TXW;          // tile width: TXW wavefronts side by side
    constexpr int NWZ = NW / TXW;          // wavefront rows
    constexpr int PW_ = TXT + 3, PH_ = TZ + 3;
    constexpr int VXW_ = TXT + 6, VXH_ = TZ + 3;
    constexpr int VZW_ = TXT + 3, VZH_ = TZ + 6;
    constexpr int PROWS = (PH_ + NWZ - 1) / NWZ;
    constexpr int VROWS = TZ / NWZ;
    static_assert(TZ % NWZ == 0 && 3 * TZ <= kBlock, "tile rows must split evenly over the wavefront rows");
    // the wavefront index is uniform: in an SGPR, row-level conditions and loads are scalar
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int LDS_N = VXH_ * VXW_ + VZH_ * VZW_ + PH_ * PW_;
    static_assert(LDS_N >= kRecPerBlock * kRecStride, "record blocks reuse the tile's LDS");
    __shared__ __align__(16) float s_lds[LDS_N];
    if ((int)blockIdx.x < P.nrec_blocks) {
        record_block(P, blockIdx.x, s_lds);
        return;
    }
    float *const s_vx = s_lds;
    float *const s_vz = s_lds + VXH_ * VXW_;
    float *const s_p = s_vz + VZH_ * VZW_;

    const int2 t = P.tiles[blockIdx.x - P.nrec_blocks];
    const int it = P.it;
    const View S = make_view(P.shots + t.x, P.cur, it);   // registers: stores below cannot alias it
    if (it >= S.nt) return;
    const int n1e = S.n1e, n2e = S.n2e, ldx = S.ldx;
    const int ntx = (n2e + TXT - 1) / TXT;
    const int z0 = (t.y / ntx) * TZ;
    const int x0 = (t.y % ntx) * TXT;
    // stencil bounds: 3..n-2 forward (submodeling2.f90:148-149), 3..n-3 for
    // both passes of the Hessian-type sweeps (:1014-1015, :1108-1109)
    const int lo = 2;
    const int hiz = (MODE == kFwd) ? n1e - 3 : n1e - 4;
    const int hix = (MODE == kFwd) ? n2e - 3 : n2e - 4;
    const Fld vxo = S.fvxo, vzo = S.fvzo, pxo = S.fpxo, pzo = S.fpzo;
    const Fld pxn = S.fpxn, pzn = S.fpzn, vxn = S.fvxn, vzn = S.fvzn;
    const int ic = S.nt - 1 - it;   // time-reversed snapshot index (icount = nt-it+1)
    const int cl = (wv % TXW) * 64 + lane;   // this lane's column in the tile
    const int rz = wv / TXW;                  // this wavefront's first row
    const int gxl = x0 + cl;
    const bool colok = gxl < n2e;
    // this wavefront's 64-column snapshot segment (lanes past the grid read
    // the last one and discard the value)
    const int wseg = __builtin_amdgcn_readfirstlane(min((x0 >> 6) + (wv % TXW), S.nseg - 1));
    {   // interior tiles take the branch-free body (the same bits, see tile_interior)
        const int pzl = z0 - 1, pzh = z0 + TZ + 1, pxl = x0 - 1, pxh = x0 + TXT + 1;
        bool fast = pzl >= S.npml && pzh < S.npml + S.n1 && pxl >= S.npml && pxh < S.npml + S.n2 && pzl >= lo &&
                    pzh <= hiz && pxl >= lo && pxh <= hix && !(S.ofs == 1 && pzl <= S.npml);
        if (MODE != kReinject && S.sd->inj_row != nullptr && S.sd->inj_z0 <= pzh && S.sd->inj_z1 >= pzl) fast = false;
        if (MODE != kBack && S.sd->wf != nullptr) fast = false;
        if (MODE == kReinject &&
            !(S.sd->xlo1 <= pzl && pzh <= S.sd->xhi1 && S.sd->xlo2 <= pxl && pxh <= S.sd->xhi2))
            fast = false;
        if (fast) {
            tile_interior<MODE, TZ, TXW, SNAP>(S, s_vx, s_vz, s_p, z0, x0, it, ic, wv, lane, wseg);
            return;
        }
    }

    // ---- 1. every global load of the step is issued here, into registers,
    // before anything waits on one: velocity halos, the pressure rectangle,
    // and (re-injection) the snapshot values
    constexpr int XR = (VXH_ + NWZ - 1) / NWZ, ZR = (VZH_ + NWZ - 1) / NWZ;
    const int npml = S.npml, n1 = S.n1, n2 = S.n2;
    // loads outside the grid read a clamped in-grid address and are zeroed
    // after: no branch around any load
    const int gxc = colok ? gxl : n2e - 1;
    float vxl[XR], vzl[ZR];
#pragma unroll
    for (int j = 0; j < XR; ++j) {
        const int r = rz + NWZ * j, gz = z0 - 1 + r;
        const float v = ld(vxo, (unsigned)(clampi(gz, 0, n1e - 1) * ldx + gxc));
        vxl[j] = (r < VXH_ && gz >= 0 && gz < n1e && colok) ? v : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < ZR; ++j) {
        const int r = rz + NWZ * j, gz = z0 - 3 + r;
        const float v = ld(vzo, (unsigned)(clampi(gz, 0, n1e - 1) * ldx + gxc));
        vzl[j] = (r < VZH_ && gz >= 0 && gz < n1e && colok) ? v : 0.0f;
    }
    constexpr int HXN = (VXH_ * 6 + kBlock - 1) / kBlock, HZN = (VZH_ * 3 + kBlock - 1) / kBlock;
    float hx[HXN], hz[HZN];
    int hx_i[HXN], hz_i[HZN];
#pragma unroll
    for (int q = 0; q < HXN; ++q) {
        const int i = tid + q * kBlock;
        hx[q] = 0.0f;
        hx_i[q] = -1;
        if (i < VXH_ * 6) {
            const int r = i / 6, k = i - 6 * r;
            const int c = k < 3 ? k : TXT + k;
            const int gz = z0 - 1 + r, gx = x0 - 3 + c;
            hx_i[q] = r * VXW_ + c;
            if (gz >= 0 && gz < n1e && gx >= 0 && gx < n2e) hx[q] = ld(vxo, (unsigned)(gz * ldx + gx));
        }
    }
#pragma unroll
    for (int q = 0; q < HZN; ++q) {
        const int i = (kBlock - 1 - tid) + q * kBlock;
        hz[q] = 0.0f;
        hz_i[q] = -1;
        if (i < VZH_ * 3) {
            const int r = i / 3, k = i - 3 * r;
            const int c = k == 0 ? 0 : TXT + k;
            const int gz = z0 - 3 + r, gx = x0 - 1 + c;
            hz_i[q] = r * VZW_ + c;
            if (gz >= 0 && gz < n1e && gx >= 0 && gx < n2e) hz[q] = ld(vzo, (unsigned)(gz * ldx + gx));
        }
    }
    float pxr[PROWS], pzr[PROWS], kar[PROWS], spr[PROWS];
#pragma unroll
    for (int j = 0; j < PROWS; ++j) {
        const int r = rz + NWZ * j;
        const int gz = z0 - 1 + r;
        const bool ok = r < PH_ && gz >= 0 && gz < n1e && colok;
        const unsigned g = (unsigned)(clampi(gz, 0, n1e - 1) * ldx + gxc);
        const float a = ld(pxo, g), b = ld(pzo, g), c = ld(S.kappa, g);
        pxr[j] = ok ? a : 0.0f;
        pzr[j] = ok ? b : 0.0f;
        kar[j] = ok ? c : 0.0f;
        spr[j] = 0.0f;
        if (MODE == kReinject) {
            const bool in = ok && gz >= S.sd->xlo1 && gz <= S.sd->xhi1 && gxl >= S.sd->xlo2 && gxl <= S.sd->xhi2;
            const float v = snap_load_w<SNAP>(S, S.sp, S.sp16, 0, ic, clampi(gz, S.sd->xlo1, S.sd->xhi1),
                                              clampi(gxl, S.sd->xlo2, S.sd->xhi2), wseg);
            spr[j] = in ? v : 0.0f;
        }
    }
    // side points of the ring: columns x0-1, x0+TX, x0+TX+1 on the tile rows,
    // taken by the highest threads (the wavefront with the fewest core rows)
    const int sid = kBlock - 1 - tid;
    int s_r = 0, s_c = 0, s_gz = -1, s_gx = -1;
    float s_px = 0.0f, s_pz = 0.0f, s_ka = 0.0f, s_sp = 0.0f;
    const bool side = sid < 3 * TZ;
    if (side) {
        const int k = sid / TZ;
        s_r = 1 + (sid - k * TZ);
        s_c = k == 0 ? 0 : TXT + k;
        s_gz = z0 - 1 + s_r;
        s_gx = x0 - 1 + s_c;
        if (s_gz < n1e && s_gx >= 0 && s_gx < n2e) {
            const int g = s_gz * ldx + s_gx;
            s_px = ld(pxo, (unsigned)g);
            s_pz = ld(pzo, (unsigned)g);
            s_ka = ld(S.kappa, (unsigned)g);
            if (MODE == kReinject && s_gz >= S.sd->xlo1 && s_gz <= S.sd->xhi1 && s_gx >= S.sd->xlo2 && s_gx <= S.sd->xhi2)
                s_sp = snap_load<SNAP>(S, S.sp, S.sp16, 0, ic, s_gz, s_gx);
        }
    }
    float svxr[VROWS], svzr[VROWS];
#pragma unroll
    for (int j = 0; j < VROWS; ++j) {
        const int gz = z0 + rz + NWZ * j;
        svxr[j] = 0.0f; svzr[j] = 0.0f;
        if (MODE == kReinject) {
            const bool in = colok && gz >= npml && gz < npml + n1 && gxl >= npml && gxl < npml + n2;
            const int cz = clampi(gz, npml, npml + n1 - 1), cx = clampi(gxl, npml, npml + n2 - 1);
            const float a = snap_load_w<SNAP>(S, S.svx, S.svx16, 1, ic, cz, cx, wseg);
            const float b = snap_load_w<SNAP>(S, S.svz, S.svz16, 2, ic, cz, cx, wseg);
            svxr[j] = in ? a : 0.0f;
            svzr[j] = in ? b : 0.0f;
        }
    }
#pragma unroll
    for (int j = 0; j < XR; ++j) {
        const int r = rz + NWZ * j;
        if (r < VXH_) s_vx[r * VXW_ + cl + 3] = vxl[j];
    }
#pragma unroll
    for (int j = 0; j < ZR; ++j) {
        const int r = rz + NWZ * j;
        if (r < VZH_) s_vz[r * VZW_ + cl + 1] = vzl[j];
    }
#pragma unroll
    for (int q = 0; q < HXN; ++q)
        if (hx_i[q] >= 0) s_vx[hx_i[q]] = hx[q];
#pragma unroll
    for (int q = 0; q < HZN; ++q)
        if (hz_i[q] >= 0) s_vz[hz_i[q]] = hz[q];
    const float c_d2a = colok ? S.fd2a[gxl] : 1.0f, c_d2a1 = colok ? S.fd2a1[gxl] : 1.0f;
    __syncthreads();

    // ---- 2. pressure half-step; a tile whose pressure rows (z0-1 .. z0+TZ+1)
    // miss the rows of the Hicks windows looks up no injection entry
    const bool injt = MODE != kReinject && S.sd->inj_row != nullptr && !S.sd->vinj && S.sd->inj_z0 <= z0 + TZ + 1 &&
                      S.sd->inj_z1 >= z0 - 1;
#pragma unroll
    for (int j = 0; j < PROWS; ++j) {
        const int r = rz + NWZ * j;
        if (r < PH_) {
            const int gz = z0 - 1 + r;
            float p = 0.0f;
            if (gz >= 0 && gz < n1e && colok) {
                const int injq = injt ? glb(S.sd->inj_row)[gz] : -1;
                p = p_point<MODE, VZW_>(S, it, ic, gz, gxl, pxr[j], pzr[j], kar[j], &s_vx[r * VXW_ + cl + 3],
                                        &s_vz[(r + 2) * VZW_ + cl + 1], S.fd1a[gz], S.fd1a1[gz], c_d2a, c_d2a1,
                                        injq, r >= 1 && r <= TZ, lo, hiz, hix, spr[j], pxn, pzn);
            }
            s_p[r * PW_ + cl + 1] = p;
            if (MODE == kBack && r >= 1 && r <= TZ) {      // snapshot sp(:,:,it) (:1060-1064)
                const bool rowi = gz >= npml && gz < npml + n1;
                if (rowi)
                    snap_store_wave<SNAP>(S, S.sp, S.sp16, 0, it, gz, gxl, p,
                                          colok && gxl >= npml && gxl < npml + n2, rowi);
            }
        }
    }
    if (side) {
        float p = 0.0f;
        if (s_gz < n1e && s_gx >= 0 && s_gx < n2e) {
            const int injq = injt ? glb(S.sd->inj_row)[s_gz] : -1;
            p = p_point<MODE, VZW_>(S, it, ic, s_gz, s_gx, s_px, s_pz, s_ka, &s_vx[s_r * VXW_ + s_c + 2],
                                    &s_vz[(s_r + 2) * VZW_ + s_c], S.fd1a[s_gz], S.fd1a1[s_gz], S.fd2a[s_gx],
                                    S.fd2a1[s_gx], injq, false, lo, hiz, hix, s_sp, pxn, pzn);
        }
        s_p[s_r * PW_ + s_c] = p;
    }
    __syncthreads();

    // ---- 3. velocity half-step of the owned points (lanes past the grid
    // stay in the loop: the snapshot store reduces over whole wavefronts)
    const float a0 = 1.125f;
    const float a1 = -(1.0f / 24.0f);
    const float c_d2b = colok ? S.fd2b[gxl] : 1.0f, c_bw2 = colok ? S.fbw2[gxl] : 0.0f;
#pragma unroll
    for (int j = 0; j < VROWS; ++j) {
        const int r = rz + NWZ * j;
        const int gz = z0 + r, gx = gxl;
        if (gz >= n1e) break;
        const unsigned g = (unsigned)(gz * ldx + gx);
        const bool rowi = gz >= npml && gz < npml + n1;
        const bool interior = rowi && gx >= npml && gx < npml + n2;
        float vxv = 0.0f, vzv = 0.0f;
        if (colok) {
            vxv = s_vx[(r + 1) * VXW_ + cl + 3];
            vzv = s_vz[(r + 3) * VZW_ + cl + 1];
            if (in_box(gz, gx, lo, hiz, hix)) {
                const float *pp = &s_p[(r + 1) * PW_ + cl + 1];         // p(gz, gx)
                const float der2 = a0 * (pp[1] - pp[0]) + a1 * (pp[2] - pp[-1]);
                const float der1 = a0 * (pp[PW_] - pp[0]) + a1 * (pp[2 * PW_] - pp[-PW_]);
                if (MODE == kBack) {
                    vxv = c_d2b * vxv - c_bw2 * der2;
                    vzv = S.fd1b[gz] * vzv - S.fbw1[gz] * der1;
                } else {
                    vxv = c_d2b * vxv + c_bw2 * der2;
                    vzv = S.fd1b[gz] * vzv + S.fbw1[gz] * der1;
                }
            }
            if (MODE == kFwd && S.sd->vinj && S.sd->inj_row != nullptr) {
                const int q = glb(S.sd->inj_row)[gz];
                if (q >= 0) {
                    const GLB int *ptr = glb(S.sd->inj_ptr) + q * (n2e + 1);
                    for (int k = ptr[gx]; k < ptr[gx + 1]; ++k)
                        vzv = vzv + glb(S.sd->inj_cdt)[k] * glb(S.sd->inj_amp)[glb(S.sd->inj_r)[k] * S.sd->amp_stride + it * S.sd->amp_tstride];   // * bue (= 1), :252
                }
            }
            if (MODE == kReinject && interior) {
                vxv = vxv + (-S.dt) * svxr[j];                         // * bue (= 1), :1177
                vzv = vzv + (-S.dt) * svzr[j];                         // :1184
            }
            st(vxn, g, vxv);
            if (S.ofs == 1 && gz == npml - 1) {
                // vz(npml,:) is overwritten by the mirror below; its owner writes nothing
            } else {
                st(vzn, g, vzv);
                if (S.ofs == 1 && gz == npml) st(vzn, g - ldx, vzv);  // vz(npml,:) = vz(npml+1,:)
            }
        }
        if (MODE == kBack && rowi) {                          // svx, svz(:,:,it) (:1082-1087), whole wavefront
            snap_store_wave<SNAP>(S, S.svx, S.svx16, 1, it, gz, gx, vxv, interior, true);
            snap_store_wave<SNAP>(S, S.svz, S.svz16, 2, it, gz, gx, vzv, interior, true);
        }
    }
}

// ---------------------------------------------------------------------------
// Exact discrete adjoint of one forward time step (reverse time).  The state
// buffers hold the adjoint fields (qpx, qpz, qvx, qvz) of the step's outputs
// and receive those of its inputs.  With x the forward step
//   A  px1 = damp2a px + kappaw2 D-x vx,  pz1 = damp1a pz + kappaw1 D-z vz (box)
//   B  px2 = px1 + source;  C  p = M (px2 + pz1);  D  record R p
//   E  vx1 = damp2b vx + buw2 D+x p,  vz1 = damp1b vz + buw1 D+z p (box)
//   F  vz2 = vz1 with row npml taking row npml+1 (1-based)
// the kernel applies F^T, E^T, D^T, C^T, B^T, A^T.  Per tile it has the same
// halo structure as k_step: F^T is applied while staging qvz, the adjoint
// pressure qp = (D+x)^T(buw2 qvx) + (D+z)^T(buw1 qvz) + R^T d is formed on the
// tile plus its ring, and (D-)^T(kappaw q) on the ring feeds qvx, qvz.
//
// HESS = false: transpose of submodeling (stencil box 3..n-2); the source
//   gradient dt*kappa*qpx2 goes to the record band (B^T, summed by the
//   record blocks with the source's Hicks tile).
// HESS = true: transpose of the re-injection pass (box 3..n-3); the adjoint of
//   its extended sources, -dt*kappa*qpx2 on the pressure range and -dt*qvx,
//   -dt*qvz on the interior, is stored as the snapshot of index nt-1-it.
// ---------------------------------------------------------------------------
template <int TZ, int TXW, bool HESS>
__global__ __launch_bounds__(kBlock) void k_adj(StepParams P)
{
    constexpr int NW = kBlock / 64;
    constexpr int TXT = 64 * TXW;
    constexpr int NWZ = NW / TXW;
    constexpr int PW_ = TXT + 3, PH_ = TZ + 3;
    constexpr int VXW_ = TXT + 6, VXH_ = TZ + 3;
    constexpr int VZW_ = TXT + 3, VZH_ = TZ + 6;
    constexpr int VROWS = TZ / NWZ;
    // the wavefront index is uniform: in an SGPR, row-level conditions and loads are scalar
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int LDS_N = VXH_ * VXW_ + VZH_ * VZW_ + 2 * PH_ * PW_;
    static_assert(LDS_N >= kRecPerBlock * kRecStride, "record blocks reuse the tile's LDS");
    __shared__ __align__(16) float s_lds[LDS_N];
    if ((int)blockIdx.x < P.nrec_blocks) {
        record_block(P, blockIdx.x, s_lds);
        return;
    }
    float *const s_qvx = s_lds;
    float *const s_qvz = s_lds + VXH_ * VXW_;
    float *const s_wx = s_qvz + VZH_ * VZW_;
    float *const s_wz = s_wx + PH_ * PW_;

    const int2 t = P.tiles[blockIdx.x - P.nrec_blocks];
    const int it = P.it;
    const View S = make_view(P.shots + t.x, P.cur, it);
    if (it >= S.nt) return;
    const int n1e = S.n1e, n2e = S.n2e, ldx = S.ldx, npml = S.npml, n1 = S.n1, n2 = S.n2;
    const int ntx = (n2e + TXT - 1) / TXT;
    const int z0 = (t.y / ntx) * TZ;
    const int x0 = (t.y % ntx) * TXT;
    const int lo = 2, hiz = HESS ? n1e - 4 : n1e - 3, hix = HESS ? n2e - 4 : n2e - 3;
    const int cl = (wv % TXW) * 64 + lane;
    const int rz = wv / TXW;
    const int gxl = x0 + cl;
    const bool colok = gxl < n2e;
    const int ic = S.nt - 1 - it;
    const float a0 = 1.125f;
    const float a1 = -(1.0f / 24.0f);
    const bool fs = S.ofs == 1;
    // F^T on load: the adjoint of vz row npml-1 (0-based) moves onto row npml
    auto qvz_at = [&](int gz, int gx) -> float {
        if (gz < 0 || gz >= n1e || gx < 0 || gx >= n2e) return 0.0f;
        if (fs && gz == npml - 1) return 0.0f;
        const float v = S.vzo[gz * ldx + gx];
        return (fs && gz == npml) ? v + S.vzo[(gz - 1) * ldx + gx] : v;
    };
    for (int r = rz; r < VXH_; r += NWZ) {
        const int gz = z0 - 1 + r;
        s_qvx[r * VXW_ + cl + 3] = (gz >= 0 && gz < n1e && colok) ? S.vxo[gz * ldx + gxl] : 0.0f;
    }
    for (int r = rz; r < VZH_; r += NWZ) s_qvz[r * VZW_ + cl + 1] = qvz_at(z0 - 3 + r, gxl);
    for (int i = tid; i < VXH_ * 6; i += kBlock) {
        const int r = i / 6, k = i - 6 * r;
        const int c = k < 3 ? k : TXT + k;
        const int gz = z0 - 1 + r, gx = x0 - 3 + c;
        s_qvx[r * VXW_ + c] = (gz >= 0 && gz < n1e && gx >= 0 && gx < n2e) ? S.vxo[gz * ldx + gx] : 0.0f;
    }
    for (int i = tid; i < VZH_ * 3; i += kBlock) {
        const int r = i / 3, k = i - 3 * r;
        const int c = k == 0 ? 0 : TXT + k;
        s_qvz[r * VZW_ + c] = qvz_at(z0 - 3 + r, x0 - 1 + c);
    }
    __syncthreads();

    // ring point: qp, then the adjoint pressures and the (D-)^T weights
    auto ring = [&](int r, int c, bool owned) {
        const int gz = z0 - 1 + r, gx = x0 - 1 + c;
        float wx = 0.0f, wz = 0.0f;
        if (gz >= 0 && gz < n1e && gx >= 0 && gx < n2e) {
            const unsigned g = (unsigned)(gz * ldx + gx);
            const float kap = S.kappa[g];
            // U = buw * qv on the box, zero outside
            auto ux = [&](int xx) -> float {
                return in_box(gz, xx, lo, hiz, hix) ? S.bw2[xx] * s_qvx[r * VXW_ + (xx - (x0 - 3))] : 0.0f;
            };
            auto uz = [&](int zz) -> float {
                return in_box(zz, gx, lo, hiz, hix) ? S.bw1[zz] * s_qvz[(zz - (z0 - 3)) * VZW_ + c] : 0.0f;
            };
            float qp = a0 * ux(gx - 1) - a0 * ux(gx) + a1 * ux(gx - 2) - a1 * ux(gx + 1);
            qp += a0 * uz(gz - 1) - a0 * uz(gz) + a1 * uz(gz - 2) - a1 * uz(gz + 1);
            if (S.sd->inj_row != nullptr) {              // R^T d: adjoint data spread with the Hicks tiles
                const int q = glb(S.sd->inj_row)[gz];
                if (q >= 0) {
                    const GLB int *ptr = glb(S.sd->inj_ptr) + q * (n2e + 1);
                    const int ia = S.sd->amp_rev ? S.nt - 1 - it : it;
                    for (int k = ptr[gx]; k < ptr[gx + 1]; ++k)
                        qp += glb(S.sd->inj_cdt)[k] * glb(S.sd->inj_amp)[glb(S.sd->inj_r)[k] * S.sd->amp_stride + ia * S.sd->amp_tstride];
                }
            }
            if (fs && gz <= npml) qp = 0.0f;          // C^T: M
            const float qpx = S.pxo[g] + qp, qpz = S.pzo[g] + qp;
            const bool box = in_box(gz, gx, lo, hiz, hix);
            const float tk = S.dtdx * kap;
            wx = box ? (tk * S.d2a1[gx]) * qpx : 0.0f;
            wz = box ? (tk * S.d1a1[gz]) * qpz : 0.0f;
            if (owned) {
                S.pxn[g] = box ? S.d2a[gx] * qpx : qpx;
                S.pzn[g] = box ? S.d1a[gz] * qpz : qpz;
                const bool interior = gz >= npml && gz < npml + n1 && gx >= npml && gx < npml + n2;
                if (HESS) {
                    if (gz >= S.sd->xlo1 && gz <= S.sd->xhi1 && gx >= S.sd->xlo2 && gx <= S.sd->xhi2)
                        st_nt(splane_at(S.sp, S, ic), snap_off(S, gz, gx), -S.dt * kap * qpx);
                    else if (interior)
                        st_nt(splane_at(S.sp, S, ic), snap_off(S, gz, gx), 0.0f);
                } else if (S.sd->band_row != nullptr) {
                    const int slot = glb(S.sd->band_row)[gz];
                    if (slot >= 0) glb(S.sd->pband + (size_t)S.par * S.sd->band_stride)[slot * ldx + gx] = S.dt * kap * qpx;
                }
            }
        }
        s_wx[r * PW_ + c] = wx;
        s_wz[r * PW_ + c] = wz;
    };
    for (int r = rz; r < PH_; r += NWZ) {
        if (colok) ring(r, cl + 1, r >= 1 && r <= TZ);
        else { s_wx[r * PW_ + cl + 1] = 0.0f; s_wz[r * PW_ + cl + 1] = 0.0f; }
    }
    {
        const int sid = kBlock - 1 - tid;
        if (sid < 3 * TZ) {
            const int k = sid / TZ;
            ring(1 + (sid - k * TZ), k == 0 ? 0 : TXT + k, false);
        }
    }
    __syncthreads();

    if (!colok) return;
    for (int j = 0; j < VROWS; ++j) {
        const int r = rz + NWZ * j;
        const int gz = z0 + r, gx = gxl;
        if (gz >= n1e) break;
        const unsigned g = (unsigned)(gz * ldx + gx);
        const float qvx = s_qvx[(r + 1) * VXW_ + cl + 3];
        const float qvz = s_qvz[(r + 3) * VZW_ + cl + 1];
        const bool box = in_box(gz, gx, lo, hiz, hix);
        const float *wxr = &s_wx[(r + 1) * PW_ + cl + 1];
        const float *wzc = &s_wz[(r + 1) * PW_ + cl + 1];
        const float ax = a0 * wxr[0] - a0 * wxr[1] + a1 * wxr[-1] - a1 * wxr[2];          // (D-x)^T
        const float az = a0 * wzc[0] - a0 * wzc[PW_] + a1 * wzc[-PW_] - a1 * wzc[2 * PW_]; // (D-z)^T
        S.vxn[g] = (box ? S.d2b[gx] * qvx : qvx) + ax;
        S.vzn[g] = (box ? S.d1b[gz] * qvz : qvz) + az;
        if (HESS && gz >= npml && gz < npml + n1 && gx >= npml && gx < npml + n2) {
            const unsigned sidx = snap_off(S, gz, gx);
            st_nt(splane_at(S.svx, S, ic), sidx, -S.dt * qvx);
            st_nt(splane_at(S.svz, S, ic), sidx, -S.dt * qvz);
        }
    }
}

// ---------------------------------------------------------------------------
// Two time steps per launch (it and it+1) for every active shot of a batch:
// the wave state is read from HBM once and written once per two steps.
//
// A workgroup owns a TZ x 64*NC core tile.  It stages the state of the core
// plus a 6-cell ring in LDS (vx, vz, and p once formed; px, pz and kappa stay
// in registers of the thread that owns the point) and advances the shrinking
// regions each half-step needs (region rows / columns relative to the ring
// origin z0-6, x0-6, inclusive):
//   P_A [2, TZ+11] x [2, TX+10]    V_A [3, TZ+9] x [3, TX+8]
//   P_B [5, TZ+7]  x [5, TX+7]     V_B [6, TZ+5] x [6, TX+5] (the core)
// V_A keeps one row more than P_B reads so that the free-surface mirror
// vz(npml) = vz(npml+1) always has its source row.  Every point is updated
// with the same operations in the same order as in k_step, so the results are
// bit-identical; only the core's final state, record band, snapshots and
// wavefield are written.
//
// Thread layout: "main" items are (region row, 64-column core segment) pairs,
// one per wavefront with lane = column (aligned 256-byte rows, and the 16-bit
// snapshot segment is exactly one wavefront); the 12 ring columns of every
// row are "halo" items spread over all threads.
// ---------------------------------------------------------------------------
template <int MODE, int RW>
__device__ __forceinline__ float p_half(const View &S, int t, int gz, int gx, float &pxv, float &pzv, float kap,
                                        const float *vr, const float *vc, float d1a, float d1a1, float d2a, float d2a1,
                                        int injq, int lo, int hiz, int hix, float spv)
{
    const float a0 = 1.125f;
    const float a1 = -(1.0f / 24.0f);
    if (in_box(gz, gx, lo, hiz, hix)) {
        const float der2 = a0 * (vr[0] - vr[-1]) + a1 * (vr[1] - vr[-2]);
        const float der1 = a0 * (vc[0] - vc[-RW]) + a1 * (vc[RW] - vc[-2 * RW]);
        const float tk = S.dtdx * kap;                      // kappaw = dtdx*kappa*damp1 (:105-106)
        const float kw2 = tk * d2a1;
        const float kw1 = tk * d1a1;
        if (MODE == kBack) {
            pxv = d2a * pxv - kw2 * der2;
            pzv = d1a * pzv - kw1 * der1;
        } else {
            pxv = d2a * pxv + kw2 * der2;
            pzv = d1a * pzv + kw1 * der1;
        }
    }
    if (MODE != kReinject && injq >= 0) {
        const GLB int *ptr = glb(S.sd->inj_ptr) + injq * (S.n2e + 1);
        const int kb = ptr[gx], ke = ptr[gx + 1];
        for (int k = kb; k < ke; ++k) {
            const float term = (glb(S.sd->inj_cdt)[k] * glb(S.sd->inj_amp)[glb(S.sd->inj_r)[k] * S.sd->amp_stride + t * S.sd->amp_tstride]) * kap;
            if (MODE == kBack) pxv = pxv - term;   // :1042
            else pxv = pxv + term;                 // :175
        }
    }
    if (MODE == kReinject) pxv = pxv + ((-S.dt) * spv) * kap;   // ss*dt*sp*kappa, ss = -1 (:1123); spv = 0 off range
    float p = pxv + pzv;
    if (S.ofs == 1 && gz <= S.npml) p = 0.0f;           // p(1:npml+1,:) = 0
    return p;
}

template <int MODE, int RW>
__device__ __forceinline__ void v_half(const View &S, int t, int gz, int gx, float &vxv, float &vzv, const float *pp,
                                       float d1b, float bw1, float d2b, float bw2, int lo, int hiz, int hix, bool interior,
                                       float svx, float svz)
{
    const float a0 = 1.125f;
    const float a1 = -(1.0f / 24.0f);
    if (in_box(gz, gx, lo, hiz, hix)) {
        const float der2 = a0 * (pp[1] - pp[0]) + a1 * (pp[2] - pp[-1]);
        const float der1 = a0 * (pp[RW] - pp[0]) + a1 * (pp[2 * RW] - pp[-RW]);
        if (MODE == kBack) {
            vxv = d2b * vxv - bw2 * der2;
            vzv = d1b * vzv - bw1 * der1;
        } else {
            vxv = d2b * vxv + bw2 * der2;
            vzv = d1b * vzv + bw1 * der1;
        }
    }
    if (MODE == kFwd && S.sd->vinj && S.sd->inj_row != nullptr) {
        const int q = glb(S.sd->inj_row)[gz];
        if (q >= 0) {
            const GLB int *ptr = glb(S.sd->inj_ptr) + q * (S.n2e + 1);
            for (int k = ptr[gx]; k < ptr[gx + 1]; ++k)
                vzv = vzv + glb(S.sd->inj_cdt)[k] * glb(S.sd->inj_amp)[glb(S.sd->inj_r)[k] * S.sd->amp_stride + t * S.sd->amp_tstride];   // :252
        }
    }
    if (MODE == kReinject && interior) {
        vxv = vxv + (-S.dt) * svx;                         // * bue (= 1), :1177
        vzv = vzv + (-S.dt) * svz;                         // :1184
    }
}

// p of an owned point: record band and (single-shot calls) the wavefield
template <int MODE>
__device__ __forceinline__ void p_out(const View &S, int t, int gz, int gx, float p)
{
    if (MODE == kBack) return;
    const int npml = S.npml;
    if (S.sd->band_row != nullptr) {
        const int slot = glb(S.sd->band_row)[gz];
        if (slot >= 0) glb(S.sd->pband + (size_t)(t & (kBandSlots - 1)) * S.sd->band_stride)[slot * S.ldx + gx] = p;
    }
    if (S.sd->wf != nullptr && gz >= npml && gz < npml + S.n1 && gx >= npml && gx < npml + S.n2)
        plane_at(glb(S.sd->wf), S, t)[(gx - npml) * S.n1 + (gz - npml)] = p;
}

template <int MODE, int SNAP, int TZ, int NC>
__global__ __launch_bounds__(kBlock) void k_step2(StepParams P)
{
    constexpr int TXT = 64 * NC;
    constexpr int RH = TZ + 13, RW = TXT + 12;
    constexpr int NTASK = RH * NC, TPW = (NTASK + 3) / 4;
    constexpr int NHALO = 12 * RH, HPT = (NHALO + kBlock - 1) / kBlock;
    constexpr int LDS_N = 3 * RH * RW;
    static_assert(LDS_N >= kRecPerBlock * kRecStride, "record blocks reuse the tile's LDS");
    __shared__ __align__(16) float s_lds[LDS_N];
    // the wavefront index is uniform: in an SGPR, row-level conditions and loads are scalar
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    if ((int)blockIdx.x < P.nrec_blocks) {
        record_block(P, blockIdx.x, s_lds);
        return;
    }
    float *const s_vx = s_lds;
    float *const s_vz = s_lds + RH * RW;
    float *const s_p = s_vz + RH * RW;

    const int2 tl = P.tiles[blockIdx.x - P.nrec_blocks];
    const int itA = P.it, itB = P.it + 1;
    const View S = make_view(P.shots + tl.x, P.cur, itA);
    if (itA >= S.nt) return;
    const bool two = P.nsub == 2 && itB < S.nt;
    const int n1e = S.n1e, n2e = S.n2e, ldx = S.ldx, npml = S.npml, n1 = S.n1, n2 = S.n2;
    const int ntx = (n2e + TXT - 1) / TXT;
    const int z0 = (tl.y / ntx) * TZ, x0 = (tl.y % ntx) * TXT;
    const int zr = z0 - 6, xr = x0 - 6;                    // region origin
    // stencil bounds: 3..n-2 forward (submodeling2.f90:148-149), 3..n-3 for
    // both passes of the Hessian-type sweeps (:1014-1015, :1108-1109)
    const int lo = 2;
    const int hiz = (MODE == kFwd) ? n1e - 3 : n1e - 4;
    const int hix = (MODE == kFwd) ? n2e - 3 : n2e - 4;
    const int icA = S.nt - 1 - itA, icB = S.nt - 2 - itA;   // time-reversed snapshot index
    const bool inj = MODE != kReinject && S.sd->inj_row != nullptr && !S.sd->vinj;
    const int xlo1 = S.sd->xlo1, xhi1 = S.sd->xhi1, xlo2 = S.sd->xlo2, xhi2 = S.sd->xhi2;

    // per-lane column profiles of the core segments
    float c_d2a[NC], c_d2a1[NC], c_d2b[NC], c_bw2[NC];
#pragma unroll
    for (int sg = 0; sg < NC; ++sg) {
        const unsigned gx = (unsigned)min(x0 + 64 * sg + lane, n2e - 1);
        c_d2a[sg] = ld(S.d2a, gx);
        c_d2a1[sg] = ld(S.d2a1, gx);
        c_d2b[sg] = ld(S.d2b, gx);
        c_bw2[sg] = ld(S.bw2, gx);
    }

    // ---- 0. the state of the region: vx, vz to LDS, px, pz, kappa to registers
    float m_px[TPW], m_pz[TPW], m_ka[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        m_px[i] = 0.0f; m_pz[i] = 0.0f; m_ka[i] = 0.0f;
        const int q = wv + 4 * i;
        if (q < NTASK) {
            const int r = q / NC, sg = q - NC * r;
            const int gz = zr + r, gx = x0 + 64 * sg + lane, c = 6 + 64 * sg + lane;
            const bool ok = gz >= 0 && gz < n1e && gx < n2e;
            const unsigned g = (unsigned)(clampi(gz, 0, n1e - 1) * ldx + min(gx, n2e - 1));
            const float a = ld(S.vxo, g), b = ld(S.vzo, g);
            s_vx[r * RW + c] = ok ? a : 0.0f;
            s_vz[r * RW + c] = ok ? b : 0.0f;
            if (r >= 2 && r <= TZ + 11) {
                const float u = ld(S.pxo, g), v = ld(S.pzo, g), k = ld(S.kappa, g);
                m_px[i] = ok ? u : 0.0f;
                m_pz[i] = ok ? v : 0.0f;
                m_ka[i] = ok ? k : 0.0f;
            }
        }
    }
    float h_px[HPT], h_pz[HPT], h_ka[HPT];
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
        h_px[j] = 0.0f; h_pz[j] = 0.0f; h_ka[j] = 0.0f;
        const int k = tid + kBlock * j;
        if (k < NHALO) {
            const int r = k / 12, kk = k - 12 * r, c = kk < 6 ? kk : TXT + kk;
            const int gz = zr + r, gx = xr + c;
            const bool ok = gz >= 0 && gz < n1e && gx >= 0 && gx < n2e;
            const unsigned g = (unsigned)(clampi(gz, 0, n1e - 1) * ldx + clampi(gx, 0, n2e - 1));
            const float a = ld(S.vxo, g), b = ld(S.vzo, g), u = ld(S.pxo, g), v = ld(S.pzo, g), w = ld(S.kappa, g);
            s_vx[r * RW + c] = ok ? a : 0.0f;
            s_vz[r * RW + c] = ok ? b : 0.0f;
            h_px[j] = ok ? u : 0.0f;
            h_pz[j] = ok ? v : 0.0f;
            h_ka[j] = ok ? w : 0.0f;
        }
    }
    __syncthreads();

    // ---- 1. and 3. pressure half-steps of steps A (region P_A) and B (P_B)
#pragma unroll
    for (int stp = 0; stp < 2; ++stp) {
        if (stp == 1 && !two) break;
        const int t = stp == 0 ? itA : itB, ic = stp == 0 ? icA : icB;
        const int rlo = stp == 0 ? 2 : 5, rhi = stp == 0 ? TZ + 11 : TZ + 7;
        const int clo = rlo, chi = stp == 0 ? TXT + 10 : TXT + 7;
        const bool last = stp == 1 || !two;
#pragma unroll
        for (int i = 0; i < TPW; ++i) {
            const int q = wv + 4 * i;
            if (q >= NTASK) break;
            const int r = q / NC, sg = q - NC * r;
            if (r < rlo || r > rhi) continue;
            const int gz = zr + r, gx = x0 + 64 * sg + lane, c = 6 + 64 * sg + lane;
            const bool ok = gz >= 0 && gz < n1e && gx < n2e;
            const int wseg = min((x0 >> 6) + sg, S.nseg - 1);
            float spv = 0.0f;
            if (MODE == kReinject) {
                const bool in = ok && gz >= xlo1 && gz <= xhi1 && gx >= xlo2 && gx <= xhi2;
                const float v = snap_load_w<SNAP>(S, S.sp, S.sp16, 0, ic, clampi(gz, xlo1, xhi1), clampi(gx, xlo2, xhi2),
                                                  wseg);
                spv = in ? v : 0.0f;
            }
            const int gzc = clampi(gz, 0, n1e - 1);
            const float d1a = ld(S.d1a, (unsigned)gzc), d1a1 = ld(S.d1a1, (unsigned)gzc);
            const int injq = inj ? glb(S.sd->inj_row)[gzc] : -1;
            float p = 0.0f;
            if (ok)
                p = p_half<MODE, RW>(S, t, gz, gx, m_px[i], m_pz[i], m_ka[i], &s_vx[r * RW + c], &s_vz[r * RW + c], d1a,
                                     d1a1, c_d2a[sg], c_d2a1[sg], injq, lo, hiz, hix, spv);
            s_p[r * RW + c] = p;
            if (r >= 6 && r <= TZ + 5) {                          // the core
                if (ok) {
                    p_out<MODE>(S, t, gz, gx, p);
                    if (last) {
                        const unsigned g = (unsigned)(gz * ldx + gx);
                        st(S.pxn, g, m_px[i]);
                        st(S.pzn, g, m_pz[i]);
                    }
                }
                if (MODE == kBack && gz >= npml && gz < npml + n1)   // snapshot sp(:,:,t) (:1060-1064)
                    snap_store_wave<SNAP>(S, S.sp, S.sp16, 0, t, gz, gx, p, ok && gx >= npml && gx < npml + n2, true);
            }
        }
#pragma unroll
        for (int j = 0; j < HPT; ++j) {
            const int k = tid + kBlock * j;
            if (k < NHALO) {
                const int r = k / 12, kk = k - 12 * r, c = kk < 6 ? kk : TXT + kk;
                const int gz = zr + r, gx = xr + c;
                if (r >= rlo && r <= rhi && c >= clo && c <= chi) {
                    const bool ok = gz >= 0 && gz < n1e && gx >= 0 && gx < n2e;
                    float p = 0.0f;
                    if (ok) {
                        float spv = 0.0f;
                        if (MODE == kReinject && gz >= xlo1 && gz <= xhi1 && gx >= xlo2 && gx <= xhi2)
                            spv = snap_load<SNAP>(S, S.sp, S.sp16, 0, ic, gz, gx);
                        const int injq = inj ? glb(S.sd->inj_row)[gz] : -1;
                        p = p_half<MODE, RW>(S, t, gz, gx, h_px[j], h_pz[j], h_ka[j], &s_vx[r * RW + c],
                                             &s_vz[r * RW + c], ld(S.d1a, (unsigned)gz), ld(S.d1a1, (unsigned)gz),
                                             ld(S.d2a, (unsigned)gx), ld(S.d2a1, (unsigned)gx), injq, lo, hiz, hix, spv);
                    }
                    s_p[r * RW + c] = p;
                }
            }
        }
        __syncthreads();

        // ---- 2. and 4. velocity half-steps of steps A (region V_A) and B (the core)
        const int vlo = stp == 0 ? 3 : 6, vhi = stp == 0 ? TZ + 9 : TZ + 5;
#pragma unroll
        for (int i = 0; i < TPW; ++i) {
            const int q = wv + 4 * i;
            if (q >= NTASK) break;
            const int r = q / NC, sg = q - NC * r;
            if (r < vlo || r > vhi) continue;
            const int gz = zr + r, gx = x0 + 64 * sg + lane, c = 6 + 64 * sg + lane;
            const bool ok = gz >= 0 && gz < n1e && gx < n2e;
            const bool rowi = gz >= npml && gz < npml + n1;
            const bool interior = rowi && gx >= npml && gx < npml + n2;
            float svx = 0.0f, svz = 0.0f;
            if (MODE == kReinject) {
                const int wseg = min((x0 >> 6) + sg, S.nseg - 1);
                const int cz = clampi(gz, npml, npml + n1 - 1), cx = clampi(gx, npml, npml + n2 - 1);
                const float a = snap_load_w<SNAP>(S, S.svx, S.svx16, 1, ic, cz, cx, wseg);
                const float b = snap_load_w<SNAP>(S, S.svz, S.svz16, 2, ic, cz, cx, wseg);
                svx = interior && ok ? a : 0.0f;
                svz = interior && ok ? b : 0.0f;
            }
            const int gzc = clampi(gz, 0, n1e - 1);
            const float d1b = ld(S.d1b, (unsigned)gzc), bw1 = ld(S.bw1, (unsigned)gzc);
            float vxv = 0.0f, vzv = 0.0f;
            if (ok) {
                vxv = s_vx[r * RW + c];
                vzv = s_vz[r * RW + c];
                v_half<MODE, RW>(S, t, gz, gx, vxv, vzv, &s_p[r * RW + c], d1b, bw1, c_d2b[sg], c_bw2[sg], lo, hiz, hix,
                                 interior, svx, svz);
                s_vx[r * RW + c] = vxv;
                s_vz[r * RW + c] = vzv;
                if (last && r >= 6 && r <= TZ + 5) {             // final state of the core
                    const unsigned g = (unsigned)(gz * ldx + gx);
                    st(S.vxn, g, vxv);
                    if (S.ofs == 1 && gz == npml - 1) {
                        // vz(npml,:) is overwritten by the mirror below; its owner writes nothing
                    } else {
                        st(S.vzn, g, vzv);
                        if (S.ofs == 1 && gz == npml) st(S.vzn, g - ldx, vzv);   // vz(npml,:) = vz(npml+1,:)
                    }
                }
            }
            if (MODE == kBack && r >= 6 && r <= TZ + 5 && rowi) {   // svx, svz(:,:,t) (:1082-1087)
                snap_store_wave<SNAP>(S, S.svx, S.svx16, 1, t, gz, gx, vxv, interior && ok, true);
                snap_store_wave<SNAP>(S, S.svz, S.svz16, 2, t, gz, gx, vzv, interior && ok, true);
            }
        }
        if (stp == 0) {
#pragma unroll
            for (int j = 0; j < HPT; ++j) {
                const int k = tid + kBlock * j;
                if (k < NHALO) {
                    const int r = k / 12, kk = k - 12 * r, c = kk < 6 ? kk : TXT + kk;
                    const int gz = zr + r, gx = xr + c;
                    if (r >= 3 && r <= TZ + 9 && c >= 3 && c <= TXT + 8 && gz >= 0 && gz < n1e && gx >= 0 &&
                        gx < n2e) {
                        const bool interior = gz >= npml && gz < npml + n1 && gx >= npml && gx < npml + n2;
                        float svx = 0.0f, svz = 0.0f;
                        if (MODE == kReinject && interior) {
                            svx = snap_load<SNAP>(S, S.svx, S.svx16, 1, ic, gz, gx);
                            svz = snap_load<SNAP>(S, S.svz, S.svz16, 2, ic, gz, gx);
                        }
                        float vxv = s_vx[r * RW + c], vzv = s_vz[r * RW + c];
                        v_half<MODE, RW>(S, t, gz, gx, vxv, vzv, &s_p[r * RW + c], ld(S.d1b, (unsigned)gz),
                                         ld(S.bw1, (unsigned)gz), ld(S.d2b, (unsigned)gx), ld(S.bw2, (unsigned)gx), lo,
                                         hiz, hix, interior, svx, svz);
                        s_vx[r * RW + c] = vxv;
                        s_vz[r * RW + c] = vzv;
                    }
                }
            }
            if (!two) return;
            // free surface: vz(npml,:) = vz(npml+1,:) inside the region (1-based rows)
            const int rm = npml - 1 - zr;
            const bool mirror = S.ofs == 1 && rm >= 3 && rm + 1 <= TZ + 9;
            __syncthreads();
            if (mirror) {
                for (int c = 3 + tid; c <= TXT + 8; c += kBlock) s_vz[rm * RW + c] = s_vz[(rm + 1) * RW + c];
                __syncthreads();
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Column march: one time step of one sweep, the arithmetic of k_step (the
// general body, operation for operation), organised for latency instead of
// LDS.  A wavefront owns kMarchCols columns x0 .. x0+57 of H rows z0 ..
// z0+H-1; its 64 lanes hold columns x0-3 .. x0+60, so the x-neighbours the
// two derivative stencils need (x-2 .. x+1 for the pressure, x-1 .. x+2 for
// the velocity) come from DPP lane shifts: the pressure is valid on lanes
// 2..62 and the velocity on lanes 3..60.  The wavefront walks down its rows
// once, keeping the last four old-vz rows, the last three old-vx rows and
// the last four new-p rows in registers: iteration k loads vz of row
// zr = z0-3+k, forms p of row zr-1 (the tile rows plus the one ring row
// above and two below, as k_step recomputes them) and the velocities of
// row zr-3.  No LDS, no barrier; the only cross-lane traffic is the DPP
// shifts, which need every lane of the wavefront active.  Per-row profile
// values (d1a, d1a1, d1b, bw1) and the injection / record-band row maps are
// read once per wavefront into lanes and taken out with readlane per row.
// A workgroup's tile entry holds its shot and (z0 << 16 | rows << 8 | up << 7
// | g): its four wavefronts take strips 4g .. 4g+3 of the row segment z0 ..
// z0+rows-1 (rows <= H; the host splits rows with many Hicks injection
// terms into short segments).  With up set the walk runs bottom to top
// (iteration k loads vz of row z0+rows+2-k, forms p of row zr+2 and the
// velocities of row zr+3), the same operations on the same operands in the
// other row order; alternate segments walk up, so two neighbouring segments
// read the ring rows they share at the same time.  The record blocks come
// first as in k_step.
// FP32 snapshots only (the 16-bit codec scales 64-aligned segments).
// ---------------------------------------------------------------------------

// value of the lane on the left (column x-1) or right (x+1); lanes at the
// wavefront's edges receive 0
__device__ __forceinline__ float from_left(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));   // wave_shr:1
}
__device__ __forceinline__ float from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));   // wave_shl:1
}

// lane k's value (k a constant: the walk's iteration) as a scalar
__device__ __forceinline__ float lane_f(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ int lane_i(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ __forceinline__ unsigned lane_u(unsigned v, int k) { return (unsigned)__builtin_amdgcn_readlane((int)v, k); }
// one bit per lane of the wavefront; bit k of it
__device__ __forceinline__ unsigned long long lanes_where(bool c) { return __builtin_amdgcn_ballot_w64(c); }
__device__ __forceinline__ bool bit(unsigned long long m, int k) { return (m >> k) & 1ull; }
// Raw buffer accesses: a resource (base address) in scalar registers, this
// lane's byte offset (its column, computed once) and a scalar byte offset
// (plane and row, formed on the scalar unit), so a row's load or store costs
// no vector instruction for its address.  No range limit (0xffffffff
// records); dword 3 is the gfx9 raw 32-bit buffer format.
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc make_rsrc(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, -1, 0x00020000);
}
__device__ __forceinline__ float bld(Rsrc r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void bst(Rsrc r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
}
// non-temporal (the snapshot stream; see st_nt)
__device__ __forceinline__ void bst_nt(Rsrc r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 2);
}
// Non-temporal load, for the re-injection's snapshot reads: each is read
// once, and as a streaming access it displaces less of the wave state from
// the caches (201x801, 24 shots, 16-row march without records: 50.9 -> 47.8
// us per step).
__device__ __forceinline__ float bld_nt(Rsrc r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 2));
}

// Rows of loads in flight ahead of the row being computed (5 measured no
// faster than 3 for 24- and 32-row segments at 201x801, 24 shots).
constexpr int kMarchAhead = 3;

// Waves per SIMD the register budget must allow: short segments make many
// wavefronts (201x801 with 24 shots: 7560 at 12 rows), tall ones few.
template <int MODE, int H>
constexpr int march_min_waves()
{
    return H <= 12 ? (MODE == kReinject ? 7 : 6) : H <= 16 ? 6 : 4;
}

template <int MODE, int H>
__global__ __launch_bounds__(kBlock, (march_min_waves<MODE, H>())) void k_march(StepParams P)
{
    static_assert(H + 6 <= 64, "the iterations of a walk must fit the lanes");
    constexpr int D = kMarchAhead;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    __shared__ __align__(16) float s_lds[kRecPerBlock * kRecStride];
    if ((int)blockIdx.x < P.nrec_blocks) {
        record_block(P, blockIdx.x, s_lds);
        return;
    }
    const int2 t = P.tiles[blockIdx.x - P.nrec_blocks];
    const ShotDev *__restrict__ S = P.shots + t.x;
    const int it = P.it;
    const int nt = S->nt;
    if (it >= nt) return;
    const int n1e = S->n1e, n2e = S->n2e, ldx = S->ldx, npml = S->npml, n1 = S->n1, n2 = S->n2;
    const int nstrip = (n2e + kMarchCols - 1) / kMarchCols;
    const int z0 = t.y >> 16, nrows = (t.y >> 8) & 255, strip = 4 * (t.y & 127) + wv;
    const bool up = (t.y >> 7) & 1;              // walk the segment bottom to top
    if (strip >= nstrip) return;                 // whole wavefront: no lane of it takes part in a shift
    const int x0 = strip * kMarchCols;
    const int gx = x0 - 3 + lane;
    const bool colok = gx >= 0 && gx < n2e;
    const bool own = lane >= 3 && lane < 3 + kMarchCols && gx < n2e;
    const unsigned gxc = (unsigned)clampi(gx, 0, n2e - 1);
    const int lo = 2;
    const int hiz = (MODE == kFwd) ? n1e - 3 : n1e - 4;
    const int hix = (MODE == kFwd) ? n2e - 3 : n2e - 4;
    const bool colbox = gx >= lo && gx <= hix;
    const bool colint = gx >= npml && gx < npml + n2;
    const float dt = S->dt, dtdx = S->dtdx, mdt = -dt;
    const int ofs = S->ofs;
    const int ic = nt - 1 - it;

    // Buffers and byte offsets: the state (8 planes, read half px, pz, vx,
    // vz at +0, pe, 2pe, 3pe floats, write half likewise), kappa, and the
    // snapshot planes of the step (written at it by the back-propagation,
    // read at nt-1-it by the re-injection; row z of the interior at
    // (z-npml)*spitch, column gx at gx-sxb).  Lane offsets are columns.
    const unsigned pe = (unsigned)(n1e * ldx);
    const Rsrc rs_st = make_rsrc(S->px[0]), rs_ka = make_rsrc(S->kappa);
    const unsigned rd = P.cur ? 16 * pe : 0u, wr = P.cur ? 0u : 16 * pe;   // bytes
    const unsigned ldx4 = 4u * (unsigned)ldx, pe4 = 4u * pe;
    const unsigned vcol = 4u * gxc;
    const int xlo1 = S->xlo1, xhi1 = S->xhi1, xlo2 = S->xlo2, xhi2 = S->xhi2;
    const int spitch = S->spitch, sxb = S->sxb;
    const unsigned sp4 = 4u * (unsigned)spitch;
    const bool xcol = gx >= xlo2 && gx <= xhi2;
    const unsigned sxs = 4u * (unsigned)(clampi(gx, xlo2, xhi2) - sxb);            // extended-source read column
    const unsigned sxi = 4u * (unsigned)(clampi(gx, npml, npml + n2 - 1) - sxb);   // interior read column
    const unsigned sxo = 4u * (unsigned)(gx - sxb);                                // snapshot store column (interior lanes)
    const size_t plane = (size_t)n1 * spitch;
    const size_t tsn = (size_t)(MODE == kBack ? it : ic) * plane;
    const Rsrc rs_sp = make_rsrc(S->sp + tsn), rs_sx = make_rsrc(S->svx + tsn), rs_sz = make_rsrc(S->svz + tsn);

    // sponge profiles: one block d1a, d1a1, d1b, bw1 [n1e] then d2a, d2a1, d2b, bw2 [n2e]
    const GLB float *const prof = glb(S->d1a);
    const unsigned a1e = (unsigned)n1e;
    const float c_d2a = ld(prof + 4 * a1e, gxc);
    const float c_d2a1 = ld(prof + (4 * a1e + n2e), gxc);
    const float c_d2b = ld(prof + (4 * a1e + 2 * n2e), gxc);
    const float c_bw2 = ld(prof + (4 * a1e + 3 * n2e), gxc);
    const GLB int *const inj_row = glb(S->inj_row);
    const bool inj_p = MODE != kReinject && inj_row != nullptr && !S->vinj;   // Hicks injection into px
    const bool inj_v = MODE == kFwd && inj_row != nullptr && S->vinj;         // vz source
    const GLB int *const band_row = glb(S->band_row);
    const bool rec_band = MODE != kBack && band_row != nullptr;
    const Rsrc rs_band = make_rsrc(S->pband + (size_t)(it & (kBandSlots - 1)) * S->band_stride);
    GLB float *const wf = glb(S->wf);

    // the rows: a full segment of H, or one of the short segments of the
    // injection band; each a fully unrolled walk (queue slots are registers)
    auto body = [&](auto hc) {
        constexpr int HH = decltype(hc)::value;
        constexpr int NK = HH + 6;                   // iterations: vz rows z0-3 .. z0+HH+2
        // What iteration k needs of its rows sits in lane k: the rows (vz of
        // zr, p of zp, velocities of zv), their clamped byte offsets, profile
        // values and row-map entries; each per-row condition is bit k of a
        // wavefront mask.  The walk takes them with constant lane and bit
        // indices, so a row costs few scalar instructions.
        const int zr_l = up ? z0 + HH + 2 - lane : z0 - 3 + lane;
        const int zp_l = up ? zr_l + 2 : zr_l - 1, zv_l = up ? zr_l + 3 : zr_l - 3;
        const bool pin = zp_l >= 0 && zp_l < n1e, vin = zv_l >= 0 && zv_l < n1e;
        const unsigned zpc = (unsigned)clampi(zp_l, 0, n1e - 1), zvc = (unsigned)clampi(zv_l, 0, n1e - 1);
        const unsigned o_vz = rd + 3 * pe4 + (unsigned)clampi(zr_l, 0, n1e - 1) * ldx4;   // vz load
        const unsigned o_p = zpc * ldx4, o_v = zvc * ldx4;                                // state rows zp, zv
        const unsigned o_sp = (unsigned)(clampi(zp_l, xlo1, xhi1) - npml) * sp4;           // extended source row zp
        const unsigned o_sv = (unsigned)(clampi(zv_l, npml, npml + n1 - 1) - npml) * sp4;  // velocity snapshot row zv
        const float l_d1a = pin ? ld(prof, zpc) : 1.0f, l_d1a1 = pin ? ld(prof + a1e, zpc) : 1.0f;
        const float l_d1b = vin ? ld(prof + 2 * a1e, zvc) : 1.0f, l_bw1 = vin ? ld(prof + 3 * a1e, zvc) : 0.0f;
        const int l_injp = inj_p && pin ? ld(inj_row, zpc) : -1;
        const int l_injv = inj_v && vin ? ld(inj_row, zvc) : -1;
        const int l_band = rec_band && pin ? ld(band_row, zpc) : -1;
        const unsigned long long mPv = lanes_where(zp_l < n1e);                      // p row exists (zp >= -1)
        const unsigned long long mPbox = lanes_where(zp_l >= lo && zp_l <= hiz);     // p row in the stencil box
        const unsigned long long mPinj = lanes_where(l_injp >= 0);                   // p row with injection terms
        const unsigned long long mPsrc = lanes_where(zp_l >= xlo1 && zp_l <= xhi1);  // p row of the extended source
        const unsigned long long mPfs = lanes_where(ofs == 1 && zp_l <= npml);       // p row zeroed (free surface)
        const unsigned long long mPst = lanes_where(zp_l >= z0 && zp_l < z0 + HH);   // p row of the tile (stored)
        const unsigned long long mPint = lanes_where(zp_l >= npml && zp_l < npml + n1);
        const unsigned long long mPband = lanes_where(l_band >= 0);
        const unsigned long long mVv = lanes_where(zv_l < n1e);
        const unsigned long long mVbox = lanes_where(zv_l >= lo && zv_l <= hiz);
        const unsigned long long mVint = lanes_where(zv_l >= npml && zv_l < npml + n1);
        const unsigned long long mVinj = lanes_where(l_injv >= 0);
        const unsigned long long mVfs0 = lanes_where(ofs == 1 && zv_l == npml - 1);  // vz row written by the mirror
        const unsigned long long mVfs1 = lanes_where(ofs == 1 && zv_l == npml);      // vz row mirrored upwards
        // ---- the loads of iteration k (vz of row zr; vx, px, pz, kappa and
        // the extended source of row zp; the velocity snapshots of row zv)
        // into slot k % D, issued D iterations before they are used.
        // No branch around a load (the wait counts stay exact), and no mask on
        // a loaded value: rows and columns outside the grid read a clamped
        // in-grid address, and nothing stored depends on what they hold (a point
        // in the stencil box reads neighbours inside the grid; rows and lanes
        // outside it are never stored).
        float Lvz[D] = {}, Lvx[D] = {}, Lpx[D] = {}, Lpz[D] = {}, Lka[D] = {}, Lsp[D] = {}, Lsx[D] = {}, Lsz[D] = {};
        auto issue = [&](int k) {
            const int s = k % D;
            Lvz[s] = bld(rs_st, vcol, lane_u(o_vz, k));
            if (k >= 3) {
                const unsigned r = lane_u(o_p, k);
                Lvx[s] = bld(rs_st, vcol, rd + 2 * pe4 + r);
                Lpx[s] = bld(rs_st, vcol, rd + r);
                Lpz[s] = bld(rs_st, vcol, rd + pe4 + r);
                Lka[s] = bld(rs_ka, vcol, r);
                if (MODE == kReinject) Lsp[s] = bld_nt(rs_sp, sxs, lane_u(o_sp, k));
            }
            if (MODE == kReinject && k >= 6) {
                const unsigned r = lane_u(o_sv, k);
                Lsx[s] = bld_nt(rs_sx, sxi, r);
                Lsz[s] = bld_nt(rs_sz, sxi, r);
            }
        };
    #pragma unroll
        for (int k = 0; k < D; ++k) issue(k);


        const float a0 = 1.125f;
        const float a1 = -(1.0f / 24.0f);
        // queues as named registers, newest first (z0: vz loaded at k, z1 at
        // k-1, ...; x0.. vx and p0q.. p likewise), shifted each iteration (in
        // the unrolled walk a shift is a renaming).  Walking down the rows in
        // a queue descend from the newest; walking up they ascend, and each
        // operand is selected by direction between two named registers.
        float z0q = 0.0f, z1q = 0.0f, z2q = 0.0f, z3q = 0.0f, x0q = 0.0f, x1q = 0.0f, x2q = 0.0f;
        float p0q = 0.0f, p1q = 0.0f, p2q = 0.0f, p3q = 0.0f;

    #pragma unroll
        for (int k = 0; k < NK; ++k) {
            const int s = k % D;
            // this iteration's values out of slot s, then the slot's next loads
            const float r_vz = Lvz[s], r_vx = Lvx[s], r_px = Lpx[s], r_pz = Lpz[s], r_ka = Lka[s], r_sp = Lsp[s];
            const float r_sx = Lsx[s], r_sz = Lsz[s];
            if (k + D < NK) issue(k + D);
            z3q = z2q; z2q = z1q; z1q = z0q; z0q = r_vz;     // old vz of row zr
            // ---- pressure of row zp = zr - 1 (zr + 2 walking up: the tile rows and
            // the ring), k >= 3; the queue holds vz of rows zr .. zr -/+ 3
            if (k >= 3) {
                float p = 0.0f;
                if (bit(mPv, k)) {                            // zp >= z0 - 1 >= -1; row -1 is outside
                    const float vx = r_vx, kap = r_ka;
                    float pxv = r_px, pzv = r_pz;
                    // every lane shifts (inactive lanes would feed the bound value)
                    const float vxm1 = from_left(vx), vxm2 = from_left(vxm1), vxp1 = from_right(vx);
                    if (bit(mPbox, k)) {
                        const float der2 = a0 * (vx - vxm1) + a1 * (vxp1 - vxm2);
                        const float vz0 = up ? z2q : z1q;     // vz(zp)
                        const float vzm1 = up ? z1q : z2q;    // vz(zp-1)
                        const float vzp1 = up ? z3q : z0q;    // vz(zp+1)
                        const float vzm2 = up ? z0q : z3q;    // vz(zp-2)
                        const float der1 = a0 * (vz0 - vzm1) + a1 * (vzp1 - vzm2);
                        const float tk = dtdx * kap;
                        const float kw2 = tk * c_d2a1;
                        const float kw1 = tk * lane_f(l_d1a1, k);
                        const float d1a = lane_f(l_d1a, k);
                        float nx, nz;
                        if (MODE == kBack) {
                            nx = c_d2a * pxv - kw2 * der2;
                            nz = d1a * pzv - kw1 * der1;
                        } else {
                            nx = c_d2a * pxv + kw2 * der2;
                            nz = d1a * pzv + kw1 * der1;
                        }
                        pxv = colbox ? nx : pxv;
                        pzv = colbox ? nz : pzv;
                    }
                    if (inj_p && bit(mPinj, k)) {
                        const int injq = lane_i(l_injp, k);
                        if (colok) {                                  // :1042 (subtracted), :175 (added)
                            const GLB int *ptr = glb(S->inj_ptr) + injq * (n2e + 1);
                            pxv = inject_terms<MODE == kBack ? -1 : 1>(pxv, kap, ptr[gx], ptr[gx + 1], glb(S->inj_r),
                                                                       glb(S->inj_cdt), glb(S->inj_amp), S->amp_stride,
                                                                       it * S->amp_tstride);
                        }
                    }
                    if (MODE == kReinject && bit(mPsrc, k))
                        pxv = xcol ? pxv + (mdt * r_sp) * kap : pxv;   // ss*dt*sp*kappa, ss = -1 (:1123)
                    p = pxv + pzv;
                    if (bit(mPfs, k)) p = 0.0f;                  // p(1:npml+1,:) = 0
                    if (bit(mPst, k) && own) {
                        const unsigned r = lane_u(o_p, k);
                        bst(rs_st, vcol, wr + r, pxv);
                        bst(rs_st, vcol, wr + pe4 + r, pzv);
                        if (MODE != kBack) {
                            if (rec_band && bit(mPband, k))
                                bst(rs_band, vcol, (unsigned)lane_i(l_band, k) * ldx4, p);
                            if (wf != nullptr && bit(mPint, k) && colint) {
                                const int zp = up ? z0 + HH + 4 - k : z0 - 4 + k;
                                (wf + (size_t)it * ((size_t)n1 * n2))[(gx - npml) * n1 + (zp - npml)] = p;
                            }
                        } else if (bit(mPint, k) && colint) {
                            const int zp = up ? z0 + HH + 4 - k : z0 - 4 + k;
                            bst_nt(rs_sp, sxo, (unsigned)(zp - npml) * sp4, p);   // sp(:,:,it) (:1060-1064)
                        }
                    }
                }
                p3q = p2q; p2q = p1q; p1q = p0q; p0q = p;
                x2q = x1q; x1q = x0q; x0q = r_vx;
            }
            // ---- velocities of row zv = zr - 3 (zr + 3 walking up: the tile
            // rows), k >= 6; the queue holds p of rows zp .. zp -/+ 3
            if (k >= 6) {
                if (bit(mVv, k)) {
                    const bool rowi = bit(mVint, k);
                    const float p0 = up ? p1q : p2q;     // p(zv)
                    const float pr1 = from_right(p0), pr2 = from_right(pr1), pl1 = from_left(p0);
                    float vxv = up ? x1q : x2q, vzv = z3q;
                    if (bit(mVbox, k)) {
                        const float der2 = a0 * (pr1 - p0) + a1 * (pr2 - pl1);
                        const float pm1 = up ? p0q : p3q;    // p(zv-1)
                        const float pp1 = up ? p2q : p1q;    // p(zv+1)
                        const float pp2 = up ? p3q : p0q;    // p(zv+2)
                        const float der1 = a0 * (pp1 - p0) + a1 * (pp2 - pm1);
                        const float d1b = lane_f(l_d1b, k), bw1 = lane_f(l_bw1, k);
                        float nx, nz;
                        if (MODE == kBack) {
                            nx = c_d2b * vxv - c_bw2 * der2;
                            nz = d1b * vzv - bw1 * der1;
                        } else {
                            nx = c_d2b * vxv + c_bw2 * der2;
                            nz = d1b * vzv + bw1 * der1;
                        }
                        vxv = colbox ? nx : vxv;
                        vzv = colbox ? nz : vzv;
                    }
                    if (inj_v && bit(mVinj, k)) {
                        const int q = lane_i(l_injv, k);
                        if (colok) {
                            const GLB int *ptr = glb(S->inj_ptr) + q * (n2e + 1);
                            for (int m = ptr[gx]; m < ptr[gx + 1]; ++m)
                                vzv = vzv + glb(S->inj_cdt)[m] * glb(S->inj_amp)[glb(S->inj_r)[m] * S->amp_stride + it * S->amp_tstride];   // * bue (= 1), :252
                        }
                    }
                    if (MODE == kReinject && rowi) {
                        vxv = colint ? vxv + mdt * r_sx : vxv;         // * bue (= 1), :1177
                        vzv = colint ? vzv + mdt * r_sz : vzv;         // :1184
                    }
                    if (own) {
                        const unsigned r = lane_u(o_v, k);
                        bst(rs_st, vcol, wr + 2 * pe4 + r, vxv);
                        if (!bit(mVfs0, k)) {                          // vz(npml,:) is written by the mirror below
                            bst(rs_st, vcol, wr + 3 * pe4 + r, vzv);
                            if (bit(mVfs1, k)) bst(rs_st, vcol, wr + 3 * pe4 + r - ldx4, vzv);   // vz(npml,:) = vz(npml+1,:)
                        }
                        if (MODE == kBack && rowi && colint) {         // svx, svz(:,:,it) (:1082-1087)
                            const unsigned rs = lane_u(o_sv, k);
                            bst_nt(rs_sx, sxo, rs, vxv);
                            bst_nt(rs_sz, sxo, rs, vzv);
                        }
                    }
                }
            }
        }
    };
    if (MODE != kBack || nrows == H) body(std::integral_constant<int, H>{});   // only the back-propagation injects
    else body(std::integral_constant<int, kMarchBandRows>{});                 // receiver rows
}

// ---------------------------------------------------------------------------
// Two-step column march of the re-injection: steps it (A) and it+1 (B) in one
// walk, so the wave state crosses HBM once per two steps.  A wavefront owns
// kMarch2Cols = 52 columns x0 .. x0+51 of the rows z0 .. z0+H-1 (H = the
// tile entry's rows); its lanes hold columns x0-6 .. x0+57: step A's pressure
// is valid on lanes 2..62 and its velocities on 3..60, step B's pressure on
// 5..59 and its velocities on 6..57, the owned lanes.  Iteration k loads vz
// of row r0 (walking down r0 = z0-6+k) and the other fields of row r0-1, and
// forms A's pressure of row r0-1, A's velocities of row r0-3, B's pressure of
// row r0-4 and B's velocities of row r0-6 (walking up: r0 = z0+H+5-k and rows
// r0+2, r0+3, r0+5, r0+6: each stencil takes its rows on the side it needs).
// Step A's fields stay in register queues until step B has used them; only
// B's state is stored, plus both steps' record band rows and wavefield (a
// shot whose last step is step A, in a batch of shots with different nt,
// stores A's state from the same walk and skips B).  The
// operations of each point are those of k_march, in the same order, so the
// results are bit-identical to two one-step launches.  The vz free-surface
// mirror of step A (vz(npml-1) = vz(npml)) must reach B's pressure at row
// npml-2 before it is formed, which a walk up guarantees, so a segment whose
// A velocity rows reach rows npml-1 or npml walks up (and the host keeps a
// segment's last A velocity row off npml-1, so such a segment holds both).
// The loop runs four iterations per trip (the queue shifts are renamings
// within a trip); what an iteration needs of its rows is in lane k of the
// tables, its per-row conditions in the bits of a flag word in lane k.
// ---------------------------------------------------------------------------
// per-row conditions of a pressure row (P*) and a velocity row (V*); B's rows
// of iteration k are A's rows of iteration k-3, so one table serves both
enum March2Flag : unsigned {
    kP = 1u << 0, kPBox = 1u << 1, kPSrc = 1u << 2, kPFs = 1u << 3, kPOwn = 1u << 4, kPInt = 1u << 5, kPBand = 1u << 6,
    kV = 1u << 7, kVBox = 1u << 8, kVInt = 1u << 9, kVFs0 = 1u << 10, kVFs1 = 1u << 11, kVOwn = 1u << 12
};

// Six waves per SIMD (75 VGPRs with one row of loads ahead): every wavefront
// of a 24-shot launch at 201x801 (4800) is resident at once.
__global__ __launch_bounds__(kBlock, 6) void k_march2_reinject(StepParams P)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int2 t = P.tiles[blockIdx.x];
    const ShotDev *__restrict__ S = P.shots + t.x;
    const int it = P.it;
    const int nt = S->nt;
    if (it >= nt) return;
    // the shot's last step when its nt is odd (a batch's shots may differ in
    // nt): step A alone, stored from A's rows as the one-step march stores
    const bool two = it + 1 < nt;
    const int n1e = S->n1e, n2e = S->n2e, ldx = S->ldx, npml = S->npml, n1 = S->n1, n2 = S->n2;
    const int nstrip = (n2e + kMarch2Cols - 1) / kMarch2Cols;
    const int z0 = t.y >> 16, H = (t.y >> 8) & 255, strip = 4 * (t.y & 127) + wv;
    const int ofs = S->ofs;
    const bool up = ((t.y >> 7) & 1) || (ofs == 1 && z0 - 3 <= npml && npml - 1 <= z0 + H + 2);
    if (strip >= nstrip) return;                 // whole wavefront: no lane of it takes part in a shift
    const int x0 = strip * kMarch2Cols;
    const int gx = x0 - 6 + lane;
    const bool own = lane >= 6 && lane < 6 + kMarch2Cols && gx < n2e;
    const unsigned gxc = (unsigned)clampi(gx, 0, n2e - 1);
    const int lo = 2, hiz = n1e - 4, hix = n2e - 4;
    const bool colbox = gx >= lo && gx <= hix;
    const bool colint = gx >= npml && gx < npml + n2;
    const float dtdx = S->dtdx, mdt = -S->dt;

    const unsigned pe = (unsigned)(n1e * ldx);
    const Rsrc rs_st = make_rsrc(S->px[0]), rs_ka = make_rsrc(S->kappa);
    const unsigned rd = P.cur ? 16 * pe : 0u, wr = P.cur ? 0u : 16 * pe;   // bytes
    const unsigned ldx4 = 4u * (unsigned)ldx, pe4 = 4u * pe;
    const unsigned vcol = 4u * gxc;
    const int xlo1 = S->xlo1, xhi1 = S->xhi1, xlo2 = S->xlo2, xhi2 = S->xhi2;
    const int spitch = S->spitch, sxb = S->sxb;
    const unsigned sp4 = 4u * (unsigned)spitch;
    const bool xcol = gx >= xlo2 && gx <= xhi2;
    const unsigned sxs = 4u * (unsigned)(clampi(gx, xlo2, xhi2) - sxb);
    const unsigned sxi = 4u * (unsigned)(clampi(gx, npml, npml + n2 - 1) - sxb);
    const size_t plane = (size_t)n1 * spitch;
    // snapshot planes of the two steps; for a lone step A, B's loads (issued
    // unconditionally, their values unused) read A's plane, never a plane
    // before the shot's store
    const size_t tA = (size_t)(nt - 1 - it) * plane, tB = two ? (size_t)(nt - 2 - it) * plane : tA;
    const Rsrc rs_spA = make_rsrc(S->sp + tA), rs_sxA = make_rsrc(S->svx + tA), rs_szA = make_rsrc(S->svz + tA);
    const Rsrc rs_spB = make_rsrc(S->sp + tB), rs_sxB = make_rsrc(S->svx + tB), rs_szB = make_rsrc(S->svz + tB);

    const GLB float *const prof = glb(S->d1a);
    const unsigned a1e = (unsigned)n1e;
    const float c_d2a = ld(prof + 4 * a1e, gxc);
    const float c_d2a1 = ld(prof + (4 * a1e + n2e), gxc);
    const float c_d2b = ld(prof + (4 * a1e + 2 * n2e), gxc);
    const float c_bw2 = ld(prof + (4 * a1e + 3 * n2e), gxc);
    const GLB int *const band_row = glb(S->band_row);
    const bool rec_band = band_row != nullptr;
    const Rsrc rs_bandA = make_rsrc(S->pband + (size_t)(it & (kBandSlots - 1)) * S->band_stride);
    const Rsrc rs_bandB = make_rsrc(S->pband + (size_t)((it + 1) & (kBandSlots - 1)) * S->band_stride);
    GLB float *const wf = glb(S->wf);

    // lane k: the rows of iteration k (A's pressure row ap, A's velocity row
    // av; B's rows of iteration k are those of iteration k-3), their clamped
    // byte offsets, profile values, band slots and condition flags
    const int NK = H + 12;
    const int r0_l = up ? z0 + H + 5 - lane : z0 - 6 + lane;
    const int ap_l = up ? r0_l + 2 : r0_l - 1, av_l = up ? r0_l + 3 : r0_l - 3;
    const unsigned apc = (unsigned)clampi(ap_l, 0, n1e - 1), avc = (unsigned)clampi(av_l, 0, n1e - 1);
    const bool apin = ap_l >= 0 && ap_l < n1e, avin = av_l >= 0 && av_l < n1e;
    const unsigned o_vz = rd + 3 * pe4 + (unsigned)clampi(r0_l, 0, n1e - 1) * ldx4;
    const unsigned o_ap = apc * ldx4, o_av = avc * ldx4;
    const unsigned o_sp = (unsigned)(clampi(ap_l, xlo1, xhi1) - npml) * sp4;
    const unsigned o_sv = (unsigned)(clampi(av_l, npml, npml + n1 - 1) - npml) * sp4;
    const float l_d1a = apin ? ld(prof, apc) : 1.0f, l_d1a1 = apin ? ld(prof + a1e, apc) : 1.0f;
    const float l_d1b = avin ? ld(prof + 2 * a1e, avc) : 1.0f, l_bw1 = avin ? ld(prof + 3 * a1e, avc) : 0.0f;
    const int l_band = rec_band && apin ? ld(band_row, apc) : -1;
    auto fl = [](bool c, unsigned f) { return c ? f : 0u; };
    const unsigned l_flags =
        fl(ap_l < n1e, kP) | fl(ap_l >= lo && ap_l <= hiz, kPBox) | fl(ap_l >= xlo1 && ap_l <= xhi1, kPSrc) |
        fl(ofs == 1 && ap_l <= npml, kPFs) | fl(ap_l >= z0 && ap_l < z0 + H, kPOwn) |
        fl(ap_l >= npml && ap_l < npml + n1, kPInt) | fl(l_band >= 0, kPBand) |
        fl(av_l < n1e, kV) | fl(av_l >= lo && av_l <= hiz, kVBox) | fl(av_l >= npml && av_l < npml + n1, kVInt) |
        fl(ofs == 1 && av_l == npml - 1, kVFs0) | fl(ofs == 1 && av_l == npml, kVFs1) |
        fl(av_l >= z0 && av_l < z0 + H, kVOwn);

    const float a0 = 1.125f;
    const float a1 = -(1.0f / 24.0f);
    // rows of loads in flight ahead: one (six waves per SIMD) measured 1 %
    // faster than two (96 VGPRs, five waves) at 201x801 with 24 shots
    constexpr int D = 1;
    constexpr int NS = 4;                        // load slots (k mod 4), one trip's iterations
    float Lvz[NS] = {}, Lvx[NS] = {}, Lpx[NS] = {}, Lpz[NS] = {}, Lka[NS] = {};
    float LspA[NS] = {}, LsxA[NS] = {}, LszA[NS] = {}, LspB[NS] = {}, LsxB[NS] = {}, LszB[NS] = {};
    // every load of iteration k, unconditionally (rows outside the walk read
    // a clamped in-grid row; nothing stored depends on them)
    auto issue = [&](int s, int k) {
        const int kk = k < NK ? k : NK - 1;
        Lvz[s] = bld(rs_st, vcol, lane_u(o_vz, kk));
        const unsigned r = lane_u(o_ap, kk);
        Lvx[s] = bld(rs_st, vcol, rd + 2 * pe4 + r);
        Lpx[s] = bld(rs_st, vcol, rd + r);
        Lpz[s] = bld(rs_st, vcol, rd + pe4 + r);
        Lka[s] = bld(rs_ka, vcol, r);
        const int kb = kk >= 3 ? kk - 3 : 0;
        LspA[s] = bld_nt(rs_spA, sxs, lane_u(o_sp, kk));
        const unsigned ra = lane_u(o_sv, kk);
        LsxA[s] = bld_nt(rs_sxA, sxi, ra);
        LszA[s] = bld_nt(rs_szA, sxi, ra);
        LspB[s] = bld_nt(rs_spB, sxs, lane_u(o_sp, kb));
        const unsigned rb = lane_u(o_sv, kb);
        LsxB[s] = bld_nt(rs_sxB, sxi, rb);
        LszB[s] = bld_nt(rs_szB, sxi, rb);
    };
#pragma unroll
    for (int k = 0; k < D; ++k) issue(k, k);

    // queues, newest first: S0 vz (z), S0 vx of A's pressure rows (x), A's
    // pressure (p), A's px, pz and kappa (ax, az, ka), A's velocities (bx,
    // bz), B's pressure (q)
    float z0q = 0.f, z1q = 0.f, z2q = 0.f, z3q = 0.f, x0q = 0.f, x1q = 0.f, x2q = 0.f;
    float p0q = 0.f, p1q = 0.f, p2q = 0.f, p3q = 0.f;
    float ax0 = 0.f, ax1 = 0.f, ax2 = 0.f, ax3 = 0.f, az0 = 0.f, az1 = 0.f, az2 = 0.f, az3 = 0.f;
    float ka0 = 0.f, ka1 = 0.f, ka2 = 0.f, ka3 = 0.f;
    float bx0 = 0.f, bx1 = 0.f, bx2 = 0.f, bx3 = 0.f, bz0 = 0.f, bz1 = 0.f, bz2 = 0.f, bz3 = 0.f;
    float q0 = 0.f, q1 = 0.f, q2 = 0.f, q3 = 0.f;

    for (int k0 = 0; k0 < NK; k0 += NS) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int k = k0 + j;
            const float r_vz = Lvz[j], r_vx = Lvx[j], r_px = Lpx[j], r_pz = Lpz[j], r_ka = Lka[j];
            const float r_spA = LspA[j], r_sxA = LsxA[j], r_szA = LszA[j];
            const float r_spB = LspB[j], r_sxB = LsxB[j], r_szB = LszB[j];
            issue((j + D) % NS, k + D);
            // the walk is the kernel's last work: ending here, rather than
            // leaving the trip, keeps the trip's end the loop's only back-edge
            if (k >= NK) return;
            const unsigned f = (unsigned)__builtin_amdgcn_readlane((int)l_flags, k);
            const unsigned fb = (unsigned)__builtin_amdgcn_readlane((int)l_flags, k >= 3 ? k - 3 : 0);
            z3q = z2q; z2q = z1q; z1q = z0q; z0q = r_vz;
            // The four phases run at every iteration, so that every queue
            // shifts once per iteration and a trip's shifts are renamings.
            // A phase's results are valid from iteration 3 (A pressure), 6
            // (A velocities), 9 (B pressure) and 12 (B velocities): each
            // reads queue entries of its last four iterations only, so values
            // formed earlier leave the queues unread, and each store is kept
            // to owned rows, which no iteration before those reaches.
            // ---- step A, pressure
            {
                float pxv = r_px, pzv = r_pz, p = 0.0f;
                const float kap = r_ka;
                if (f & kP) {
                    const float vx = r_vx;
                    const float vxm1 = from_left(vx), vxm2 = from_left(vxm1), vxp1 = from_right(vx);
                    if (f & kPBox) {
                        const float der2 = a0 * (vx - vxm1) + a1 * (vxp1 - vxm2);
                        const float vz0 = up ? z2q : z1q, vzm1 = up ? z1q : z2q;
                        const float vzp1 = up ? z3q : z0q, vzm2 = up ? z0q : z3q;
                        const float der1 = a0 * (vz0 - vzm1) + a1 * (vzp1 - vzm2);
                        const float tk = dtdx * kap;
                        const float kw2 = tk * c_d2a1;
                        const float kw1 = tk * lane_f(l_d1a1, k);
                        const float d1a = lane_f(l_d1a, k);
                        const float nx = c_d2a * pxv + kw2 * der2;
                        const float nz = d1a * pzv + kw1 * der1;
                        pxv = colbox ? nx : pxv;
                        pzv = colbox ? nz : pzv;
                    }
                    if (f & kPSrc) pxv = xcol ? pxv + (mdt * r_spA) * kap : pxv;
                    p = pxv + pzv;
                    if (f & kPFs) p = 0.0f;
                    if ((f & kPOwn) && own) {
                        if (!two) {
                            const unsigned r = lane_u(o_ap, k);
                            bst(rs_st, vcol, wr + r, pxv);
                            bst(rs_st, vcol, wr + pe4 + r, pzv);
                        }
                        if ((f & kPBand) && rec_band) bst(rs_bandA, vcol, (unsigned)lane_i(l_band, k) * ldx4, p);
                        if (wf != nullptr && (f & kPInt) && colint)
                            (wf + (size_t)it * ((size_t)n1 * n2))[(gx - npml) * n1 + ((up ? z0 + H + 7 - k : z0 - 7 + k) - npml)] = p;
                    }
                }
                p3q = p2q; p2q = p1q; p1q = p0q; p0q = p;
                x2q = x1q; x1q = x0q; x0q = r_vx;
                ax3 = ax2; ax2 = ax1; ax1 = ax0; ax0 = pxv;
                az3 = az2; az2 = az1; az1 = az0; az0 = pzv;
                ka3 = ka2; ka2 = ka1; ka1 = ka0; ka0 = kap;
            }
            // ---- step A, velocities
            {
                float vxv = up ? x1q : x2q, vzv = z3q;
                if (f & kV) {
                    const float p0 = up ? p1q : p2q;
                    const float pr1 = from_right(p0), pr2 = from_right(pr1), pl1 = from_left(p0);
                    if (f & kVBox) {
                        const float der2 = a0 * (pr1 - p0) + a1 * (pr2 - pl1);
                        const float pm1 = up ? p0q : p3q, pp1 = up ? p2q : p1q, pp2 = up ? p3q : p0q;
                        const float der1 = a0 * (pp1 - p0) + a1 * (pp2 - pm1);
                        const float d1b = lane_f(l_d1b, k), bw1 = lane_f(l_bw1, k);
                        const float nx = c_d2b * vxv + c_bw2 * der2;
                        const float nz = d1b * vzv + bw1 * der1;
                        vxv = colbox ? nx : vxv;
                        vzv = colbox ? nz : vzv;
                    }
                    if (f & kVInt) {
                        vxv = colint ? vxv + mdt * r_sxA : vxv;
                        vzv = colint ? vzv + mdt * r_szA : vzv;
                    }
                    if (!two && (f & kVOwn) && own) {
                        const unsigned r = lane_u(o_av, k);
                        bst(rs_st, vcol, wr + 2 * pe4 + r, vxv);
                        if (!(f & kVFs0)) {
                            bst(rs_st, vcol, wr + 3 * pe4 + r, vzv);
                            if (f & kVFs1) bst(rs_st, vcol, wr + 3 * pe4 + r - ldx4, vzv);
                        }
                    }
                }
                bx3 = bx2; bx2 = bx1; bx1 = bx0; bx0 = vxv;
                bz3 = bz2; bz2 = bz1; bz1 = bz0; bz0 = vzv;
                // vz(npml-1) = vz(npml) after step A; walking up, row npml
                // came one iteration before row npml-1
                if (f & kVFs0) bz0 = bz1;
            }
            // ---- step B, pressure (stored when the shot takes step B)
            {
                float pxv = ax3, pzv = az3, p = 0.0f;
                const float kap = ka3;
                if (fb & kP) {
                    const float vx = up ? bx2 : bx1;
                    const float vxm1 = from_left(vx), vxm2 = from_left(vxm1), vxp1 = from_right(vx);
                    if (fb & kPBox) {
                        const float der2 = a0 * (vx - vxm1) + a1 * (vxp1 - vxm2);
                        const float vz0 = up ? bz2 : bz1, vzm1 = up ? bz1 : bz2;
                        const float vzp1 = up ? bz3 : bz0, vzm2 = up ? bz0 : bz3;
                        const float der1 = a0 * (vz0 - vzm1) + a1 * (vzp1 - vzm2);
                        const float tk = dtdx * kap;
                        const float kw2 = tk * c_d2a1;
                        const float kw1 = tk * lane_f(l_d1a1, k - 3);
                        const float d1a = lane_f(l_d1a, k - 3);
                        const float nx = c_d2a * pxv + kw2 * der2;
                        const float nz = d1a * pzv + kw1 * der1;
                        pxv = colbox ? nx : pxv;
                        pzv = colbox ? nz : pzv;
                    }
                    if (fb & kPSrc) pxv = xcol ? pxv + (mdt * r_spB) * kap : pxv;
                    p = pxv + pzv;
                    if (fb & kPFs) p = 0.0f;
                    if (two && (fb & kPOwn) && own) {
                        const unsigned r = lane_u(o_ap, k - 3);
                        bst(rs_st, vcol, wr + r, pxv);
                        bst(rs_st, vcol, wr + pe4 + r, pzv);
                        if ((fb & kPBand) && rec_band) bst(rs_bandB, vcol, (unsigned)lane_i(l_band, k - 3) * ldx4, p);
                        if (wf != nullptr && (fb & kPInt) && colint)
                            (wf + (size_t)(it + 1) * ((size_t)n1 * n2))[(gx - npml) * n1 + ((up ? z0 + H + 10 - k : z0 - 10 + k) - npml)] = p;
                    }
                }
                q3 = q2; q2 = q1; q1 = q0; q0 = p;
            }
            // ---- step B, velocities (B's velocity rows are the owned rows
            // from iteration 12 on)
            if (fb & kV) {
                const float p0 = up ? q1 : q2;
                const float pr1 = from_right(p0), pr2 = from_right(pr1), pl1 = from_left(p0);
                float vxv = bx3, vzv = bz3;
                if (fb & kVBox) {
                    const float der2 = a0 * (pr1 - p0) + a1 * (pr2 - pl1);
                    const float pm1 = up ? q0 : q3, pp1 = up ? q2 : q1, pp2 = up ? q3 : q0;
                    const float der1 = a0 * (pp1 - p0) + a1 * (pp2 - pm1);
                    const float d1b = lane_f(l_d1b, k - 3), bw1 = lane_f(l_bw1, k - 3);
                    const float nx = c_d2b * vxv + c_bw2 * der2;
                    const float nz = d1b * vzv + bw1 * der1;
                    vxv = colbox ? nx : vxv;
                    vzv = colbox ? nz : vzv;
                }
                if (fb & kVInt) {
                    vxv = colint ? vxv + mdt * r_sxB : vxv;
                    vzv = colint ? vzv + mdt * r_szB : vzv;
                }
                if (two && (fb & kVOwn) && own) {
                    const unsigned r = lane_u(o_av, k - 3);
                    bst(rs_st, vcol, wr + 2 * pe4 + r, vxv);
                    if (!(fb & kVFs0)) {
                        bst(rs_st, vcol, wr + 3 * pe4 + r, vzv);
                        if (fb & kVFs1) bst(rs_st, vcol, wr + 3 * pe4 + r - ldx4, vzv);
                    }
                }
            }
        }
    }
}

void launch_march2(const StepParams &P, hipStream_t st)
{
    hipLaunchKernelGGL(k_march2_reinject, dim3(P.ntiles), dim3(kBlock), 0, st, P);
}

// Row-segment heights of the column march the autotuner may pick.
#define DAW_MARCH_ROWS(X) X(8) X(12) X(16) X(24) X(32)

template <int MODE>
static void launch_march_mode(const StepParams &P, int h, hipStream_t st)
{
    const int grid = P.ntiles + P.nrec_blocks;
#define DAW_MARCH(R)                                                                                           \
    if (h == R) {                                                                                              \
        hipLaunchKernelGGL((k_march<MODE, R>), dim3(grid), dim3(kBlock), 0, st, P);                            \
        return;                                                                                                \
    }
    DAW_MARCH_ROWS(DAW_MARCH)
#undef DAW_MARCH
}

bool march_rows_ok(int h)
{
#define DAW_OK_MARCH(R) if (h == R) return true;
    DAW_MARCH_ROWS(DAW_OK_MARCH)
#undef DAW_OK_MARCH
    return false;
}

void launch_march(int mode, const StepParams &P, int h, hipStream_t st)
{
    if (P.ntiles + P.nrec_blocks == 0) return;
    if (mode == kFwd) launch_march_mode<kFwd>(P, h, st);
    else if (mode == kBack) launch_march_mode<kBack>(P, h, st);
    else launch_march_mode<kReinject>(P, h, st);
}

// ---------------------------------------------------------------------------
// Batched record: the Hicks record of groups of (shot, point) items at
// steps it0 .. it0+nsteps-1, read from the band's time slots after the steps
// ran (a step's slot is not overwritten for kBandSlots steps).  A group is
// up to 64 consecutive points of one point set of one shot whose windows
// fit a kRecRows x kRecCols block of the grid; a workgroup stages that block
// of the band for kRecSteps steps in LDS with coalesced loads, then
// wavefront w runs step w: one lane per item, the item's 441-term sum in the
// order of submodeling2.f90:214-219 and subwavefieldresample.f90:32-37 (i1
// outer, i2 inner; a term outside the item's clip window contributes
// nothing, as the record blocks' +0 does), coefficients read from the
// transposed table (neighbouring points' coefficient k side by side).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rec_group(const ShotDev *__restrict__ shots, const RecItem *__restrict__ items,
                                                  const RecGroup *__restrict__ groups, int it0, int nsteps)
{
    // rows of kRecCols + 20 floats, the block's columns from offset 10: a
    // window column ix2-10 .. ix2+10 of any item stays inside its row
    constexpr int RW = kRecCols + 20;
    __shared__ float sb[kRecSteps][kRecRows][RW];
    __shared__ unsigned soff[kRecSteps * kRecRows];
    const RecGroup G = groups[blockIdx.x];
    const int j0 = (int)blockIdx.y * kRecSteps;
    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
    const ShotDev *__restrict__ S = shots + items[G.first].shot;
    const int nt = S->nt, ldx = S->ldx;
    constexpr unsigned kNone = 0xffffffffu;
    // byte offset in the band of staged row jr = (step j, row r), or kNone
    // where the step is past the sweep or the row holds no band slot
    if (tid < kRecSteps * G.nz) {
        const int j = tid / G.nz, r = tid - j * G.nz, it = it0 + j0 + j;
        const int slot = glb(S->band_row)[G.z0 + r];
        const bool live = j0 + j < nsteps && it < nt && slot >= 0;
        soff[tid] = live ? 4u * (unsigned)((it & (kBandSlots - 1)) * S->band_stride + slot * ldx) : kNone;
    }
    __syncthreads();
    // ---- stage: wavefront w takes rows jr = w, w+4, ... of the block, its
    // lanes the columns; all of a wavefront's rows are loaded before the
    // first is stored, so their loads are in flight together (buffer loads:
    // lane offset = the row's offset from LDS + the column's)
    const Rsrc rs_band = make_rsrc(S->pband);
    constexpr int RPW = kRecSteps * kRecRows / 4, CPL = (kRecCols + 63) / 64;
    unsigned vo[CPL];
#pragma unroll
    for (int m = 0; m < CPL; ++m) vo[m] = 4u * (unsigned)(l + 64 * m < G.nx ? G.x0 + l + 64 * m : 0);
    float v[RPW][CPL];
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        const int jr = w + 4 * i;
        const unsigned o = jr < kRecSteps * G.nz ? soff[jr] : kNone;
        const unsigned ob = o == kNone ? 0u : o;
#pragma unroll
        for (int m = 0; m < CPL; ++m) v[i][m] = bld(rs_band, vo[m] + ob, 0u);
    }
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        const int jr = w + 4 * i;
        if (jr >= kRecSteps * G.nz) break;
        const int j = jr / G.nz, r = jr - j * G.nz;             // uniform
        const bool live = soff[jr] != kNone;
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
            const int c = l + 64 * m;
            if (c < G.nx) sb[j][r][10 + c] = live ? v[i][m] : 0.0f;
        }
    }
    __syncthreads();
    const int it = it0 + j0 + w;
    if (l >= G.count || j0 + w >= nsteps || it >= nt) return;
    const RecItem R = items[G.first + l];
    int lo1, hi1, lo2, hi2;
    if (R.clip == 0) {
        lo1 = -R.ix1; hi1 = S->n1e - 1 - R.ix1;
        lo2 = -R.ix2; hi2 = S->n2e - 1 - R.ix2;
    } else {
        lo1 = S->npml - R.ix1; hi1 = S->npml + S->n1 - 1 - R.ix1;
        lo2 = S->npml - R.ix2; hi2 = S->npml + S->n2 - 1 - R.ix2;
    }
    // columns of the window inside the clip range: bit q for i2 = q - 10
    unsigned m2 = 0;
#pragma unroll
    for (int q = 0; q < 21; ++q) m2 |= (q - 10 >= lo2 && q - 10 <= hi2) ? (1u << q) : 0u;
    // the group's items have consecutive coefficient columns of one table:
    // item l's coefficient k at coefT(first) + l + k*cstride
    const Rsrc rs_ct = make_rsrc(items[G.first].coefT);
    const unsigned cs4 = 4u * (unsigned)items[G.first].cstride, co = 4u * (unsigned)l;
    const int cbase = R.ix2 - 10 - G.x0 + 10;                 // column of i2 = -10 in a staged row
    float acc = 0.0f;
#pragma unroll 2
    for (int a = 0; a < 21; ++a) {
        const int i1 = a - 10;
        const unsigned mr = (i1 >= lo1 && i1 <= hi1) ? m2 : 0u;
        const float *row = &sb[w][clampi(R.ix1 + i1 - G.z0, 0, kRecRows - 1)][cbase];
        float c[21];
#pragma unroll
        for (int q = 0; q < 21; ++q) c[q] = bld(rs_ct, co + (unsigned)(a * 21 + q) * cs4, 0u);
        // a term outside the window becomes +0 (bits cleared), which adds
        // nothing: acc is never -0; no branch, so every read is in flight
#pragma unroll
        for (int q = 0; q < 21; ++q) {
            const unsigned keep = 0u - ((mr >> q) & 1u);
            acc = acc + __uint_as_float(__float_as_uint(c[q] * row[q]) & keep);
        }
    }
    R.out[R.reverse ? nt - 1 - it : it] = acc;
}

void launch_rec_batch(const ShotDev *shots, const RecItem *items, const RecGroup *groups, int ngroups, int it0,
                      int nsteps, hipStream_t st)
{
    if (ngroups <= 0 || nsteps <= 0) return;
    hipLaunchKernelGGL(k_rec_group, dim3(ngroups, (nsteps + kRecSteps - 1) / kRecSteps), dim3(256), 0, st, shots,
                       items, groups, it0, nsteps);
}

// ---------------------------------------------------------------------------
// Transposed copy of traces: out[t][r] = in[r][t] for r < nr, t < nt (32x32
// tiles through LDS, both sides coalesced).  The back-propagation injects
// from the copy, so the receivers' amplitudes of one step are contiguous.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_transpose(const float *__restrict__ in, float *__restrict__ out, int nr, int nt)
{
    __shared__ float tile[32][33];
    const int t0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, t = t0 + tx;
        if (r < nr && t < nt) tile[k][tx] = in[(size_t)r * nt + t];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int t = t0 + k, r = r0 + tx;
        if (r < nr && t < nt) out[(size_t)t * nr + r] = tile[tx][k];
    }
}

void launch_transpose(const float *in, float *out, int nr, int nt, hipStream_t st)
{
    if (nr <= 0 || nt <= 0) return;
    hipLaunchKernelGGL(k_transpose, dim3((nt + 31) / 32, (nr + 31) / 32), dim3(256), 0, st, in, out, nr, nt);
}

// Tile shapes the autotuner may pick: (rows, wavefronts across x).
#define DAW_TILE_SHAPES(X) X(8, 2) X(16, 1) X(16, 2) X(20, 1) X(24, 1) X(32, 1) X(32, 2) X(64, 1)

void launch_adjoint(bool hess, const StepParams &P, int tz, int txw, hipStream_t st)
{
    const int grid = P.ntiles + P.nrec_blocks;
    if (grid == 0) return;
#define DAW_ADJ(Z, W)                                                                                          \
    if (tz == Z && txw == W) {                                                                                 \
        if (hess) hipLaunchKernelGGL((k_adj<Z, W, true>), dim3(grid), dim3(kBlock), 0, st, P);                 \
        else hipLaunchKernelGGL((k_adj<Z, W, false>), dim3(grid), dim3(kBlock), 0, st, P);                     \
        return;                                                                                                \
    }
    DAW_TILE_SHAPES(DAW_ADJ)
#undef DAW_ADJ
}

template <int MODE, int SNAP>
static void launch_mode(const StepParams &P, int tz, int txw, hipStream_t st)
{
    const int grid = P.ntiles + P.nrec_blocks;
#define DAW_STEP(Z, W)                                                                                         \
    if (tz == Z && txw == W) {                                                                                 \
        hipLaunchKernelGGL((k_step<MODE, Z, W, SNAP>), dim3(grid), dim3(kBlock), 0, st, P);                    \
        return;                                                                                                \
    }
    DAW_TILE_SHAPES(DAW_STEP)
#undef DAW_STEP
}

bool tile_shape_ok(int tz, int txw)
{
#define DAW_OK_SHAPE(Z, W) if (tz == Z && txw == W) return true;
    DAW_TILE_SHAPES(DAW_OK_SHAPE)
#undef DAW_OK_SHAPE
    return false;
}

template <int MODE>
static void launch_snap(int snap, const StepParams &P, int tz, int txw, hipStream_t st)
{
    if (snap == kSnapBF16) launch_mode<MODE, kSnapBF16>(P, tz, txw, st);
    else if (snap == kSnapF16) launch_mode<MODE, kSnapF16>(P, tz, txw, st);
    else launch_mode<MODE, kSnapF32>(P, tz, txw, st);
}

// Shapes of the two-step kernel: (core rows, 64-column segments).
#define DAW_TILE2_SHAPES(X) X(16, 1) X(32, 1) X(16, 2)

template <int MODE, int SNAP>
static void launch_mode2(const StepParams &P, int tz, int nc, hipStream_t st)
{
    const int grid = P.ntiles + P.nrec_blocks;
#define DAW_STEP2(Z, W)                                                                                        \
    if (tz == Z && nc == W) {                                                                                  \
        hipLaunchKernelGGL((k_step2<MODE, SNAP, Z, W>), dim3(grid), dim3(kBlock), 0, st, P);                   \
        return;                                                                                                \
    }
    DAW_TILE2_SHAPES(DAW_STEP2)
#undef DAW_STEP2
}

bool tile2_shape_ok(int tz, int nc)
{
#define DAW_OK_SHAPE2(Z, W) if (tz == Z && nc == W) return true;
    DAW_TILE2_SHAPES(DAW_OK_SHAPE2)
#undef DAW_OK_SHAPE2
    return false;
}

template <int MODE>
static void launch_snap2(int snap, const StepParams &P, int tz, int nc, hipStream_t st)
{
    if (snap == kSnapBF16) launch_mode2<MODE, kSnapBF16>(P, tz, nc, st);
    else if (snap == kSnapF16) launch_mode2<MODE, kSnapF16>(P, tz, nc, st);
    else launch_mode2<MODE, kSnapF32>(P, tz, nc, st);
}

void launch_step2(int mode, int snap, const StepParams &P, int tz, int nc, hipStream_t st)
{
    const int grid = P.ntiles + P.nrec_blocks;
    if (grid == 0) return;
    if (mode == kFwd) launch_mode2<kFwd, kSnapF32>(P, tz, nc, st);
    else if (mode == kBack) launch_snap2<kBack>(snap, P, tz, nc, st);
    else launch_snap2<kReinject>(snap, P, tz, nc, st);
}

void launch_step(int mode, int snap, const StepParams &P, int tz, int txw, hipStream_t st)
{
    const int grid = P.ntiles + P.nrec_blocks;
    if (grid == 0) return;
    if (mode == kFwd) launch_mode<kFwd, kSnapF32>(P, tz, txw, st);   // no snapshots
    else if (mode == kBack) launch_snap<kBack>(snap, P, tz, txw, st);
    else launch_snap<kReinject>(snap, P, tz, txw, st);
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

// subresample (subresample_hicks.f90:57-75): output sample i1 of every trace
// is the sum over j = max(1-it0,-4) .. min(n1in-it0,4) of coef(j,ia) *
// xin(it0+j), in that order from 0.0, with (it0, ia) = plan[i1] (it0 1-based).
// A thread owns one output sample: its window, clipping and nine
// coefficients stay in registers while it walks the traces of its grid row
// (blockIdx.y + k*gridDim.y); lanes of a wave read one stretch of input.
__global__ void __launch_bounds__(256) k_time_resample(const float *xin, float *xout, int n1in, int n1out, int ntr,
                                                       const int2 *plan, const float *coef)
{
    const int i1 = blockIdx.x * 256 + threadIdx.x;
    if (i1 >= n1out) return;
    const int2 p = plan[i1];
    const int jmin = max(1 - p.x, -4), jmax = min(n1in - p.x, 4);
    float c[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) c[j] = coef[j + 9 * (p.y - 1)];
    for (int tr = blockIdx.y; tr < ntr; tr += gridDim.y) {
        const float *x = xin + (size_t)tr * n1in + (p.x - 1);
        float acc = 0.0f;
        if (jmin == -4 && jmax == 4) {
#pragma unroll
            for (int j = -4; j <= 4; ++j) acc = acc + c[j + 4] * x[j];
        } else {
            for (int j = jmin; j <= jmax; ++j) acc = acc + c[j + 4] * x[j];
        }
        xout[(size_t)tr * n1out + i1] = acc;
    }
}

void launch_time_resample(const float *xin, float *xout, int n1in, int n1out, size_t ntr, const int *plan,
                          const float *coef, hipStream_t st)
{
    if (ntr == 0 || n1out == 0) return;
    const unsigned bx = (unsigned)((n1out + 255) / 256);
    // enough rows that the grid holds several thousand workgroups
    const unsigned by = (unsigned)std::min<size_t>(ntr, std::max<size_t>(1, 8192 / bx));
    hipLaunchKernelGGL(k_time_resample, dim3(bx, by), dim3(256), 0, st, xin, xout, n1in, n1out, (int)ntr,
                       reinterpret_cast<const int2 *>(plan), coef);
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
    kItn, kArnorm, kFac, kStopNext, kD0, kD1, kD2, kScal
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
__global__ void k_minres_scalars(const double *part, double *scal, double *hist, double *trace, const int *len,
                                 int itnlim, double rtol, int nshots)
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
    if (trace) {
        // the values the log line of this iteration prints and whether it is
        // printed (:270-285, :308-325); x(1) follows from k_minres_x
        const double an = q[kAnorm], epsx = an * q[kYnorm] * eps, epsr = an * q[kYnorm] * rtol;
        const double rel = sqrt(gbar * gbar + dbar2 * dbar2) / an;
        const bool prnt = len[s] <= 40 || itn <= 10 || itn >= itnlim - 10 || itn % 10 == 0 ||
                          phibar <= 10.0 * epsx || phibar <= 10.0 * epsr || rel <= 10.0 * epsx ||
                          rel <= 10.0 * epsr || q[kAcond] >= 1.0e-2 / eps || istop != 0;
        double *t = trace + ((size_t)s * itnlim + (itn - 1)) * kTraceStride;
        t[1] = phibar / q[kBeta1];
        t[2] = an;
        t[3] = q[kAcond];
        t[4] = prnt ? 1.0 : 0.0;
    }
    q[kItn] = itn;
    // the x update of this iteration still runs (:250-257 precede the exit);
    // k_minres_close applies the stop afterwards
    q[kStopNext] = istop;
}

// w1 <- w2, w2 <- w, w <- (v - oldeps*w1 - delta*w2) * denom, x += phi*w (:252-257).
// Buffer roles rotate on the host; here w_new is written over w1_old.
__global__ void k_minres_x(float *x, float *w_new_over_w1, const float *w2_is_old_w, const float *w1_is_old_w2,
                           const float *v, const size_t *off, const int *len, double *scal, double *trace,
                           int itnlim)
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
        const float xi = (float)((double)xs[i] + phi * (double)wi);
        xs[i] = xi;
        if (trace && i == 0) trace[((size_t)s * itnlim + ((int)q[kItn] - 1)) * kTraceStride] = (double)xi;
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

// ---------------------------------------------------------------------------
// MINRES in the reference's own arithmetic (minresModule.f90:10-360 with
// sublinearminres.f90:64-72,117-133, all REAL): the opt-in parity mode.
// Vector updates are the reference's FP32 expressions, rounded after every
// operation in its order (-ffp-contract=off).  Its dot products and datanorm
// are sequential FP32 sums, which no parallel reduction reproduces: one
// workgroup per shot forms the rounded products in LDS (waves 1-3, one chunk
// ahead) while one lane of wave 0 adds them in the reference's order.
// Scalars are FP32, kept in the shot's scalar row as exactly representable
// doubles so that the host-side stop test and the log readers are shared
// with the FP64 path.
// ---------------------------------------------------------------------------
constexpr int kSeqChunk = 2048;

// ORDER 0: sum over i ascending of a[i]*b[i] (dot_product, minresModule.f90:87,189,195).
// ORDER 1: datanorm's order over [r][t] traces, t outer and r inner
// (sublinearminres.f90:125-129).  Result to scal[s*kScal + slot].
template <int ORDER>
__global__ __launch_bounds__(256) void k_seqdot32(const float *a, const float *b, const size_t *off, const int *len,
                                                  const int *nts, double *scal, int slot, int skip_inactive)
{
    __shared__ float4 buf4[2][kSeqChunk / 4];
    const int s = blockIdx.x;
    double *q = scal + (size_t)s * kScal;
    if (skip_inactive && !active(q)) return;
    const unsigned n = (unsigned)len[s];
    const float *as = a + off[s], *bs = b + off[s];
    const unsigned nt = (unsigned)nts[s];
    const unsigned nr = n / nt;
    const int tid = threadIdx.x;
    const unsigned nchunks = (n + kSeqChunk - 1) / kSeqChunk;
    auto fill = [&](unsigned c) {
        float *p = reinterpret_cast<float *>(buf4[c & 1]);
        const unsigned k0 = c * kSeqChunk;
        const int cnt = (int)min((unsigned)kSeqChunk, n - k0);
        for (int j = tid - 64; j < cnt; j += kBlock - 64) {
            unsigned i = k0 + (unsigned)j;
            if (ORDER == 1) {
                const unsigned t = i / nr, r = i - t * nr;
                i = r * nt + t;
            }
            p[j] = as[i] * bs[i];
        }
    };
    float acc = 0.0f;
    if (tid >= 64 && nchunks > 0) fill(0);
    __syncthreads();
    for (unsigned c = 0; c < nchunks; ++c) {
        if (tid >= 64) {
            if (c + 1 < nchunks) fill(c + 1);
        } else if (tid == 0) {
            const int cnt = (int)min((unsigned)kSeqChunk, n - c * kSeqChunk);
            const float4 *p4 = buf4[c & 1];
            int j = 0;
#pragma unroll 8
            for (; j + 4 <= cnt; j += 4) {
                const float4 v = p4[j >> 2];
                acc = acc + v.x;
                acc = acc + v.y;
                acc = acc + v.z;
                acc = acc + v.w;
            }
            const float *p = reinterpret_cast<const float *>(p4);
            for (; j < cnt; ++j) acc = acc + p[j];
        }
        __syncthreads();
    }
    if (tid == 0) q[slot] = (double)acc;
}

// fac = datanorm(deltadr) / datanorm(deltade) * mu, shift = -fac
// (sublinearminres.f90:64-72), beta1 = sqrt(b.b) (minresModule.f90:83-153),
// from kD0 = datanorm(b)^2, kD1 = datanorm(H b)^2, kD2 = b.b.
__global__ void k_minres32_init(double *scal, float mu, int nshots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double *q = scal + (size_t)s * kScal;
    const float ndr = sqrtf((float)q[kD0]), nde = sqrtf((float)q[kD1]);
    const float fac = ndr / nde * mu;
    float beta1 = (float)q[kD2];
    for (int i = 0; i < kScal; ++i) q[i] = 0.0;
    q[kFac] = fac;
    q[kShift] = -fac;
    if (beta1 < 0.0f) { q[kIstop] = 8; return; }
    if (beta1 == 0.0f) { q[kIstop] = 0; q[kItn] = -1; return; }
    beta1 = sqrtf(beta1);
    q[kBeta1] = beta1;
    q[kBeta] = beta1;
    q[kQrnorm] = beta1;
    q[kPhibar] = beta1;
    q[kRhs1] = beta1;
    q[kCs] = -1.0;
}

// v = (1/beta) y (:181-182)
__global__ void k_minres32_v(const float *y, float *v, const size_t *off, const int *len, const double *scal)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    if (!active(q)) return;
    const float sc = 1.0f / (float)q[kBeta];
    const size_t n = len[s];
    const float *ys = y + off[s];
    float *vs = v + off[s];
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock)
        vs[i] = sc * ys[i];
}

// y = y - shift v, then y = y - (beta/oldb) r1 from the second iteration (:185-188)
__global__ void k_minres32_l1(float *y, const float *v, const float *r1, const size_t *off, const int *len,
                              const double *scal)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    if (!active(q)) return;
    const float shift = (float)q[kShift];
    const bool second = q[kItn] >= 1.0;    // kItn counts finished iterations
    const float c = second ? (float)q[kBeta] / (float)q[kOldb] : 0.0f;
    const size_t n = len[s];
    float *ys = y + off[s];
    const float *vs = v + off[s], *rs = r1 + off[s];
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock) {
        float yi = ys[i] - shift * vs[i];
        if (second) yi = yi - c * rs[i];
        ys[i] = yi;
    }
}

// y = y - (alfa/beta) r2 (:191)
__global__ void k_minres32_l2(float *y, const float *r2, const size_t *off, const int *len, const double *scal)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    if (!active(q)) return;
    const float c = (float)q[kAlfa] / (float)q[kBeta];
    const size_t n = len[s];
    float *ys = y + off[s];
    const float *rs = r2 + off[s];
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock)
        ys[i] = ys[i] - c * rs[i];
}

// Scalar part of one iteration (:196-303) in FP32; alfa in kAlfa, r2.y in kD0.
__global__ void k_minres32_scalars(double *scal, double *hist, double *trace, const int *len, int itnlim, float rtol,
                                   int nshots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double *q = scal + (size_t)s * kScal;
    if (!active(q)) return;
    const float eps = 1.1920928955078125e-07f;   // epsilon(real), minresModule.f90:64
    const int itn = (int)q[kItn] + 1;
    const float alfa = (float)q[kAlfa];
    const float oldb = (float)q[kBeta];
    float beta = (float)q[kD0];
    q[kOldb] = oldb;
    if (beta < 0.0f) { q[kIstop] = 6; q[kItn] = itn; return; }
    beta = sqrtf(beta);
    const float tnorm2 = (float)q[kTnorm2] + alfa * alfa + oldb * oldb + beta * beta;
    const float beta1 = (float)q[kBeta1];
    float gmax = (float)q[kGmax], gmin = (float)q[kGmin];
    int istop = 0;
    if (itn == 1) {
        if (beta / beta1 <= 10.0f * eps) istop = -1;
        gmax = alfa < 0.0f ? -alfa : alfa;
        gmin = gmax;
    }
    const float cs = (float)q[kCs], sn = (float)q[kSn], dbar = (float)q[kDbar];
    const float oldeps = (float)q[kEpsln];
    const float delta = cs * dbar + sn * alfa;
    const float gbar = sn * dbar - cs * alfa;
    const float epsln = sn * beta;
    const float dbar2 = -cs * beta;
    const float gamma = sqrtf(gbar * gbar + beta * beta);
    const float cs2 = gbar / gamma, sn2 = beta / gamma;
    const float phibar0 = (float)q[kPhibar];
    const float phi = cs2 * phibar0;
    const float phibar = sn2 * phibar0;
    const float denom = 1.0f / gamma;
    gmax = gmax > gamma ? gmax : gamma;
    gmin = gmin < gamma ? gmin : gamma;
    const float z = (float)q[kRhs1] / gamma;
    const float ynorm2 = z * z + (float)q[kYnorm2];
    const float rhs1 = (float)q[kRhs2] - delta * z;
    const float rhs2 = -epsln * z;
    const float anorm = sqrtf(tnorm2);
    const float ynorm = sqrtf(ynorm2);
    const float rnorml = (float)q[kRnorm];
    const float rootl = sqrtf(gbar * gbar + dbar2 * dbar2);
    const float acond = gmax / gmin;
    const float rel_r = phibar / beta1;
    q[kTnorm2] = tnorm2;
    q[kGmax] = gmax;
    q[kGmin] = gmin;
    q[kOldeps] = oldeps;
    q[kDelta] = delta;
    q[kEpsln] = epsln;
    q[kDbar] = dbar2;
    q[kCs] = cs2;
    q[kSn] = sn2;
    q[kPhi] = phi;
    q[kPhibar] = phibar;
    q[kDenom] = denom;
    q[kBeta] = beta;
    q[kYnorm2] = ynorm2;
    q[kRhs1] = rhs1;
    q[kRhs2] = rhs2;
    q[kAnorm] = anorm;
    q[kYnorm] = ynorm;
    q[kQrnorm] = phibar;
    q[kRnorm] = phibar;
    q[kArnorm] = rnorml * rootl;
    q[kAcond] = acond;
    if (hist) hist[(size_t)s * itnlim + (itn - 1)] = rel_r;
    if (istop == 0) {
        if (itn >= itnlim) istop = 5;
        if (rel_r <= rtol) istop = 1;
    }
    if (trace) {
        const float epsx = anorm * ynorm * eps, epsr = anorm * ynorm * rtol, rel = rootl / anorm;
        const bool prnt = len[s] <= 40 || itn <= 10 || itn >= itnlim - 10 || itn % 10 == 0 ||
                          phibar <= 10.0f * epsx || phibar <= 10.0f * epsr || rel <= 10.0f * epsx ||
                          rel <= 10.0f * epsr || acond >= 1.0e-2f / eps || istop != 0;
        double *t = trace + ((size_t)s * itnlim + (itn - 1)) * kTraceStride;
        t[1] = rel_r;
        t[2] = anorm;
        t[3] = acond;
        t[4] = prnt ? 1.0 : 0.0;
    }
    q[kItn] = itn;
    q[kStopNext] = istop;
}

// w1 <- w2, w2 <- w, w <- (v - oldeps w1 - delta w2) denom, x <- x + phi w (:252-257), FP32
__global__ void k_minres32_x(float *x, float *w_new_over_w1, const float *w2_is_old_w, const float *w1_is_old_w2,
                             const float *v, const size_t *off, const int *len, double *scal, double *trace,
                             int itnlim)
{
    const int s = blockIdx.y;
    double *q = scal + (size_t)s * kScal;
    if (!(q[kIstop] == 0.0 && q[kItn] >= 1.0)) return;
    const float oldeps = (float)q[kOldeps], delta = (float)q[kDelta], denom = (float)q[kDenom];
    const float phi = (float)q[kPhi];
    const size_t n = len[s];
    float *xs = x + off[s], *wn = w_new_over_w1 + off[s];
    const float *w2 = w2_is_old_w + off[s], *w1 = w1_is_old_w2 + off[s], *vs = v + off[s];
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock) {
        const float wi = (vs[i] - oldeps * w1[i] - delta * w2[i]) * denom;
        wn[i] = wi;
        const float xi = xs[i] + phi * wi;
        xs[i] = xi;
        if (trace && i == 0) trace[((size_t)s * itnlim + ((int)q[kItn] - 1)) * kTraceStride] = (double)xi;
    }
}

// ---------------------------------------------------------------------------
// MINRES with its Krylov vectors r1, r2 and w, w1, w2 (minresModule.f90:29)
// stored in 16 bits (BF16 or FP16): each 256-element chunk of a vector is
// scaled by the power of two at or above its max |value| (an exact
// rescaling, the snapshot codec's) and rounded to nearest even.  Arithmetic
// is that of the FP64 path (FP64 dot products and scalars, FP32 results):
// v, y and x stay FP32; a 16-bit vector is decoded where it is read and the
// value that is encoded is the FP32 one just computed.  Chunk c of shot s has
// its scale at sc[coff[s] + c].  The loops run over whole chunks so that
// every thread of a workgroup reaches the chunk's barriers.
// ---------------------------------------------------------------------------
struct Vec16 {
    unsigned short *q;
    float *sc;
};

// max |v| over the workgroup (256 threads), the same value in every thread
__device__ __forceinline__ float block_max_abs(float v)
{
    __shared__ float s_m[kBlock / 64];
    const float w = wave_max_abs(v);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = w;
    __syncthreads();
    float m = s_m[0];
    for (int i = 1; i < kBlock / 64; ++i) m = fmaxf(m, s_m[i]);
    __syncthreads();
    return m;
}

template <int F>
__device__ __forceinline__ float k16_load(const Vec16 &V, size_t base, size_t cbase, size_t i)
{
    return snap_decode<F>(V.q[base + i]) * V.sc[cbase + (i >> 8)];
}

// encode chunk element i (all threads of the block call it for one chunk)
template <int F>
__device__ __forceinline__ void k16_store(const Vec16 &V, size_t base, size_t cbase, size_t i, bool ok, float v)
{
    const float sc = pow2_scale(block_max_abs(ok ? v : 0.0f));
    if (ok) V.q[base + i] = snap_encode<F>(v * (1.0f / sc));
    if (threadIdx.x == 0) V.sc[cbase + (i >> 8)] = sc;
}

// r1 = r2 = encode(b)
template <int F>
__global__ void k_mk_init(const float *b, Vec16 ra, Vec16 rb, const size_t *off, const size_t *coff, const int *len)
{
    const int s = blockIdx.y;
    const size_t n = len[s];
    for (size_t c = blockIdx.x; c * kBlock < n; c += kVecBlocks) {
        const size_t i = c * kBlock + threadIdx.x;
        const bool ok = i < n;
        const float v = ok ? b[off[s] + i] : 0.0f;
        k16_store<F>(ra, off[s], coff[s], i, ok, v);
        k16_store<F>(rb, off[s], coff[s], i, ok, v);
    }
}

// v = y / beta with y = r2 (:181-182)
template <int F>
__global__ void k_mk_v(Vec16 r2, float *v, const size_t *off, const size_t *coff, const int *len, const double *scal)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    if (!active(q)) return;
    const double sc = 1.0 / q[kBeta];
    const size_t n = len[s];
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock)
        v[off[s] + i] = (float)(sc * (double)k16_load<F>(r2, off[s], coff[s], i));
}

// y = Hv - shift*v - (beta/oldb)*r1; partial v.y (:185-190)
template <int F>
__global__ void k_mk_l1(float *y, const float *v, Vec16 r1, const size_t *off, const size_t *coff, const int *len,
                        const double *scal, double *part)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    double acc = 0.0;
    if (active(q)) {
        const double shift = q[kShift];
        const double c = (q[kItn] >= 1.0) ? q[kBeta] / q[kOldb] : 0.0;
        const size_t n = len[s];
        for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)kVecBlocks * kBlock) {
            const double vi = v[off[s] + i];
            const float ri = c != 0.0 ? k16_load<F>(r1, off[s], coff[s], i) : 0.0f;
            const float yi = (float)((double)y[off[s] + i] - shift * vi - c * (double)ri);
            y[off[s] + i] = yi;
            acc += vi * (double)yi;
        }
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[(size_t)s * kVecBlocks + blockIdx.x] = acc;
}

// y = y - (alfa/beta)*r2; partial y.y; the new r2 = encode(y) goes over r1 (:191-197)
template <int F>
__global__ void k_mk_l2(float *y, Vec16 r2, Vec16 rout, const size_t *off, const size_t *coff, const int *len,
                        const double *scal, double *part)
{
    const int s = blockIdx.y;
    const double *q = scal + (size_t)s * kScal;
    const bool act = active(q);
    double acc = 0.0;
    if (act) {
        const double c = q[kAlfa] / q[kBeta];
        const size_t n = len[s];
        for (size_t ch = blockIdx.x; ch * kBlock < n; ch += kVecBlocks) {
            const size_t i = ch * kBlock + threadIdx.x;
            const bool ok = i < n;
            float yi = 0.0f;
            if (ok) {
                yi = (float)((double)y[off[s] + i] - c * (double)k16_load<F>(r2, off[s], coff[s], i));
                y[off[s] + i] = yi;
                acc += (double)yi * (double)yi;
            }
            k16_store<F>(rout, off[s], coff[s], i, ok, yi);
        }
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[(size_t)s * kVecBlocks + blockIdx.x] = acc;
}

// w_new = (v - oldeps*w1 - delta*w2)*denom stored over w1's slot, x += phi*w_new (:252-257)
template <int F>
__global__ void k_mk_x(float *x, Vec16 wn, Vec16 w2v, Vec16 w1v, const float *v, const size_t *off, const size_t *coff,
                       const int *len, double *scal, double *trace, int itnlim)
{
    const int s = blockIdx.y;
    double *q = scal + (size_t)s * kScal;
    if (!(q[kIstop] == 0.0 && q[kItn] >= 1.0)) return;
    const double oldeps = q[kOldeps], delta = q[kDelta], denom = q[kDenom], phi = q[kPhi];
    const size_t n = len[s];
    for (size_t ch = blockIdx.x; ch * kBlock < n; ch += kVecBlocks) {
        const size_t i = ch * kBlock + threadIdx.x;
        const bool ok = i < n;
        float wi = 0.0f;
        if (ok) {
            const double w1 = k16_load<F>(w1v, off[s], coff[s], i), w2 = k16_load<F>(w2v, off[s], coff[s], i);
            wi = (float)(((double)v[off[s] + i] - oldeps * w1 - delta * w2) * denom);
            const float xi = (float)((double)x[off[s] + i] + phi * (double)wi);
            x[off[s] + i] = xi;
            if (trace && i == 0) trace[((size_t)s * itnlim + ((int)q[kItn] - 1)) * kTraceStride] = (double)xi;
        }
        k16_store<F>(wn, off[s], coff[s], i, ok, wi);
    }
}

// w = w1 = w2 = 0
__global__ void k_mk_zero(Vec16 a, Vec16 b, Vec16 c, size_t nq, size_t nsc)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += (size_t)gridDim.x * blockDim.x) {
        a.q[i] = 0; b.q[i] = 0; c.q[i] = 0;
        if (i < nsc) { a.sc[i] = 1.0f; b.sc[i] = 1.0f; c.sc[i] = 1.0f; }
    }
}

#define K16_DISPATCH(F, KERNEL, ...)                                                                  \
    do {                                                                                             \
        if ((F) == kSnapBF16) hipLaunchKernelGGL(KERNEL<kSnapBF16>, __VA_ARGS__);                    \
        else hipLaunchKernelGGL(KERNEL<kSnapF16>, __VA_ARGS__);                                      \
    } while (0)

static Vec16 v16(unsigned short *q, float *sc) { return Vec16{q, sc}; }

void launch_mk_init(int fmt, const float *b, unsigned short *qa, float *sa, unsigned short *qb, float *sb,
                    const size_t *off, const size_t *coff, const int *len, int nshots, hipStream_t st)
{
    K16_DISPATCH(fmt, k_mk_init, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, b, v16(qa, sa), v16(qb, sb), off, coff,
                 len);
}

void launch_mk_zero(unsigned short *q0, float *s0, unsigned short *q1, float *s1, unsigned short *q2, float *s2,
                    size_t nq, size_t nsc, hipStream_t st)
{
    const size_t n = std::max(nq, nsc);
    const unsigned blocks = (unsigned)std::min<size_t>(65536, (n + 255) / 256);
    hipLaunchKernelGGL(k_mk_zero, dim3(std::max(1u, blocks)), dim3(256), 0, st, v16(q0, s0), v16(q1, s1), v16(q2, s2),
                       nq, nsc);
}

void launch_mk_v(int fmt, const unsigned short *q, const float *sc, float *v, const size_t *off, const size_t *coff,
                 const int *len, const double *scal, int nshots, hipStream_t st)
{
    K16_DISPATCH(fmt, k_mk_v, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st,
                 v16(const_cast<unsigned short *>(q), const_cast<float *>(sc)), v, off, coff, len, scal);
}

void launch_mk_l1(int fmt, float *y, const float *v, const unsigned short *q1, const float *s1, const size_t *off,
                  const size_t *coff, const int *len, const double *scal, double *part, int nshots, hipStream_t st)
{
    K16_DISPATCH(fmt, k_mk_l1, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, y, v,
                 v16(const_cast<unsigned short *>(q1), const_cast<float *>(s1)), off, coff, len, scal, part);
    hipLaunchKernelGGL(k_minres_alfa, dim3((nshots + 63) / 64), dim3(64), 0, st, part, const_cast<double *>(scal),
                       nshots);
}

void launch_mk_l2(int fmt, float *y, const unsigned short *q2, const float *s2, unsigned short *qo, float *so,
                  const size_t *off, const size_t *coff, const int *len, const double *scal, double *part, int nshots,
                  hipStream_t st)
{
    K16_DISPATCH(fmt, k_mk_l2, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, y,
                 v16(const_cast<unsigned short *>(q2), const_cast<float *>(s2)), v16(qo, so), off, coff, len, scal,
                 part);
}

void launch_mk_x(int fmt, float *x, unsigned short *qn, float *sn, const unsigned short *q2, const float *s2,
                 const unsigned short *q1, const float *s1, const float *v, const size_t *off, const size_t *coff,
                 const int *len, double *scal, double *trace, int itnlim, int nshots, hipStream_t st)
{
    K16_DISPATCH(fmt, k_mk_x, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, x, v16(qn, sn),
                 v16(const_cast<unsigned short *>(q2), const_cast<float *>(s2)),
                 v16(const_cast<unsigned short *>(q1), const_cast<float *>(s1)), v, off, coff, len, scal, trace,
                 itnlim);
    hipLaunchKernelGGL(k_minres_close, dim3((nshots + 63) / 64), dim3(64), 0, st, scal, nshots);
}

void launch_seqdot32(const float *a, const float *b, const size_t *off, const int *len, const int *nts, double *scal,
                     int which, int order, bool skip_inactive, int nshots, hipStream_t st)
{
    static const int slots[4] = {kD0, kD1, kD2, kAlfa};
    const int slot = slots[which & 3];
    if (order == 1)
        hipLaunchKernelGGL(k_seqdot32<1>, dim3(nshots), dim3(kBlock), 0, st, a, b, off, len, nts, scal, slot,
                           skip_inactive ? 1 : 0);
    else
        hipLaunchKernelGGL(k_seqdot32<0>, dim3(nshots), dim3(kBlock), 0, st, a, b, off, len, nts, scal, slot,
                           skip_inactive ? 1 : 0);
}

void launch_minres32_init(double *scal, float mu, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres32_init, dim3((nshots + 63) / 64), dim3(64), 0, st, scal, mu, nshots);
}

void launch_minres32_v(const float *y, float *v, const size_t *off, const int *len, const double *scal, int nshots,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_minres32_v, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, y, v, off, len, scal);
}

void launch_minres32_l1(float *y, const float *v, const float *r1, const size_t *off, const int *len,
                        const double *scal, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres32_l1, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, y, v, r1, off, len, scal);
}

void launch_minres32_l2(float *y, const float *r2, const size_t *off, const int *len, const double *scal,
                        int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres32_l2, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, y, r2, off, len, scal);
}

void launch_minres32_scalars(double *scal, double *hist, double *trace, const int *len, int itnlim, float rtol,
                             int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres32_scalars, dim3((nshots + 63) / 64), dim3(64), 0, st, scal, hist, trace, len, itnlim,
                       rtol, nshots);
}

void launch_minres32_x(float *x, float *wn, const float *w2, const float *w1, const float *v, const size_t *off,
                       const int *len, double *scal, double *trace, int itnlim, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres32_x, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, x, wn, w2, w1, v, off, len, scal,
                       trace, itnlim);
    hipLaunchKernelGGL(k_minres_close, dim3((nshots + 63) / 64), dim3(64), 0, st, scal, nshots);
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

void launch_minres_scalars(const double *part, double *scal, double *hist, double *trace, const int *len, int itnlim,
                           double rtol, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres_scalars, dim3((nshots + 63) / 64), dim3(64), 0, st, part, scal, hist, trace, len,
                       itnlim, rtol, nshots);
}

void launch_minres_x(float *x, float *wn, const float *w2, const float *w1, const float *v, const size_t *off,
                     const int *len, double *scal, double *trace, int itnlim, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_minres_x, dim3(kVecBlocks, nshots), dim3(kBlock), 0, st, x, wn, w2, w1, v, off, len, scal,
                       trace, itnlim);
    hipLaunchKernelGGL(k_minres_close, dim3((nshots + 63) / 64), dim3(64), 0, st, scal, nshots);
}

// ---------------------------------------------------------------------------
// Band-pass of filmartin2d (minresModule.f90:486-614) on device traces.
// The transfer function W (real, W[nto-k] = W[k]) is applied in FP64 like the
// reference's FFT.  Because W is real and even, one complex FFT of length nto
// filters two real traces at once: x1 + i x2 -> IFFT(W FFT(x1 + i x2)) =
// y1 + i y2.  One workgroup per trace pair, the whole transform in LDS
// (nto <= 8192: 128 KB).  tw[m] = exp(-2 pi i m / nto), m < nto/2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bitrev(int v, int bits) { return (int)(__brev((unsigned)v) >> (32 - bits)); }

__device__ void lds_fft(double2 *a, const double2 *__restrict__ tw, int nto, int bits, bool inverse)
{
    for (int len = 2; len <= nto; len <<= 1) {
        const int half = len >> 1, stride = nto / len;
        for (int j = threadIdx.x; j < nto / 2; j += blockDim.x) {
            const int grp = j / half, k = j - grp * half;
            const int i0 = grp * len + k, i1 = i0 + half;
            double2 w = tw[k * stride];
            if (inverse) w.y = -w.y;
            const double2 u = a[i0], v = a[i1];
            const double2 vw = make_double2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            a[i0] = make_double2(u.x + vw.x, u.y + vw.y);
            a[i1] = make_double2(u.x - vw.x, u.y - vw.y);
        }
        __syncthreads();
    }
}

// din and dout may be the same traces (in-place filtering): every load of a
// pair precedes the barriers that precede its stores, so no __restrict__.
__global__ __launch_bounds__(256) void k_filter(const float *din, float *dout, int ntr,
                                                int nt, int nto, int bits, const double *__restrict__ wgt,
                                                const double2 *__restrict__ tw)
{
    extern __shared__ double2 a[];
    const int t1 = 2 * blockIdx.x, t2 = t1 + 1;
    const bool has2 = t2 < ntr;
    for (int i = threadIdx.x; i < nto; i += blockDim.x) {
        const double x1 = i < nt ? (double)din[(size_t)t1 * nt + i] : 0.0;
        const double x2 = (i < nt && has2) ? (double)din[(size_t)t2 * nt + i] : 0.0;
        a[bitrev(i, bits)] = make_double2(x1, x2);
    }
    __syncthreads();
    lds_fft(a, tw, nto, bits, false);
    for (int k = threadIdx.x; k < nto; k += blockDim.x) {
        const int kk = k <= nto / 2 ? k : nto - k;
        // bin kk < nto/2 takes dmartin(kk+1); the Nyquist bin takes dmartin(1)
        const double g = kk == nto / 2 ? wgt[0] : wgt[kk];
        const double2 v = a[k];
        a[k] = make_double2(v.x * g, v.y * g);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nto; i += blockDim.x) {
        const int r = bitrev(i, bits);
        if (i < r) { const double2 tmp = a[i]; a[i] = a[r]; a[r] = tmp; }
    }
    __syncthreads();
    lds_fft(a, tw, nto, bits, true);
    const double sc = 1.0 / (double)nto;
    for (int i = threadIdx.x; i < nt; i += blockDim.x) {
        dout[(size_t)t1 * nt + i] = (float)(a[i].x * sc);
        if (has2) dout[(size_t)t2 * nt + i] = (float)(a[i].y * sc);
    }
}

int launch_filter(const float *din, float *dout, int ntr, int nt, int nto, const double *wgt, const double2 *tw,
                  hipStream_t st)
{
    if (ntr <= 0) return 0;
    int bits = 0;
    while ((1 << bits) < nto) ++bits;
    const size_t lds = (size_t)nto * sizeof(double2);
    (void)hipFuncSetAttribute((const void *)k_filter, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_filter, dim3((ntr + 1) / 2), dim3(256), lds, st, din, dout, ntr, nt, nto, bits, wgt, tw);
    return 0;
}

// ---------------------------------------------------------------------------
// Band-pass for FFT lengths 16384..65536 (the rest of filmartin2d's length
// list, minresModule.f90:501-510): a trace pair's transform no longer fits
// LDS, so it runs as a four-step FFT, N = N1 * N2 (N1, N2 <= 256), in three
// passes through an HBM work area Y[pair][k1][n2] (complex FP64):
//   pass 1 (columns): x[N2*n1 + n2] -> FFT over n1 -> times W_N^(n2*k1) -> Y[k1][n2]
//   pass 2 (rows):    FFT over n2 gives X[k1 + N1*k2]; times the filter;
//                     inverse FFT over k2 -> Y[k1][n2]
//   pass 3 (columns): times W_N^(-n2*k1) -> inverse FFT over k1 -> / N -> traces
// Each workgroup stages kFftChunk complex values (64 KB) of one pair in LDS:
// kFftChunk/N1 columns or kFftChunk/N2 rows.  tw[m] = exp(-2 pi i m / N),
// m < N (the butterflies of a length-L stage use tw[k * N/L]).
// ---------------------------------------------------------------------------
constexpr int kFftChunk = 4096;

// nb independent radix-2 FFTs of length L = 2^bits stored one after another
// in a, each in bit-reversed order on entry and natural order on exit.
__device__ void lds_fft_batch(double2 *a, int L, int nb, const double2 *tw, int tstride, bool inverse)
{
    const int hl = L >> 1;
    for (int len = 2; len <= L; len <<= 1) {
        const int half = len >> 1, stride = (L / len) * tstride;
        for (int j = threadIdx.x; j < nb * hl; j += blockDim.x) {
            const int b = j / hl, jj = j - b * hl;
            const int grp = jj / half, k = jj - grp * half;
            const int i0 = b * L + grp * len + k, i1 = i0 + half;
            double2 w = tw[k * stride];
            if (inverse) w.y = -w.y;
            const double2 u = a[i0], v = a[i1];
            const double2 vw = make_double2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            a[i0] = make_double2(u.x + vw.x, u.y + vw.y);
            a[i1] = make_double2(u.x - vw.x, u.y - vw.y);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b, bool conj_b)
{
    const double by = conj_b ? -b.y : b.y;
    return make_double2(a.x * b.x - a.y * by, a.x * by + a.y * b.x);
}

// Pass 1 (INV = false) and pass 3 (INV = true).  grid (N2 / C, pairs in the group).
template <bool INV>
__global__ __launch_bounds__(256) void k_bp_cols(const float *din, float *dout, double2 *Y, int ntr, int nt, int N1,
                                                 int bits1, int N2, int pair0, const double2 *tw)
{
    __shared__ double2 a[kFftChunk];
    const int C = kFftChunk / N1;
    const int p = blockIdx.y, pair = pair0 + p;
    const int t1 = 2 * pair, t2 = t1 + 1;
    const bool has2 = t2 < ntr;
    const int N = N1 * N2, c0 = blockIdx.x * C;
    double2 *Yp = Y + (size_t)p * N;
    for (int j = threadIdx.x; j < kFftChunk; j += blockDim.x) {
        const int c = j % C, i = j / C, n2 = c0 + c;          // i: n1 (pass 1) or k1 (pass 3)
        double2 v;
        if (!INV) {
            const int n = N2 * i + n2;
            v = make_double2(n < nt ? (double)din[(size_t)t1 * nt + n] : 0.0,
                             (n < nt && has2) ? (double)din[(size_t)t2 * nt + n] : 0.0);
        } else {
            v = cmul(Yp[(size_t)i * N2 + n2], tw[(n2 * i) & (N - 1)], true);
        }
        a[c * N1 + bitrev(i, bits1)] = v;
    }
    __syncthreads();
    lds_fft_batch(a, N1, C, tw, N / N1, INV);
    const double sc = 1.0 / (double)N;
    for (int j = threadIdx.x; j < kFftChunk; j += blockDim.x) {
        const int c = j % C, i = j / C, n2 = c0 + c;          // i: k1 (pass 1) or n1 (pass 3)
        const double2 v = a[c * N1 + i];
        if (!INV) {
            Yp[(size_t)i * N2 + n2] = cmul(v, tw[(n2 * i) & (N - 1)], false);
        } else {
            const int n = N2 * i + n2;
            if (n < nt) {
                dout[(size_t)t1 * nt + n] = (float)(v.x * sc);
                if (has2) dout[(size_t)t2 * nt + n] = (float)(v.y * sc);
            }
        }
    }
}

// Pass 2: rows k1 of the group's pairs.  grid (N1 / R, pairs in the group).
__global__ __launch_bounds__(256) void k_bp_rows(double2 *Y, int N1, int N2, int bits2, const double *wgt,
                                                 const double2 *tw)
{
    __shared__ double2 a[kFftChunk];
    const int R = kFftChunk / N2;
    const int p = blockIdx.y, r0 = blockIdx.x * R;
    const int N = N1 * N2;
    double2 *Yp = Y + (size_t)p * N;
    for (int j = threadIdx.x; j < kFftChunk; j += blockDim.x) {
        const int r = j / N2, n2 = j - r * N2;
        a[r * N2 + bitrev(n2, bits2)] = Yp[(size_t)(r0 + r) * N2 + n2];
    }
    __syncthreads();
    lds_fft_batch(a, N2, R, tw, N / N2, false);
    for (int j = threadIdx.x; j < kFftChunk; j += blockDim.x) {
        const int r = j / N2, k2 = j - r * N2;
        const int k = (r0 + r) + N1 * k2;                     // frequency index of X
        const int kk = k <= N / 2 ? k : N - k;
        // bin kk < N/2 takes dmartin(kk+1); the Nyquist bin takes dmartin(1)
        const double gain = kk == N / 2 ? wgt[0] : wgt[kk];
        a[j] = make_double2(a[j].x * gain, a[j].y * gain);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kFftChunk; j += blockDim.x) {
        const int r = j / N2, i = j - r * N2, rv = bitrev(i, bits2);
        if (i < rv) {
            const double2 tmp = a[r * N2 + i];
            a[r * N2 + i] = a[r * N2 + rv];
            a[r * N2 + rv] = tmp;
        }
    }
    __syncthreads();
    lds_fft_batch(a, N2, R, tw, N / N2, true);
    for (int j = threadIdx.x; j < kFftChunk; j += blockDim.x) {
        const int r = j / N2, n2 = j - r * N2;
        Yp[(size_t)(r0 + r) * N2 + n2] = a[j];
    }
}

int filter_long_split(int nto, int &n1, int &n2)
{
    int bits = 0;
    while ((1 << bits) < nto) ++bits;
    if ((1 << bits) != nto || bits < 14 || bits > 16) return -1;
    n1 = 1 << ((bits + 1) / 2);      // 16384 = 128 x 128, 32768 = 256 x 128, 65536 = 256 x 256
    n2 = nto / n1;
    return 0;
}

int launch_filter_long(const float *din, float *dout, int ntr, int nt, int nto, const double *wgt, const double2 *tw,
                       double2 *work, int pairs_per_group, hipStream_t st)
{
    int n1 = 0, n2 = 0;
    if (filter_long_split(nto, n1, n2) != 0 || nt > nto || pairs_per_group < 1) return -1;
    int b1 = 0, b2 = 0;
    while ((1 << b1) < n1) ++b1;
    while ((1 << b2) < n2) ++b2;
    const int npairs = (ntr + 1) / 2;
    for (int p0 = 0; p0 < npairs; p0 += pairs_per_group) {
        const int np = std::min(pairs_per_group, npairs - p0);
        hipLaunchKernelGGL(k_bp_cols<false>, dim3(n2 / (kFftChunk / n1), np), dim3(256), 0, st, din, dout, work, ntr,
                           nt, n1, b1, n2, p0, tw);
        hipLaunchKernelGGL(k_bp_rows, dim3(n1 / (kFftChunk / n2), np), dim3(256), 0, st, work, n1, n2, b2, wgt, tw);
        hipLaunchKernelGGL(k_bp_cols<true>, dim3(n2 / (kFftChunk / n1), np), dim3(256), 0, st, din, dout, work, ntr,
                           nt, n1, b1, n2, p0, tw);
    }
    return 0;
}

int minres_scal_stride() { return kScal; }
int minres_vec_blocks() { return kVecBlocks; }

}  // namespace daw
