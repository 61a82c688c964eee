// Device-side data layout shared by the kernels and the batch engine.
//
// Layout in HBM (all FP32 unless noted):
//   * every extended-grid field is row-major with x fastest: f[z*ldx + x],
//     z in [0,n1e), x in [0,n2e), row pitch ldx = n2e rounded up to 64 floats
//     (256 B) so each wavefront row load is one aligned 256-byte segment;
//   * the wave state px,pz,vx,vz is double-buffered (read [cur], write
//     [cur^1]) so one fused launch per time step never reads a value a
//     neighbouring workgroup has already overwritten;
//   * back-propagation snapshots sp,svx,svz are interior-only [it][z][x],
//     x fastest, n1*n2 per step;
//   * data traces follow the reference: trace-major, time fastest [r][t].
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace daw {

enum SweepMode { kFwd = 0, kBack = 1, kReinject = 2 };
enum SnapFormat { kSnapF32 = 0, kSnapBF16 = 1, kSnapF16 = 2 };

// Owned tile per workgroup and the halo widths the 4th-order staggered
// stencil needs (two cells each side for the derivative, and one more
// pressure ring recomputed locally so the velocity half-step needs no
// second launch).
// TX columns (one wavefront width) by TZ rows; TZ is 16 or 32, a launch
// parameter the tile autotuner picks.
constexpr int TX = 64;
// Record blocks: items per workgroup and the LDS stride of one item's products.
constexpr int kRecPerBlock = 8;
constexpr int kRecStride = 444;   // 441 window products padded to whole float4s
constexpr int kBlock = 256;
// Time slots of the record band (slot of step it: it & (kBandSlots-1)): a
// step's band is recorded within kBandSlots steps of being written.
constexpr int kBandSlots = 16;

struct ShotDev {
    // grid
    int n1, n2, npml, n1e, n2e, ldx, nt, ofs;
    float dt, dtdx;
    // model (extended grid) and 1-D sponge profiles
    const float *kappa;                 // [n1e][ldx]
    const float *d1a, *d1a1, *d1b, *bw1; // [n1e]; bw1 = dtdx*damp1b1
    const float *d2a, *d2a1, *d2b, *bw2; // [n2e]; bw2 = dtdx*damp2b1
    // double-buffered wave state
    float *px[2], *pz[2], *vx[2], *vz[2];
    // Hicks injection gather into px (source or receiver residuals):
    // inj_row[z] = q >= 0 for rows touched by a window; entries
    // inj_ptr[q*(n2e+1)+x] .. [+1] list (receiver, c*dt) in receiver order.
    const int *inj_row;
    const int *inj_ptr;
    const int *inj_r;
    const float *inj_cdt;
    int inj_z0, inj_z1;                 // first and last row with inj_row[z] >= 0 (z1 < z0: none)
    const float *inj_amp;               // amplitude of receiver r at step it: inj_amp[r*amp_stride + it*amp_tstride]
    int amp_stride;
    int amp_tstride;                    // 1 for [r][t] traces; the receiver count for a [t][r] (transposed) copy
    int amp_rev;                        // amplitude time index nt-1-it (adjoint of a time-reversed input)
    int vinj;                           // 1: the source is a vz source (type_src = 1), applied in the velocity phase
    // record band: band_row[z] = slot >= 0 for rows some record window touches
    const int *band_row;
    float *pband;                       // kBandSlots time slots of [nslot][ldx], it & (kBandSlots-1)
    int band_stride;                    // floats per time slot
    // snapshots (kBack writes, kReinject reads) of the interior rows, [nt][n1][spitch]:
    // element (z, extended column gx) at z*spitch + gx - sxb, sxb = npml
    // rounded down to 64, so a wavefront's 64 columns are one aligned run.
    // FP32, or 16-bit (BF16 / FP16) with one power-of-two scale per 64-column
    // row segment of each snapshot plane: ssc[((it*3 + field)*n1 + z)*nseg + gx/64]
    float *sp, *svx, *svz;
    unsigned short *sp16, *svx16, *svz16;
    float *ssc;
    int nseg;
    int spitch, sxb;
    int xlo1, xhi1, xlo2, xhi2;         // pressure extended-source range, extended coords inclusive
    // optional full interior wavefield, Fortran (n1,n2,nt)
    float *wf;
};

struct RecItem {
    int shot;
    int ix1, ix2;       // 0-based extended-grid centre node
    int clip;           // 0: window clipped to the extended grid, 1: to the interior
    int reverse;        // write time index nt-1-it instead of it
    int cstride;        // coefficient k of the point at coefT[k * cstride]
    const float *coef;  // 441 coefficients, (i1+10)*21 + (i2+10) (record order)
    const float *coefT; // the same, transposed: the point set's coefficient k of all points contiguous
    float *out;         // trace base (nt floats)
};

// Batched record groups: items first .. first+count-1 (count <= 64, one
// shot, consecutive points of one point set) whose windows lie in grid rows
// z0 .. z0+nz-1 and columns x0 .. x0+nx-1 (nz <= kRecRows, nx <= kRecCols).
// (4 x 21 staged rows of 158 floats: 53 KB of LDS, three workgroups per CU)
constexpr int kRecSteps = 4;
constexpr int kRecRows = 21;
constexpr int kRecCols = 138;
struct RecGroup {
    int first, count;
    int z0, nz, x0, nx;
};

struct StepParams {
    const ShotDev *shots;
    const int2 *tiles;  // (shot, tile) in XCD-aware launch order
    int ntiles;
    int it;             // first step advanced by the tile blocks
    int nsub;           // steps advanced by this launch (1, or 2 for the two-step kernel)
    int cur;            // state buffers read
    const RecItem *rec; // record of steps rec_it .. rec_it+rec_nsub-1 (pband slot step & 3)
    int nrec;           // items per recorded step
    int rec_it;
    int rec_nsub;
    int nrec_blocks;
};

}  // namespace daw
