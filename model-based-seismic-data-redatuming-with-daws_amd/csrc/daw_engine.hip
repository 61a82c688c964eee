// Batch engine and C-ABI of the DAW redatuming library (include/daw.h).
//
// The engine advances many shots together: every launch of k_step moves all
// active shots of a batch one time step, so a C2-size batch of 32 shots is
// one launch of ~7,000 workgroups instead of 32 launches of 224.  Geometry
// tables (Hicks coefficients, injection gathers, record bands) are built on
// the host once per acquisition and shared by every shot that uses them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "daw.h"
#include "daw_device.h"
#include "daw_host.h"
#include "daw_io.h"
#include "daw_kernels.h"

namespace daw {

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(DAW_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------ state
// Launch configuration of one sweep mode: time steps per launch and tile;
// march = 1: the column-march kernel, tz rows per wavefront (txw unused);
// march = 2: the two-step column march of the re-injection (steps = 2).
// rec_batch = 1: the Hicks records of a sweep run as batched launches every
// few steps (k_rec_group) instead of as record blocks inside the step launches.
struct LaunchCfg {
    int steps = 1, tz = 16, txw = 1, march = 0, rec_batch = 0;
};
// The tile's columns as the API reports them: 64 per wavefront, or the 58
// owned columns of a marching wavefront.
static int tile_cols(const LaunchCfg &c) { return c.march == 2 ? kMarch2Cols : c.march ? kMarchCols : 64 * c.txw; }

struct Global {
    std::mutex mu;
    int hicks_ofs = -1000;     // ofs given to kwsinc2d; -1000: not initialised
    hipStream_t stream = nullptr;
    bool timing = false;
    double t_ms[3] = {0, 0, 0};
    long long t_cnt[3] = {0, 0, 0};
    double t_rec_ms[3] = {0, 0, 0};   // the batched record launches within those sweeps
    std::vector<std::pair<hipEvent_t, hipEvent_t>> rec_ev;   // events around them (timing on)
    double sweep_ms = 0.0;
    long long pts = 0;         // point-updates executed (sum n1e*n2e per shot per step)
    long long launches = 0;
    int max_batch_used = 0;
    int tz = 16;               // tile rows of k_step
    int txw = 1;               // tile width in wavefronts (64 columns each)
    int march = 0;             // 1: the column-march kernel with tz rows per wavefront
    int rec_batch = 1;         // 1: batched record launches (k_rec_group) instead of in-launch record blocks
    bool tz_auto = true;       // pick tz per batch shape with a timed probe
    int steps = 1;             // time steps per launch of the sweep kernels (1: k_step, 2: k_step2)
    bool steps_auto = true;    // let the probe pick the steps per launch too
    bool hess_exact = false;   // Hessian pass 1 = exact transpose of pass 2
    int snap_fmt = kSnapF32;   // snapshot store of the reference Hessian (SnapFormat)
    bool graph = false;        // replay full sweeps as captured hipGraphs
    int minres_arith = 0;      // 0: FP64 dot products and scalars; 1: the reference's FP32 arithmetic
    int krylov_fmt = kSnapF32; // storage of MINRES r1, r2, w, w1, w2 in the FP64 path (SnapFormat)
    std::map<std::vector<int>, std::vector<int>> tz_cache;   // batch shape -> {steps, tz, txw, march} per mode
    LaunchCfg last_cfg[3];     // launch configuration of each mode in the last tuned batch
};
static Global g;

static int ensure_stream()
{
    if (!g.stream) HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    return 0;
}

// RAII device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p; bytes = o.bytes;
            o.p = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t n)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = n;
        if (n == 0) return 0;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) { p = nullptr; return fail(DAW_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n, hipGetErrorString(e)); }
        return 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

template <class T>
static int upload(DevBuf &b, const std::vector<T> &v)
{
    if (int rc = b.alloc(std::max<size_t>(v.size(), 1) * sizeof(T))) return rc;
    if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// ------------------------------------------------------------------ geometry
struct Grid {
    int n1, n2, npml, n1e, n2e, ldx, ofs;
    float h, dt;
    size_t plane_e() const { return (size_t)n1e * ldx; }
    size_t plane_i() const { return (size_t)n1 * n2; }
    // snapshot rows: extended columns sxb .. npml+n2-1, padded to whole 64-column segments
    int sxb() const { return npml & ~63; }
    int spitch() const { return (npml + n2 - sxb() + 63) & ~63; }
    size_t plane_s() const { return (size_t)n1 * spitch(); }
};

static Grid make_grid(int n1, int n2, int npml, float h, float dt, int ofs)
{
    Grid G;
    G.n1 = n1; G.n2 = n2; G.npml = npml; G.h = h; G.dt = dt; G.ofs = ofs;
    G.n1e = n1 + 2 * npml;
    G.n2e = n2 + 2 * npml;
    G.ldx = (G.n2e + 63) / 64 * 64;
    return G;
}

static int check_grid(int n1, int n2, int npml, int nt, float h, float dt)
{
    if (n1 < 1 || n2 < 1 || npml < 3 || nt < 1) return fail(DAW_ERR_ARG, "bad grid n1=%d n2=%d npml=%d nt=%d (npml >= 3)", n1, n2, npml, nt);
    if (!(h > 0.0f) || !(dt > 0.0f)) return fail(DAW_ERR_ARG, "h and dt must be positive");
    if ((long long)(n1 + 2 * npml) * (n2 + 2 * npml) > (1LL << 31)) return fail(DAW_ERR_ARG, "grid too large");
    return 0;
}

// Model on the device: kappa (x fastest, pitch ldx) and the 1-D profiles.
struct DevModel {
    Grid G;
    DevBuf kappa, prof;   // prof: d1a d1a1 d1b bw1 [n1e each], d2a d2a1 d2b bw2 [n2e each]
    const float *d1a() const { return prof.as<float>(); }
    const float *d1a1() const { return prof.as<float>() + G.n1e; }
    const float *d1b() const { return prof.as<float>() + 2 * G.n1e; }
    const float *bw1() const { return prof.as<float>() + 3 * G.n1e; }
    const float *d2a() const { return prof.as<float>() + 4 * G.n1e; }
    const float *d2a1() const { return prof.as<float>() + 4 * G.n1e + G.n2e; }
    const float *d2b() const { return prof.as<float>() + 4 * G.n1e + 2 * G.n2e; }
    const float *bw2() const { return prof.as<float>() + 4 * G.n1e + 3 * G.n2e; }
};

// vp is the reference's vp(n1,n2) (z fastest) on the host.
static int make_model(DevModel &M, const Grid &G, const float *vp)
{
    M.G = G;
    std::vector<float> kap((size_t)G.n1e * G.ldx, 0.0f);
    for (int z = 0; z < G.n1e; ++z) {
        const int j1 = std::min(std::max(z - G.npml, 0), G.n1 - 1);          // submodext edge replication
        for (int x = 0; x < G.n2e; ++x) {
            const int j2 = std::min(std::max(x - G.npml, 0), G.n2 - 1);
            const double v = (double)vp[(size_t)j1 + (size_t)j2 * G.n1];
            kap[(size_t)z * G.ldx + x] = (float)(1.0 * v * v);             // kappa = dble(rho)*dble(vp)**2, rho = 1
        }
    }
    std::vector<float> d1a, d1a1, d1b, d1b1, d2a, d2a1, d2b, d2b1;
    sponge_a(G.npml, G.n1, d1a, d1a1);
    sponge_b(G.npml, G.n1, d1b, d1b1);
    sponge_a(G.npml, G.n2, d2a, d2a1);
    sponge_b(G.npml, G.n2, d2b, d2b1);
    const float dtdx = G.dt / G.h;
    std::vector<float> prof;
    prof.reserve(4 * (G.n1e + G.n2e));
    prof.insert(prof.end(), d1a.begin(), d1a.end());
    prof.insert(prof.end(), d1a1.begin(), d1a1.end());
    prof.insert(prof.end(), d1b.begin(), d1b.end());
    for (float v : d1b1) prof.push_back(dtdx * 1.0f * v);      // buw1 = dtdx*bu1*damp1b1, bu1 = 1
    prof.insert(prof.end(), d2a.begin(), d2a.end());
    prof.insert(prof.end(), d2a1.begin(), d2a1.end());
    prof.insert(prof.end(), d2b.begin(), d2b.end());
    for (float v : d2b1) prof.push_back(dtdx * 1.0f * v);      // buw2
    if (int rc = upload(M.kappa, kap)) return rc;
    return upload(M.prof, prof);
}

// Hicks tables of one point set (sources, receivers or datum receivers).
struct PointSet {
    int n = 0;
    std::vector<int> ix1, ix2;      // 0-based extended centre
    std::vector<float> coef;        // n * 441
};

// Coefficient tiles for the record blocks: each 21x21 tile in the order the
// record sums it, (i1+10)*21 + (i2+10) (PointSet keeps (i1+10) + 21*(i2+10)).
// bt gets the same tiles transposed, coefficient k of every point
// contiguous (t[k*n + r]), for the batched record's lanes of neighbouring points.
static int upload_record_coef(DevBuf &b, DevBuf &bt, const std::vector<float> &coef)
{
    const size_t n = coef.size() / 441;
    std::vector<float> t(coef.size()), tt(coef.size());
    for (size_t r = 0; r < n; ++r)
        for (int a = 0; a < 21; ++a)
            for (int c = 0; c < 21; ++c) {
                t[r * 441 + a * 21 + c] = coef[r * 441 + a + 21 * c];
                tt[(size_t)(a * 21 + c) * n + r] = coef[r * 441 + a + 21 * c];
            }
    if (int rc = upload(b, t)) return rc;
    return upload(bt, tt);
}

static int make_points(PointSet &P, const int *tab, int n, const Grid &G, bool interior_clip)
{
    P.n = n;
    P.ix1.resize(n);
    P.ix2.resize(n);
    P.coef.resize((size_t)n * 441);
    for (int r = 0; r < n; ++r) {
        const int i1 = tab[r], i2 = tab[r + n], a1 = tab[r + 2 * n], a2 = tab[r + 3 * n], ip = tab[r + 4 * n];
        if (a1 < 1 || a1 > kNalpha || a2 < 1 || a2 > kNalpha || ip < 0 || ip > kIrad)
            return fail(DAW_ERR_ARG, "point %d: alpha/free-surface index out of range (%d,%d,%d)", r, a1, a2, ip);
        const int lo1 = interior_clip ? G.npml + 1 : 1, hi1 = interior_clip ? G.npml + G.n1 : G.n1e;
        const int lo2 = interior_clip ? G.npml + 1 : 1, hi2 = interior_clip ? G.npml + G.n2 : G.n2e;
        if (i1 < lo1 - kIrad || i1 > hi1 + kIrad || i2 < lo2 - kIrad || i2 > hi2 + kIrad)
            return fail(DAW_ERR_ARG, "point %d: grid index (%d,%d) outside the grid", r, i1, i2);
        P.ix1[r] = i1 - 1;
        P.ix2[r] = i2 - 1;
        hicks2d_tile(a1, a2, ip, g.hicks_ofs, &P.coef[(size_t)r * 441]);
    }
    return 0;
}

// Injection gather of a point set, window clipped to the extended grid.
struct Inject {
    std::vector<int> row;      // n1e
    int z0 = 0, z1 = -1;       // first and last row with row[z] >= 0
    std::vector<int> ptr;      // nq*(n2e+1)
    std::vector<int> rid;
    std::vector<float> cdt;
    DevBuf d_row, d_ptr, d_rid, d_cdt;
};

static int make_inject(Inject &I, const PointSet &P, const Grid &G)
{
    I.row.assign(G.n1e, -1);
    int nq = 0;
    for (int r = 0; r < P.n; ++r)
        for (int i1 = -kIrad; i1 <= kIrad; ++i1) {
            const int z = P.ix1[r] + i1;
            if (z >= 0 && z < G.n1e && I.row[z] < 0) I.row[z] = 0;
        }
    I.z0 = 0;
    I.z1 = -1;
    for (int z = 0; z < G.n1e; ++z)
        if (I.row[z] >= 0) {
            if (nq == 0) I.z0 = z;
            I.z1 = z;
            I.row[z] = nq++;
        }
    std::vector<int> cnt((size_t)nq * (G.n2e + 1), 0);
    auto for_window = [&](int r, auto &&fn) {
        const int c1 = P.ix1[r], c2 = P.ix2[r];
        const int lo1 = std::max(-c1, -kIrad), hi1 = std::min(G.n1e - 1 - c1, kIrad);
        const int lo2 = std::max(-c2, -kIrad), hi2 = std::min(G.n2e - 1 - c2, kIrad);
        for (int i1 = lo1; i1 <= hi1; ++i1)
            for (int i2 = lo2; i2 <= hi2; ++i2) fn(c1 + i1, c2 + i2, P.coef[(size_t)r * 441 + (i1 + 10) + 21 * (i2 + 10)]);
    };
    for (int r = 0; r < P.n; ++r)
        for_window(r, [&](int z, int x, float) { cnt[(size_t)I.row[z] * (G.n2e + 1) + x]++; });
    I.ptr.assign((size_t)nq * (G.n2e + 1), 0);
    int acc = 0;
    for (int q = 0; q < nq; ++q)
        for (int x = 0; x <= G.n2e; ++x) {
            const size_t k = (size_t)q * (G.n2e + 1) + x;
            I.ptr[k] = acc;
            if (x < G.n2e) acc += cnt[k];
        }
    // per-row tables are [ptr(x), ptr(x+1)) with ptr(n2e) closing the row
    I.rid.assign(acc, 0);
    I.cdt.assign(acc, 0.0f);
    std::vector<int> fillp(I.ptr);
    for (int r = 0; r < P.n; ++r)   // receivers in order: each point's list follows the reference's loop order
        for_window(r, [&](int z, int x, float c) {
            const size_t k = (size_t)I.row[z] * (G.n2e + 1) + x;
            const int e = fillp[k]++;
            I.rid[e] = r;
            I.cdt[e] = c * G.dt;                              // coeff*dt, first product of :175 / :1042
        });
    if (int rc = upload(I.d_row, I.row)) return rc;
    if (int rc = upload(I.d_ptr, I.ptr)) return rc;
    if (int rc = upload(I.d_rid, I.rid)) return rc;
    return upload(I.d_cdt, I.cdt);
}

// Record band: rows that any record window of the given sets touches.
struct Band {
    std::vector<int> row;
    int nslot = 0;
    DevBuf d_row;
};

static int make_band(Band &B, const std::vector<const PointSet *> &sets, const Grid &G)
{
    B.row.assign(G.n1e, -1);
    for (const PointSet *P : sets)
        for (int r = 0; r < P->n; ++r)
            for (int i1 = -kIrad; i1 <= kIrad; ++i1) {
                const int z = P->ix1[r] + i1;
                if (z >= 0 && z < G.n1e) B.row[z] = 0;
            }
    B.nslot = 0;
    for (int z = 0; z < G.n1e; ++z)
        if (B.row[z] >= 0) B.row[z] = B.nslot++;
    return upload(B.d_row, B.row);
}

// Launch order of the tiles of a sweep (XCD-aware: workgroup b runs on XCD
// b % 8, and each XCD gets one contiguous run of tiles so that halo re-reads
// of neighbouring tiles hit its L2).  Within each XCD's run, kTileRaster
// keeps raster order; kTileCostFirst puts the tiles in decreasing cost class
// cls[s][t], raster order inside a class.  The back-propagation's classes are
// 2 = tile whose pressure rows meet the rows of the receivers' Hicks windows
// (the receiver terms' dependent loads make these the longest tiles), 1 =
// any other tile, 0 = tile mostly past the grid.  Measured at 201x801 with
// 24 shots per launch, 16x64 tiles, one process: 41.6 us per launch with the
// band tiles first, 42.4 with the sponge and boundary tiles first as well,
// 44.5 in raster order.  Re-injection, whose interior tiles read the
// snapshots from HBM, is slower cost-first (59.8-61.2 against 57.1 us on an
// earlier kernel: its HBM reads then bunch up in the second half of the
// launch) and the forward sweep gains nothing.
enum TileOrder { kTileRaster = 0, kTileCostFirst = 1 };
static int tile_order(int mode) { return mode == kBack ? kTileCostFirst : kTileRaster; }

static std::vector<int2> make_tiles(const std::vector<int> &tiles_per_shot, const std::vector<std::vector<int>> &cls,
                                    int order)
{
    std::vector<int2> flat;
    for (size_t s = 0; s < tiles_per_shot.size(); ++s)
        for (int t = 0; t < tiles_per_shot[s]; ++t) flat.push_back(make_int2((int)s, t));
    const int N = (int)flat.size(), X = 8, q = N / X, rem = N % X;
    if (order == kTileCostFirst)
        for (int k = 0; k < X; ++k) {
            const int start = k * q + std::min(k, rem), len = q + (k < rem ? 1 : 0);
            std::stable_sort(flat.begin() + start, flat.begin() + start + len,
                             [&](const int2 &a, const int2 &b) { return cls[a.x][a.y] > cls[b.x][b.y]; });
        }
    std::vector<int2> out(N);
    for (int b = 0; b < N; ++b) {
        const int k = b % X, j = b / X;
        const int start = k * q + std::min(k, rem);
        out[b] = flat[start + j];
    }
    return out;
}

// ------------------------------------------------------------------ sweeps
struct TileList {
    DevBuf d;        // (shot, tile) pairs in XCD-aware launch order
    int n = 0;
};

struct SweepSet {
    int nshots = 0;
    std::vector<ShotDev> host;
    DevBuf d_shots;
    LaunchCfg cfg[3];                                   // per SweepMode
    std::vector<int2> inj_band;                         // per shot: rows with receiver injection (z0, z1)
    std::map<std::vector<int>, TileList> tiles;         // per tiles_key()
    int nt_max = 0;
    long long pts_per_step = 0;
    int snap = kSnapF32;       // SnapFormat of the shots' snapshot store
    // record items of the next sweep: one device buffer for the batch's
    // lifetime (so a captured sweep stays valid), host copy kept for the
    // asynchronous upload
    DevBuf d_items;
    std::vector<RecItem> items;
    DevBuf d_groups;                                    // record groups of the items (batched record)
    std::vector<RecGroup> groups;
    // captured sweeps, keyed by everything their launches bake in
    std::map<std::vector<long long>, hipGraphExec_t> graphs;
    SweepSet() = default;
    SweepSet(const SweepSet &) = delete;
    SweepSet &operator=(const SweepSet &) = delete;
    ~SweepSet() { clear_graphs(); }
    void clear_graphs()
    {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
};

// Steps per launch of the sweeps: two only where the tile has a two-step kernel.
static int steps_per_launch(const LaunchCfg &c)
{
    if (c.march == 2) return 2;
    return (!c.march && c.steps == 2 && tile2_shape_ok(c.tz, c.txw)) ? 2 : 1;
}

// Row heights of the two-step march (a walk of rows + 12 iterations, one per lane).
static bool march2_rows_ok(int h) { return h == 16 || h == 24 || h == 32 || h == 40 || h == 48; }

// The one-step march that runs a two-step march's last step when the sweep
// has an odd number of steps (32-row segments for the taller two-step ones).
static LaunchCfg march2_tail_cfg(const LaunchCfg &c) { return LaunchCfg{1, std::min(c.tz, 32), 0, 1, c.rec_batch}; }

static std::vector<int> tiles_key(const LaunchCfg &c, int mode)
{
    return {c.tz, c.txw, c.march ? 10 + (mode == kBack) : tile_order(mode), c.march};
}

// Row segments of the column march: h rows each, except that a segment which
// would reach rows b0..b1 (those with receiver injection terms, about 14 per
// point at 401 receivers) has kMarchBandRows rows, so that no wavefront walks
// many of them.  Every segment has exactly h or kMarchBandRows rows (the
// kernel unrolls both walks); the last may reach past the grid.
static std::vector<int2> march_segments(int n1e, int h, int b0, int b1)
{
    std::vector<int2> segs;
    for (int z = 0; z < n1e;) {
        const int r = (b0 <= b1 && z <= b1 && z + h > b0) ? kMarchBandRows : h;
        segs.push_back(make_int2(z, r));
        z += r;
    }
    return segs;
}

// The exact adjoint (k_adj) runs on the back-propagation's tile, or on 16x64
// where the back-propagation marches.
static LaunchCfg adjoint_cfg(const LaunchCfg &c) { return c.march ? LaunchCfg{1, 16, 1, 0, c.rec_batch} : c; }

// The column march runs the forward sweep, and the Hessian sweeps with FP32
// snapshots; the two-step march only the re-injection, with FP32 snapshots and
// batched records.
static bool march_ok(const LaunchCfg &c, int mode, int snap)
{
    if (c.march == 2) return mode == kReinject && snap == kSnapF32 && c.rec_batch && march2_rows_ok(c.tz);
    return !c.march || mode == kFwd || snap == kSnapF32;
}

// The configured launch of a mode (16x64 where a fixed march cannot run it).
static LaunchCfg fixed_cfg(int mode, int snap)
{
    const LaunchCfg c{g.march == 2 ? 2 : g.steps, g.tz, g.txw, g.march, g.rec_batch};
    return march_ok(c, mode, snap) ? c : LaunchCfg{1, 16, 1, 0, g.rec_batch};
}

// Row segments of the two-step march: about n1e/h segments of balanced
// heights near h (the kernel takes a segment's height from its tile entry,
// up to 52), so no segment is a sliver (241 rows at h = 24: ten of 24-25
// rows, not eleven with a last one of one row).  With the free surface a
// segment whose step-A velocity rows (z0-3 .. z0+rows+2) would end at row
// npml-1 gives its last row to the next segment, so that the segment holding
// row npml-1 also holds row npml (the kernel's vz mirror needs both).
static std::vector<int2> march2_segments(int n1e, int h, int npml, int ofs)
{
    int nseg = std::max(1, (n1e + h / 2) / h);
    while ((n1e + nseg - 1) / nseg > 51) ++nseg;
    const int base = n1e / nseg, rem = n1e % nseg;
    std::vector<int> rows(nseg);
    for (int i = 0; i < nseg; ++i) rows[i] = base + (i < rem ? 1 : 0);
    std::vector<int2> segs;
    for (size_t i = 0, z = 0; i < rows.size(); z += rows[i], ++i) {
        if (ofs == 1 && (int)z + rows[i] + 2 == npml - 1 && rows[i] > 1) {
            --rows[i];
            if (i + 1 < rows.size()) ++rows[i + 1];
            else rows.push_back(1);
        }
        segs.push_back(make_int2((int)z, rows[i]));
    }
    return segs;
}

// Copies the record items of the next sweep to the batch's item buffer
// (stream-ordered: the previous sweep has read the old items by then).
static int set_items(SweepSet &W, const std::vector<RecItem> &items)
{
    const size_t bytes = std::max<size_t>(items.size(), 1) * sizeof(RecItem);
    if (bytes > W.d_items.bytes) {
        HIPCHK(hipStreamSynchronize(g.stream));
        W.clear_graphs();
        if (int rc = W.d_items.alloc(std::max(bytes, 2 * W.d_items.bytes))) return rc;
    }
    HIPCHK(hipStreamSynchronize(g.stream));   // the host copy of the last upload is free
    // by shot, then point: the batched record's lanes of a wavefront read
    // neighbouring band columns and transposed coefficients
    W.items = items;
    std::stable_sort(W.items.begin(), W.items.end(), [](const RecItem &a, const RecItem &b) {
        return a.shot != b.shot ? a.shot < b.shot : a.coef < b.coef;
    });
    if (!items.empty())
        HIPCHK(hipMemcpyAsync(W.d_items.p, W.items.data(), items.size() * sizeof(RecItem), hipMemcpyHostToDevice,
                              g.stream));
    // record groups: runs of consecutive points of one point set of one shot
    // whose windows (clipped to the grid) fit a kRecRows x kRecCols block
    W.groups.clear();
    for (size_t i = 0; i < W.items.size();) {
        const RecItem &A = W.items[i];
        const ShotDev &S = W.host[A.shot];
        auto win = [&](const RecItem &R, int &z0, int &z1, int &x0, int &x1) {
            z0 = std::max(R.ix1 - 10, 0); z1 = std::min(R.ix1 + 10, S.n1e - 1);
            x0 = std::max(R.ix2 - 10, 0); x1 = std::min(R.ix2 + 10, S.n2e - 1);
        };
        int z0, z1, x0, x1;
        win(A, z0, z1, x0, x1);
        size_t k = i + 1;
        for (; k < W.items.size() && k - i < 64; ++k) {
            const RecItem &B = W.items[k], &P = W.items[k - 1];
            if (B.shot != A.shot || B.cstride != P.cstride || B.coefT != P.coefT + 1) break;
            int a0, a1, b0, b1;
            win(B, a0, a1, b0, b1);
            const int nz0 = std::min(z0, a0), nz1 = std::max(z1, a1), nx0 = std::min(x0, b0), nx1 = std::max(x1, b1);
            if (nz1 - nz0 + 1 > kRecRows || nx1 - nx0 + 1 > kRecCols) break;
            z0 = nz0; z1 = nz1; x0 = nx0; x1 = nx1;
        }
        W.groups.push_back(RecGroup{(int)i, (int)(k - i), z0, z1 - z0 + 1, x0, x1 - x0 + 1});
        i = k;
    }
    const size_t gbytes = std::max<size_t>(W.groups.size(), 1) * sizeof(RecGroup);
    if (gbytes > W.d_groups.bytes) {
        W.clear_graphs();
        if (int rc = W.d_groups.alloc(std::max(gbytes, 2 * W.d_groups.bytes))) return rc;
    }
    if (!W.groups.empty())
        HIPCHK(hipMemcpyAsync(W.d_groups.p, W.groups.data(), W.groups.size() * sizeof(RecGroup), hipMemcpyHostToDevice,
                              g.stream));
    return 0;
}

// Tile lists of every configured mode, sizes, and the shot descriptors on the device.
static int finalize_sweepset(SweepSet &W)
{
    W.clear_graphs();
    W.nt_max = 0;
    W.pts_per_step = 0;
    for (int s = 0; s < W.nshots; ++s) {
        W.nt_max = std::max(W.nt_max, W.host[s].nt);
        W.pts_per_step += (long long)W.host[s].n1e * W.host[s].n2e;
    }
    for (int m = 0; m < 5; ++m) {
        // modes 0-2, then the exact adjoint's tile (the back-propagation's),
        // then the one-step march of a two-step re-injection's odd last step
        if (m == 4 && W.cfg[kReinject].march != 2) continue;
        const LaunchCfg c = m < 3 ? W.cfg[m] : m == 3 ? adjoint_cfg(W.cfg[kBack]) : march2_tail_cfg(W.cfg[kReinject]);
        const int order = m < 3 ? m : m == 3 ? kBack : kReinject;
        const std::vector<int> key = tiles_key(c, order);
        if (W.tiles.count(key)) continue;
        std::vector<int> tps(W.nshots);
        std::vector<std::vector<int>> cls(W.nshots);
        if (c.march == 2) {
            // per shot: row segments times groups of four 52-column strips;
            // entry (shot, z0 << 16 | rows << 8 | up << 7 | group), odd
            // segments walking up
            std::vector<std::vector<int>> ys(W.nshots);
            for (int s = 0; s < W.nshots; ++s) {
                const ShotDev &S = W.host[s];
                const int groups = ((S.n2e + kMarch2Cols - 1) / kMarch2Cols + 3) / 4;
                if (groups > 128 || S.n1e >= 32768) return fail(DAW_ERR_ARG, "grid %d x %d too large for the column march", S.n1e, S.n2e);
                const std::vector<int2> segs = march2_segments(S.n1e, c.tz, S.npml, S.ofs);
                for (size_t i = 0; i < segs.size(); ++i)
                    for (int q = 0; q < groups; ++q) {
                        ys[s].push_back((segs[i].x << 16) | (segs[i].y << 8) | (int)(i & 1) << 7 | q);
                        cls[s].push_back(0);
                    }
                tps[s] = (int)ys[s].size();
            }
            std::vector<int2> tl = make_tiles(tps, cls, kTileCostFirst);
            for (int2 &e : tl) e.y = ys[e.x][e.y];
            TileList &T = W.tiles[key];
            T.n = (int)tl.size();
            if (int rc = upload(T.d, tl)) return rc;
            continue;
        }
        if (c.march) {
            // per shot: row segments (the back-propagation's receiver rows in
            // short ones) times groups of four strips, strips fastest; entry
            // (shot, z0 << 16 | rows << 8 | up << 7 | group).  Odd segments
            // walk bottom to top, so two neighbouring segments read the rows
            // of their common ring at the same point of their walks (both
            // first or both last) and the second read is an L2 hit.
            // The short segments of the injection band launch first within
            // each XCD's run: their wavefronts carry the receiver terms'
            // dependent loads, the longest chains of the sweep.
            std::vector<std::vector<int>> ys(W.nshots);
            for (int s = 0; s < W.nshots; ++s) {
                const ShotDev &S = W.host[s];
                const int2 b = order == kBack && s < (int)W.inj_band.size() ? W.inj_band[s] : make_int2(0, -1);
                const int groups = ((S.n2e + kMarchCols - 1) / kMarchCols + 3) / 4;
                if (groups > 128 || S.n1e >= 32768) return fail(DAW_ERR_ARG, "grid %d x %d too large for the column march", S.n1e, S.n2e);
                const std::vector<int2> segs = march_segments(S.n1e, c.tz, b.x, b.y);
                for (size_t i = 0; i < segs.size(); ++i)
                    for (int q = 0; q < groups; ++q) {
                        ys[s].push_back((segs[i].x << 16) | (segs[i].y << 8) | (int)(i & 1) << 7 | q);
                        cls[s].push_back(segs[i].y == c.tz ? 0 : 1);
                    }
                tps[s] = (int)ys[s].size();
            }
            std::vector<int2> tl = make_tiles(tps, cls, kTileCostFirst);
            for (int2 &e : tl) e.y = ys[e.x][e.y];
            TileList &T = W.tiles[key];
            T.n = (int)tl.size();
            if (int rc = upload(T.d, tl)) return rc;
            continue;
        }
        const int txt = TX * c.txw;
        for (int s = 0; s < W.nshots; ++s) {
            const ShotDev &S = W.host[s];
            const int ntx = (S.n2e + txt - 1) / txt;
            tps[s] = ((S.n1e + c.tz - 1) / c.tz) * ntx;
            cls[s].resize(tps[s]);
            // first and last row of the receivers' windows (none: z1 < z0)
            const int2 b = order == kBack && s < (int)W.inj_band.size() ? W.inj_band[s] : make_int2(0, -1);
            for (int t = 0; t < tps[s]; ++t) {
                const int z0 = (t / ntx) * c.tz, x0 = (t % ntx) * txt;
                const int rows = std::min(c.tz, S.n1e - z0), cols = std::min(txt, S.n2e - x0);
                const bool band = b.x <= z0 + c.tz + 1 && b.y >= z0 - 1;     // pressure rows z0-1 .. z0+tz+1
                cls[s][t] = (2 * rows < c.tz || 2 * cols < txt) ? 0 : band ? 2 : 1;
            }
        }
        std::vector<int2> tl = make_tiles(tps, cls, tile_order(order));
        TileList &T = W.tiles[key];
        T.n = (int)tl.size();
        if (int rc = upload(T.d, tl)) return rc;
    }
    return upload(W.d_shots, W.host);
}

static int refresh_shots(SweepSet &W)
{
    HIPCHK(hipMemcpyAsync(W.d_shots.p, W.host.data(), W.host.size() * sizeof(ShotDev), hipMemcpyHostToDevice, g.stream));
    return 0;
}

// The launches of one sweep: zero the read half of the state, then nsteps
// step launches and the trailing record.
static void launch_sweep(int mode, SweepSet &W, const RecItem *d_rec, int nrec, int nsteps)
{
    const LaunchCfg &c = W.cfg[mode];
    const TileList &T = W.tiles.at(tiles_key(c, mode));
    for (int s = 0; s < W.nshots; ++s) {
        const ShotDev &S = W.host[s];
        const size_t pe = (size_t)S.n1e * S.ldx;
        launch_fill(S.px[0], pe, 0.0f, g.stream);
        launch_fill(S.pz[0], pe, 0.0f, g.stream);
        launch_fill(S.vx[0], pe, 0.0f, g.stream);
        launch_fill(S.vz[0], pe, 0.0f, g.stream);
    }
    StepParams P;
    P.shots = W.d_shots.as<ShotDev>();
    P.tiles = T.d.as<int2>();
    P.rec = d_rec;
    P.nrec = nrec;
    const int per = steps_per_launch(c);
    int cur = 0;
    if (c.rec_batch) {
        // launch k advances steps it .. it+nsub-1; the record of the steps
        // since the last record runs as one batched launch before the band's
        // time slots would be reused, and after the last step
        P.rec_nsub = 0;
        P.nrec_blocks = 0;
        int rec_from = 0;
        for (int it = 0; it < nsteps; it += per) {
            const int n = std::min(per, nsteps - it);
            P.ntiles = T.n;
            P.it = it;
            P.nsub = n;
            P.cur = cur;
            if (c.march == 2) {
                if (n == 2) {
                    launch_march2(P, g.stream);
                } else {                                   // an odd last step
                    const LaunchCfg c1 = march2_tail_cfg(c);
                    const TileList &T1 = W.tiles.at(tiles_key(c1, kReinject));
                    StepParams P1 = P;
                    P1.tiles = T1.d.as<int2>();
                    P1.ntiles = T1.n;
                    launch_march(mode, P1, c1.tz, g.stream);
                }
            }
            else if (c.march) launch_march(mode, P, c.tz, g.stream);
            else if (per == 2) launch_step2(mode, W.snap, P, c.tz, c.txw, g.stream);
            else launch_step(mode, W.snap, P, c.tz, c.txw, g.stream);
            cur ^= 1;
            const int done = it + n;
            if (nrec > 0 && (done == nsteps || done - rec_from + per > kBandSlots)) {
                std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
                if (g.timing && !g.graph && hipEventCreate(&ev.first) == hipSuccess &&
                    hipEventCreate(&ev.second) == hipSuccess)
                    (void)hipEventRecord(ev.first, g.stream);
                launch_rec_batch(P.shots, d_rec, W.d_groups.as<RecGroup>(), (int)W.groups.size(), rec_from,
                                 done - rec_from, g.stream);
                if (ev.second) {
                    (void)hipEventRecord(ev.second, g.stream);
                    g.rec_ev.push_back(ev);
                } else if (ev.first) {
                    (void)hipEventDestroy(ev.first);
                }
                rec_from = done;
            }
        }
        return;
    }
    // record blocks inside the step launches: launch k advances steps it ..
    // it+nsub-1 and records the steps of launch k-1 (its first workgroups)
    int prev_it = 0, prev_n = 0;
    for (int it = 0;; it += per) {
        const bool more = it < nsteps;
        const int n = more ? std::min(per, nsteps - it) : 0;
        if (!more && (nrec == 0 || prev_n == 0)) break;
        P.ntiles = more ? T.n : 0;
        P.it = it;
        P.nsub = n;
        P.cur = cur;
        P.rec_it = prev_it;
        P.rec_nsub = nrec > 0 ? prev_n : 0;
        P.nrec_blocks = (nrec * P.rec_nsub + kRecPerBlock - 1) / kRecPerBlock;
        if (c.march) launch_march(mode, P, c.tz, g.stream);
        else if (per == 2) launch_step2(mode, W.snap, P, c.tz, c.txw, g.stream);
        else launch_step(mode, W.snap, P, c.tz, c.txw, g.stream);
        cur ^= 1;
        prev_it = it;
        prev_n = n;
        if (!more) break;
    }
}

// Runs one sweep of nt_max (or nsteps_limit) steps over the batch.  With
// graphs on, a full sweep is captured once per batch and launch shape and
// replayed: the shot descriptors and record items it reads are refreshed in
// place before each replay.
static int run_sweep(int mode, SweepSet &W, const RecItem *d_rec, int nrec, int nsteps_limit = 0)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g.timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, g.stream));
    }
    const int nsteps = nsteps_limit > 0 ? std::min(nsteps_limit, W.nt_max) : W.nt_max;
    if (g.graph && nsteps_limit == 0) {
        const LaunchCfg &c = W.cfg[mode];
        const std::vector<long long> key{mode, nsteps, nrec, steps_per_launch(c), c.tz, c.txw, c.march, c.rec_batch, W.snap,
                                         (long long)(intptr_t)d_rec, (long long)(intptr_t)W.tiles.at(tiles_key(c, mode)).d.p,
                                         (long long)(intptr_t)W.d_shots.p};
        auto hit = W.graphs.find(key);
        hipGraphExec_t ex = nullptr;
        if (hit == W.graphs.end()) {
            HIPCHK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
            launch_sweep(mode, W, d_rec, nrec, nsteps);
            const hipError_t el = hipGetLastError();
            hipGraph_t gr = nullptr;
            const hipError_t ec = hipStreamEndCapture(g.stream, &gr);
            if (el != hipSuccess || ec != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return fail(DAW_ERR_HIP, "sweep capture: %s", hipGetErrorString(el != hipSuccess ? el : ec));
            }
            const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ei != hipSuccess) return fail(DAW_ERR_HIP, "sweep graph instantiate: %s", hipGetErrorString(ei));
            W.graphs[key] = ex;
        } else {
            ex = hit->second;
        }
        HIPCHK(hipGraphLaunch(ex, g.stream));
    } else {
        launch_sweep(mode, W, d_rec, nrec, nsteps);
    }
    HIPCHK(hipGetLastError());
    if (g.timing) {
        HIPCHK(hipEventRecord(e1, g.stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        g.t_ms[mode] += ms;
        g.t_cnt[mode] += nsteps;
        for (auto &ev : g.rec_ev) {
            float r = 0;
            if (hipEventElapsedTime(&r, ev.first, ev.second) == hipSuccess) g.t_rec_ms[mode] += r;
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
        g.rec_ev.clear();
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    g.pts += W.pts_per_step * nsteps;
    const int per = steps_per_launch(W.cfg[mode]);
    g.launches += (nsteps + per - 1) / per;   // step launches (the batched records not counted)
    return 0;
}

// Reverse-time sweep of the exact adjoint step (k_adj), it = nt_max-1 .. 0;
// the trailing record of each launch reads the band of the step before it.
static int run_adjoint_sweep(bool hess, SweepSet &W, const RecItem *d_rec, int nrec)
{
    for (int s = 0; s < W.nshots; ++s) {
        const ShotDev &S = W.host[s];
        const size_t pe = (size_t)S.n1e * S.ldx;
        launch_fill(S.px[0], pe, 0.0f, g.stream);
        launch_fill(S.pz[0], pe, 0.0f, g.stream);
        launch_fill(S.vx[0], pe, 0.0f, g.stream);
        launch_fill(S.vz[0], pe, 0.0f, g.stream);
    }
    const LaunchCfg c = adjoint_cfg(W.cfg[kBack]);      // the one-step tile of the back-propagation
    const TileList &T = W.tiles.at(tiles_key(c, kBack));
    StepParams P;
    P.shots = W.d_shots.as<ShotDev>();
    P.tiles = T.d.as<int2>();
    P.rec = d_rec;
    const int rec_blocks = (nrec + kRecPerBlock - 1) / kRecPerBlock;
    int cur = 0;
    P.nrec = nrec;
    P.nsub = 1;
    for (int k = 0; k <= W.nt_max; ++k) {
        const int it = W.nt_max - 1 - k;
        P.ntiles = (k < W.nt_max) ? T.n : 0;
        P.it = it < 0 ? 0 : it;
        P.cur = cur;
        const bool rec_prev = k > 0 && nrec > 0;
        P.rec_nsub = rec_prev ? 1 : 0;
        P.nrec_blocks = rec_prev ? rec_blocks : 0;
        P.rec_it = it + 1;
        launch_adjoint(hess, P, c.tz, c.txw, g.stream);
        cur ^= 1;
    }
    HIPCHK(hipGetLastError());
    g.pts += W.pts_per_step * W.nt_max;
    g.launches += W.nt_max + 1;
    return 0;
}

// Per-shot device storage of one batch.
// Snapshot store: sp, svx, svz(n1,n2,nt) per shot, FP32 or 16-bit with one
// power-of-two FP32 scale per 64-column segment of each snapshot row.
struct ShotStore {
    DevBuf fields;   // 8 state arrays per shot
    DevBuf band;     // 2 slots per shot
    DevBuf snap;     // FP32 snapshots
    DevBuf snap16;   // 16-bit snapshots
    DevBuf scale;    // segment scales of the 16-bit snapshots
    int fmt = -1;    // SnapFormat, -1: no snapshots
    bool has_snap() const { return fmt >= 0; }
};

static int snap_segments(const Grid &G) { return (G.n2e + TX - 1) / TX; }

// Device bytes of one shot's snapshot store in format fmt.
static size_t snap_bytes(const Grid &G, int nt, int fmt)
{
    if (fmt < 0) return 0;
    if (fmt == kSnapF32) return 3 * G.plane_s() * nt * sizeof(float);
    return 3 * G.plane_s() * nt * 2 + 3 * (size_t)nt * G.n1 * snap_segments(G) * sizeof(float);
}

static int alloc_store(ShotStore &st, int nshots, const Grid &G, int nslot, int nt, int fmt)
{
    const size_t pe = G.plane_e();
    if (int rc = st.fields.alloc((size_t)nshots * 8 * pe * sizeof(float))) return rc;
    if (int rc = st.band.alloc((size_t)nshots * kBandSlots * std::max(nslot, 1) * G.ldx * sizeof(float))) return rc;
    st.fmt = fmt;
    if (fmt == kSnapF32) {
        if (int rc = st.snap.alloc((size_t)nshots * 3 * G.plane_s() * nt * sizeof(float))) return rc;
    } else if (fmt > 0) {
        if (int rc = st.snap16.alloc((size_t)nshots * 3 * G.plane_s() * nt * 2)) return rc;
        if (int rc = st.scale.alloc((size_t)nshots * 3 * nt * G.n1 * snap_segments(G) * sizeof(float))) return rc;
    }
    return 0;
}

static void fill_shot(ShotDev &S, const Grid &G, const DevModel &M, ShotStore &st, int s, int nt, int nslot,
                      const Band *band)
{
    std::memset(&S, 0, sizeof S);
    S.n1 = G.n1; S.n2 = G.n2; S.npml = G.npml; S.n1e = G.n1e; S.n2e = G.n2e; S.ldx = G.ldx; S.nt = nt; S.ofs = G.ofs;
    S.dt = G.dt;
    S.dtdx = G.dt / G.h;
    S.kappa = M.kappa.as<float>();
    S.d1a = M.d1a(); S.d1a1 = M.d1a1(); S.d1b = M.d1b(); S.bw1 = M.bw1();
    S.d2a = M.d2a(); S.d2a1 = M.d2a1(); S.d2b = M.d2b(); S.bw2 = M.bw2();
    const size_t pe = G.plane_e();
    float *f = st.fields.as<float>() + (size_t)s * 8 * pe;
    S.px[0] = f; S.pz[0] = f + pe; S.vx[0] = f + 2 * pe; S.vz[0] = f + 3 * pe;
    S.px[1] = f + 4 * pe; S.pz[1] = f + 5 * pe; S.vx[1] = f + 6 * pe; S.vz[1] = f + 7 * pe;
    if (band) {
        S.band_row = band->d_row.as<int>();
        S.band_stride = std::max(nslot, 1) * G.ldx;
        S.pband = st.band.as<float>() + (size_t)s * kBandSlots * S.band_stride;
    }
    const size_t pn = G.plane_s() * nt;
    S.spitch = G.spitch();
    S.sxb = G.sxb();
    if (st.fmt == kSnapF32) {
        float *sn = st.snap.as<float>() + (size_t)s * 3 * pn;
        S.sp = sn; S.svx = sn + pn; S.svz = sn + 2 * pn;
    } else if (st.fmt > 0) {
        unsigned short *sn = static_cast<unsigned short *>(st.snap16.p) + (size_t)s * 3 * pn;
        S.sp16 = sn; S.svx16 = sn + pn; S.svz16 = sn + 2 * pn;
        S.nseg = snap_segments(G);
        S.ssc = st.scale.as<float>() + (size_t)s * 3 * nt * G.n1 * S.nseg;
    }
    S.xlo1 = G.npml; S.xhi1 = G.npml + G.n1 - 1;
    S.xlo2 = G.npml; S.xhi2 = G.npml + G.n2 - 1;
}

// amp[r*amp_stride + it*amp_tstride]: amp_tstride 1 for [r][t] traces, the
// receiver count for a transposed [t][r] copy
static void set_inject(ShotDev &S, const Inject &I, const float *amp, int amp_stride, int vinj, int amp_tstride = 1)
{
    S.inj_row = I.d_row.as<int>();
    S.inj_z0 = I.z0;
    S.inj_z1 = I.z1;
    S.inj_ptr = I.d_ptr.as<int>();
    S.inj_r = I.d_rid.as<int>();
    S.inj_cdt = I.d_cdt.as<float>();
    S.inj_amp = amp;
    S.amp_stride = amp_stride;
    S.amp_tstride = amp_tstride;
    S.vinj = vinj;
}

static void add_records(std::vector<RecItem> &v, const PointSet &P, int shot, int clip, int reverse, const DevBuf &coef,
                        const DevBuf &coefT, float *out, int nt)
{
    for (int r = 0; r < P.n; ++r) {
        RecItem R;
        R.shot = shot; R.ix1 = P.ix1[r]; R.ix2 = P.ix2[r]; R.clip = clip; R.reverse = reverse;
        R.cstride = P.n;
        R.coef = coef.as<float>() + (size_t)r * 441;
        R.coefT = coefT.as<float>() + r;
        R.out = out + (size_t)r * nt;
        v.push_back(R);
    }
}

// ------------------------------------------------------------------ model cache
struct ModelKey {
    std::vector<float> vp;
    int n1, n2, npml, ofs;
    float h, dt;
    bool operator==(const ModelKey &o) const
    {
        return n1 == o.n1 && n2 == o.n2 && npml == o.npml && ofs == o.ofs && h == o.h && dt == o.dt && vp == o.vp;
    }
};
static std::unique_ptr<DevModel> g_model;
static ModelKey g_model_key;

static int get_model(const DevModel *&out, const Grid &G, const float *vp)
{
    ModelKey k;
    k.vp.assign(vp, vp + (size_t)G.n1 * G.n2);
    k.n1 = G.n1; k.n2 = G.n2; k.npml = G.npml; k.ofs = G.ofs; k.h = G.h; k.dt = G.dt;
    if (!g_model || !(k == g_model_key)) {
        auto M = std::make_unique<DevModel>();
        if (int rc = make_model(*M, G, vp)) return rc;
        g_model = std::move(M);
        g_model_key = std::move(k);
    }
    out = g_model.get();
    return 0;
}

static int check_hicks()
{
    if (g.hicks_ofs == -1000) return fail(DAW_ERR_STATE, "daw_kwsinc2d() must be called before the propagators (main_redatum_DAW.f90:92-93)");
    return 0;
}

// ------------------------------------------------------------------ batches
// Geometry of shots: grid, model, nt and acquisition tables.  An even batch
// (the reference's case: one model, one receiver table, one nt) has one
// geometry shared by all shots; an uneven batch has one per shot (its cropped
// model, its receivers, its nt).
struct Geo {
    Grid G;
    const DevModel *M = nullptr;
    std::unique_ptr<DevModel> own;  // a cropped model owned by this geometry
    int nt = 0;
    PointSet rec, dat;
    Inject inj_rec;
    Inject inj_rec_c;               // receivers with coefficient c (adjoint data spreading, R^T)
    Band band;
    DevBuf d_rec_coef, d_dat_coef;
    DevBuf d_rec_coefT, d_dat_coefT;   // transposed (batched record)
};

struct Batch {
    int nshots = 0, nt = 0;         // nt: the longest shot
    DevBuf ampT;                    // [t][r] copies of the injected traces (roff layout)
    int max_nrec = 1;
    std::vector<std::unique_ptr<Geo>> geo;
    std::vector<int> gi;            // geometry of each shot
    std::vector<Inject> inj_src;    // one per shot
    std::vector<ShotStore> store;   // one per shot
    // offsets of each shot's acquisition traces (nrec x nt) and datum traces
    // in the batch's trace blocks; the last entry is the total
    std::vector<size_t> roff, doff;
    SweepSet W;
    const Geo &geo_of(int s) const { return *geo[gi[s]]; }
};

// Hessian-type two-pass operator.  Pass 1 (kBack) back-propagates the
// per-shot amplitudes amp[roff[s] + r*nt + t] and stores snapshots; pass 2
// (kReinject) re-injects them time reversed and records the acquisition
// receivers (and optionally the datum receivers).
struct HessOut {
    float *sis = nullptr;      // acquisition traces, shot s at sis + roff[s]
    int reverse = 0;
    float *datum = nullptr;    // datum traces, shot s at datum + doff[s], or null
    float *wf = nullptr;       // single shot only: full wavefield (Fortran layout)
    int variant = 0;           // 0: esource_data ranges, 1: esource ranges (2..n-1)
    int record = 1;            // type_rec == 0
};

// Launch autotuner: for each sweep mode (forward, back-propagation,
// re-injection) time a short probe (zero amplitudes, best of three) at every
// compiled tile shape and steps-per-launch choice and keep the fastest for
// this batch shape.  DAW_TUNE_LOG=1 prints the probe times to stderr.
static int autotune_tiles(Batch &B)
{
    SweepSet &W = B.W;
    if (!g.tz_auto && !g.steps_auto) return 0;
    // candidates: {steps, rows, wavefronts across, march}; fixed settings are kept
    static const int shapes[][4] = {{1, 8, 2, 0},  {1, 16, 1, 0}, {1, 16, 2, 0}, {1, 20, 1, 0}, {1, 24, 1, 0},
                                    {1, 32, 1, 0}, {1, 32, 2, 0}, {1, 64, 1, 0}, {2, 16, 1, 0}, {2, 32, 1, 0},
                                    {2, 16, 2, 0}, {1, 8, 0, 1},  {1, 12, 0, 1}, {1, 16, 0, 1}, {1, 24, 0, 1},
                                    {1, 32, 0, 1}, {2, 16, 0, 2}, {2, 24, 0, 2}, {2, 32, 0, 2},
                                    {2, 48, 0, 2}};
    std::vector<LaunchCfg> cand;
    for (const auto &sh : shapes) {
        if (!g.steps_auto && sh[0] != g.steps) continue;
        if (!g.tz_auto && (sh[1] != g.tz || sh[2] != g.txw || sh[3] != g.march)) continue;
        cand.push_back({sh[0], sh[1], sh[2], sh[3], g.rec_batch});
    }
    if (cand.empty() || B.store.empty() || !B.store[0].has_snap()) return 0;
    const std::vector<int> key{(int)(W.pts_per_step >> 10), B.geo_of(0).G.n2e, B.nshots, g.steps_auto ? 0 : g.steps,
                               g.tz_auto ? 0 : g.tz, g.tz_auto ? 0 : g.txw, g.tz_auto ? 0 : g.march, W.snap, g.rec_batch};
    auto hit = g.tz_cache.find(key);
    if (hit == g.tz_cache.end()) {
        // zero amplitudes laid out as the real sweeps read them (the
        // back-propagation's [t][r] copy): the probes carry the gathers' cost
        DevBuf zeros;
        const size_t nz = (size_t)B.nt * B.max_nrec;
        if (int rc = zeros.alloc(nz * sizeof(float))) return rc;
        launch_fill(zeros.as<float>(), nz, 0.0f, g.stream);
        for (int s = 0; s < B.nshots; ++s) {
            set_inject(W.host[s], B.geo_of(s).inj_rec, zeros.as<float>(), 1, 0, B.geo_of(s).rec.n);
            W.host[s].amp_rev = 0;
        }
        if (int rc = refresh_shots(W)) return rc;
        // The probes carry the record blocks the real sweeps carry: the
        // acquisition receivers in the re-injection (the Hessian's data
        // record), acquisition and datum receivers in the forward sweep.  They
        // cost 12-36 % of a step and shift which tile wins.
        const int ns = B.nshots;
        DevBuf rec_out, dat_out;
        if (int rc = rec_out.alloc(std::max<size_t>(B.roff[ns], 1) * sizeof(float))) return rc;
        if (int rc = dat_out.alloc(std::max<size_t>(B.doff[ns], 1) * sizeof(float))) return rc;
        std::vector<RecItem> it_reinj, it_fwd;
        for (int s = 0; s < ns; ++s) {
            const Geo &Q = B.geo_of(s);
            add_records(it_reinj, Q.rec, s, 0, 1, Q.d_rec_coef, Q.d_rec_coefT, rec_out.as<float>() + B.roff[s], Q.nt);
            add_records(it_fwd, Q.rec, s, 0, 0, Q.d_rec_coef, Q.d_rec_coefT, rec_out.as<float>() + B.roff[s], Q.nt);
            if (Q.dat.n > 0)
                add_records(it_fwd, Q.dat, s, 1, 0, Q.d_dat_coef, Q.d_dat_coefT, dat_out.as<float>() + B.doff[s], Q.nt);
        }
        const bool timing = g.timing;
        const long long pts = g.pts, launches = g.launches;
        g.timing = false;
        // Probes of 192 steps, three rounds over all candidates (round-robin, so
        // that clock drift during the tuning does not favour the candidates
        // timed first or last), median of the three per candidate.  (64-step
        // probes tied 16x64 and 20x64 back-propagation at 201x801, 24 shots,
        // where full sweeps differ by 4 %.)
        const int probe = std::min(192, B.nt);
        const bool log = getenv("DAW_TUNE_LOG") != nullptr;
        static const char *names[3] = {"forward", "back-propagation", "re-injection"};
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        std::vector<int> best(15);
        constexpr int kRounds = 3;
        for (int mode : {kBack, kReinject, kFwd}) {
            if (mode == kFwd && B.inj_src.size() == (size_t)ns) {   // the forward sweep injects the sources
                for (int s = 0; s < ns; ++s) set_inject(W.host[s], B.inj_src[s], zeros.as<float>(), 0, 0);
                if (int rc = refresh_shots(W)) return rc;
            }
            const std::vector<RecItem> *items = mode == kBack ? nullptr : mode == kReinject ? &it_reinj : &it_fwd;
            if (items)
                if (int rc = set_items(W, *items)) return rc;
            const RecItem *d_items = items ? W.d_items.as<RecItem>() : nullptr;
            const int n_items = items ? (int)items->size() : 0;
            std::vector<float> t(cand.size() * kRounds, 1e30f);
            // rounds 2 and 3 time only the candidates within 15 % of round 1's
            // best (the rest keep their round-1 time as their median); the
            // column march has no 16-bit snapshot codec
            std::vector<char> live(cand.size(), 1);
            for (size_t k = 0; k < cand.size(); ++k)
                if (!march_ok(cand[k], mode, W.snap)) live[k] = 0;
            if (std::find(live.begin(), live.end(), 1) == live.end()) {
                const LaunchCfg d = fixed_cfg(mode, W.snap);
                best[5 * mode] = d.steps; best[5 * mode + 1] = d.tz; best[5 * mode + 2] = d.txw; best[5 * mode + 3] = d.march;
                best[5 * mode + 4] = d.rec_batch;
                continue;
            }
            for (int rep = 0; rep < kRounds; ++rep) {
                if (rep == 1) {
                    float b1 = 1e30f;
                    for (size_t k = 0; k < cand.size(); ++k) b1 = std::min(b1, t[k * kRounds]);
                    for (size_t k = 0; k < cand.size(); ++k)
                        if (t[k * kRounds] > 1.15f * b1) {
                            live[k] = 0;
                            for (int r = 1; r < kRounds; ++r) t[k * kRounds + r] = t[k * kRounds];
                        }
                }
                for (size_t k = 0; k < cand.size(); ++k) {
                    if (!live[k]) continue;
                    W.cfg[mode] = cand[k];
                    if (int rc = finalize_sweepset(W)) return rc;
                    if (int rc = run_sweep(mode, W, d_items, n_items, 4)) return rc;       // warm
                    HIPCHK(hipEventRecord(e0, g.stream));
                    if (int rc = run_sweep(mode, W, d_items, n_items, probe)) return rc;
                    HIPCHK(hipEventRecord(e1, g.stream));
                    HIPCHK(hipEventSynchronize(e1));
                    HIPCHK(hipEventElapsedTime(&t[k * kRounds + rep], e0, e1));
                }
            }
            std::vector<float> med(cand.size());
            float best_ms = 1e30f;
            for (size_t k = 0; k < cand.size(); ++k) {
                float *r = &t[k * kRounds];
                std::sort(r, r + kRounds);
                med[k] = r[kRounds / 2];
                if (log && med[k] < 1e29f)
                    fprintf(stderr, "[daw tune] %d shots %dx%d grid, %s: %d step(s)/launch, %s %dx%d  %.1f us/step\n",
                            B.nshots, B.geo_of(0).G.n1e, B.geo_of(0).G.n2e, names[mode], cand[k].steps,
                            cand[k].march == 2 ? "march2" : cand[k].march ? "march" : "tile", cand[k].tz, tile_cols(cand[k]),
                            1000.0f * med[k] / probe);
                best_ms = std::min(best_ms, med[k]);
            }
            // Among the candidates within 1.5 % of the fastest probe, the
            // smallest tile (most workgroups per launch) wins: at 201x801 with
            // 24 shots the probes tie 16x64 and 20x64 back-propagation, whose
            // full sweeps differ by 1-4 % in favour of 16x64.
            LaunchCfg pick = cand[0];
            int pick_area = 1 << 30;
            for (size_t k = 0; k < cand.size(); ++k) {
                const int area = cand[k].tz * tile_cols(cand[k]);
                if (med[k] <= 1.015f * best_ms && area < pick_area) { pick = cand[k]; pick_area = area; }
            }
            W.cfg[mode] = pick;
            best[5 * mode] = pick.steps; best[5 * mode + 1] = pick.tz; best[5 * mode + 2] = pick.txw;
            best[5 * mode + 3] = pick.march; best[5 * mode + 4] = pick.rec_batch;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        g.timing = timing;
        g.pts = pts;
        g.launches = launches;
        hit = g.tz_cache.emplace(key, best).first;
    }
    const std::vector<int> &b = hit->second;
    for (int m = 0; m < 3; ++m) W.cfg[m] = {b[5 * m], b[5 * m + 1], b[5 * m + 2], b[5 * m + 3], b[5 * m + 4]};
    for (int m = 0; m < 3; ++m) g.last_cfg[m] = W.cfg[m];
    return finalize_sweepset(W);
}

// Pressure extended-source range of the re-injection: 1..n (esource_data,
// submodeling2.f90:1121-1122) or 2..n-1 (esource, :777-778), extended coords.
static void set_range(ShotDev &S, const Grid &G, int variant)
{
    const int e = variant == 1 ? 1 : 0;
    S.xlo1 = G.npml + e; S.xhi1 = G.npml + G.n1 - 1 - e;
    S.xlo2 = G.npml + e; S.xhi2 = G.npml + G.n2 - 1 - e;
}

static int hessian(Batch &B, const float *amp, const HessOut &o)
{
    SweepSet &W = B.W;
    const bool exact = g.hess_exact;
    if (!exact && B.ampT.bytes < B.roff[B.nshots] * sizeof(float))
        if (int rc = B.ampT.alloc(B.roff[B.nshots] * sizeof(float))) return rc;
    for (int s = 0; s < B.nshots; ++s) {
        ShotDev &S = W.host[s];
        const Geo &Q = B.geo_of(s);
        S.wf = nullptr;
        if (exact) {
            // pass 1 = transpose of pass 2 applied to the time-reversed input
            // (the reference's res is time reversed, subresiduals.f90:14)
            set_inject(S, Q.inj_rec_c, amp + B.roff[s], Q.nt, 0);
            S.amp_rev = 1;
            set_range(S, Q.G, o.variant);
        } else {
            // a [t][r] copy of the shot's traces: the receivers' amplitudes
            // of one step side by side for the injection's gathers
            launch_transpose(amp + B.roff[s], B.ampT.as<float>() + B.roff[s], Q.rec.n, Q.nt, g.stream);
            set_inject(S, Q.inj_rec, B.ampT.as<float>() + B.roff[s], 1, 0, Q.rec.n);
            S.amp_rev = 0;
        }
    }
    if (int rc = refresh_shots(W)) return rc;
    if (exact) {
        if (int rc = run_adjoint_sweep(true, W, nullptr, 0)) return rc;
    } else {
        if (int rc = run_sweep(kBack, W, nullptr, 0)) return rc;
    }
    std::vector<RecItem> items;
    for (int s = 0; s < B.nshots; ++s) {
        ShotDev &S = W.host[s];
        const Geo &Q = B.geo_of(s);
        S.inj_row = nullptr;
        S.amp_rev = 0;
        S.wf = o.wf;
        set_range(S, Q.G, o.variant);
        if (o.record && o.sis) add_records(items, Q.rec, s, 0, o.reverse, Q.d_rec_coef, Q.d_rec_coefT, o.sis + B.roff[s], Q.nt);
        if (o.datum) add_records(items, Q.dat, s, 1, 0, Q.d_dat_coef, Q.d_dat_coefT, o.datum + B.doff[s], Q.nt);
    }
    if (!o.record) for (auto &S : W.host) S.wf = nullptr;
    if (int rc = set_items(W, items)) return rc;
    if (int rc = refresh_shots(W)) return rc;
    int rc = run_sweep(kReinject, W, W.d_items.as<RecItem>(), (int)items.size());
    HIPCHK(hipStreamSynchronize(g.stream));
    return rc;
}

// Forward sweep: sources of every shot, record acquisition (+ datum).
static int forward(Batch &B, const float *s_wavelet, int type_src, int record, float *sis, float *datum, float *wf)
{
    SweepSet &W = B.W;
    std::vector<RecItem> items;
    for (int s = 0; s < B.nshots; ++s) {
        ShotDev &S = W.host[s];
        const Geo &Q = B.geo_of(s);
        set_inject(S, B.inj_src[s], s_wavelet, 0, type_src == 1 ? 1 : 0);
        if (type_src != 0 && type_src != 1) S.inj_row = nullptr;
        S.wf = record ? wf : nullptr;
        if (record && sis) add_records(items, Q.rec, s, 0, 0, Q.d_rec_coef, Q.d_rec_coefT, sis + B.roff[s], Q.nt);
        if (datum) add_records(items, Q.dat, s, 1, 0, Q.d_dat_coef, Q.d_dat_coefT, datum + B.doff[s], Q.nt);
    }
    if (int rc = set_items(W, items)) return rc;
    if (int rc = refresh_shots(W)) return rc;
    int rc = run_sweep(kFwd, W, W.d_items.as<RecItem>(), (int)items.size());
    HIPCHK(hipStreamSynchronize(g.stream));
    return rc;
}

// The acquisition tables of a geometry: rec_tab (nrec,5), dat_tab (ndat,5) or null.
static int make_geo(Geo &Q, const int *rec_tab, int nrec, const int *dat_tab, int ndat)
{
    const Grid &G = Q.G;
    if (int rc = make_points(Q.rec, rec_tab, nrec, G, false)) return rc;
    if (dat_tab && ndat > 0)
        if (int rc = make_points(Q.dat, dat_tab, ndat, G, true)) return rc;
    if (int rc = make_inject(Q.inj_rec, Q.rec, G)) return rc;
    {
        Grid G1 = G;
        G1.dt = 1.0f;              // c*dt with dt = 1 is c
        if (int rc = make_inject(Q.inj_rec_c, Q.rec, G1)) return rc;
    }
    std::vector<const PointSet *> sets{&Q.rec};
    if (Q.dat.n) sets.push_back(&Q.dat);
    if (int rc = make_band(Q.band, sets, G)) return rc;
    if (int rc = upload_record_coef(Q.d_rec_coef, Q.d_rec_coefT, Q.rec.coef)) return rc;
    if (Q.dat.n)
        if (int rc = upload_record_coef(Q.d_dat_coef, Q.d_dat_coefT, Q.dat.coef)) return rc;
    return 0;
}

// Source injection of shot s from its src_tab row (5 entries at stride).
static int make_source(Batch &B, int s, const int *row5, int stride)
{
    int row[5];
    for (int k = 0; k < 5; ++k) row[k] = row5[(size_t)k * stride];
    PointSet P;
    if (int rc = make_points(P, row, 1, B.geo_of(s).G, false)) return rc;
    return make_inject(B.inj_src[s], P, B.geo_of(s).G);
}

// Storage, shot descriptors, trace offsets and tiles once geo/gi/inj_src are set.
static int finish_batch(Batch &B, bool want_snap)
{
    // the exact Hessian's transpose pass writes FP32 snapshots
    const int fmt = want_snap ? (g.hess_exact ? (int)kSnapF32 : g.snap_fmt) : -1;
    B.store.clear();
    B.store.resize(B.nshots);
    B.W.nshots = B.nshots;
    B.W.snap = fmt < 0 ? (int)kSnapF32 : fmt;
    for (int m = 0; m < 3; ++m) B.W.cfg[m] = g.last_cfg[m] = fixed_cfg(m, B.W.snap);   // the autotuner may replace them
    B.W.host.resize(B.nshots);
    B.W.inj_band.assign(B.nshots, make_int2(0, -1));
    B.roff.assign(B.nshots + 1, 0);
    B.doff.assign(B.nshots + 1, 0);
    B.nt = 0;
    int max_nrec = 1;
    for (int s = 0; s < B.nshots; ++s) {
        const Geo &Q = B.geo_of(s);
        if (int rc = alloc_store(B.store[s], 1, Q.G, Q.band.nslot, Q.nt, fmt)) return rc;
        fill_shot(B.W.host[s], Q.G, *Q.M, B.store[s], 0, Q.nt, Q.band.nslot, &Q.band);
        B.W.inj_band[s] = make_int2(Q.inj_rec.z0, Q.inj_rec.z1);
        B.roff[s + 1] = B.roff[s] + (size_t)Q.rec.n * Q.nt;
        max_nrec = std::max(max_nrec, Q.rec.n);
        B.doff[s + 1] = B.doff[s] + (size_t)Q.dat.n * Q.nt;
        B.nt = std::max(B.nt, Q.nt);
    }
    if (want_snap && B.roff[B.nshots] > 0)
        if (int rc = B.ampT.alloc(B.roff[B.nshots] * sizeof(float))) return rc;
    B.max_nrec = max_nrec;
    return finalize_sweepset(B.W);
}

// An even batch: nshots shots on one grid, model, receiver table and nt.
// src_rows5: (src_stride,5) with rows 0..nshots-1 used, or null.
static int build_batch(Batch &B, const Grid &G, const DevModel *M, int nt, int nshots, const int *src_rows5,
                       int src_stride, const int *rec_tab, int nrec, const int *dat_tab, int ndat, bool want_snap)
{
    B.nshots = nshots;
    B.geo.clear();
    B.geo.push_back(std::make_unique<Geo>());
    Geo &Q = *B.geo[0];
    Q.G = G;
    Q.M = M;
    Q.nt = nt;
    if (int rc = make_geo(Q, rec_tab, nrec, dat_tab, ndat)) return rc;
    B.gi.assign(nshots, 0);
    B.inj_src.clear();
    B.inj_src.resize(nshots);
    if (src_rows5)
        for (int s = 0; s < nshots; ++s)
            if (int rc = make_source(B, s, src_rows5 + s, src_stride)) return rc;
    return finish_batch(B, want_snap);
}

// ------------------------------------------------------------------ band-pass on device traces
// filmartin2d (minresModule.f90:486-614) for ntr traces [tr][t] in HBM, FP64
// transforms on the GPU for the whole FFT length list (nto <= 65536): in LDS
// up to 8192, as a four-step FFT through an HBM work area beyond.
static int filter_device(const float *din, float *dout, int ntr, int nt, float dt, float flow, float fhig, float llow,
                         float lhig)
{
    std::vector<double> w;
    const int nto = filter_design(nt, dt, flow, fhig, llow, lhig, w);
    if (nto < 0) return fail(DAW_ERR_ARG, "filmartin2d: nt=%d exceeds the 65536-sample FFT length list", nt);
    if (ntr <= 0) return 0;
    const bool lds = nto <= 8192;
    // twiddles exp(-2 pi i m / nto): m < nto/2 for the LDS transform, m < nto for the four-step one
    const int ntw = lds ? nto / 2 : nto;
    std::vector<double> tw((size_t)2 * ntw);
    for (int m = 0; m < ntw; ++m) {
        const double ang = 2.0 * M_PI * (double)m / (double)nto;
        tw[2 * (size_t)m] = std::cos(ang);
        tw[2 * (size_t)m + 1] = -std::sin(ang);
    }
    w.resize(nto / 2 + 1);
    DevBuf d_w, d_tw, work;
    if (int rc = upload(d_w, w)) return rc;
    if (int rc = upload(d_tw, tw)) return rc;
    if (lds) {
        launch_filter(din, dout, ntr, nt, nto, d_w.as<double>(), d_tw.as<double2>(), g.stream);
    } else {
        // work area of up to ~256 MB: trace pairs in groups
        const int npairs = (ntr + 1) / 2;
        const int group = std::max(1, std::min(npairs, (int)((256u << 20) / ((size_t)nto * sizeof(double2)))));
        if (int rc = work.alloc((size_t)group * nto * sizeof(double2))) return rc;
        if (launch_filter_long(din, dout, ntr, nt, nto, d_w.as<double>(), d_tw.as<double2>(), work.as<double2>(), group,
                               g.stream) != 0)
            return fail(DAW_ERR_ARG, "filmartin2d: FFT length %d", nto);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g.stream));
    return 0;
}

// ------------------------------------------------------------------ MINRES over a batch
struct MinresWork {
    DevBuf vec;     // x, v, Y[3], Wv[3], hb: 9 vectors per shot
    DevBuf off, len, nts, part0, part1, scal, hist, trace;
    DevBuf coff, q16, s16;   // 16-bit Krylov vectors: chunk offsets, values, chunk scales
};

// Per-shot MINRES log block handed to callers: kLogHead doubles (fac, Anorm,
// Acond, rnorm, Arnorm, ynorm, beta1, istop, itn) and then kTraceStride per
// iteration up to itnlim (x(1), qrnorm/beta1, Anorm, Acond, printed).
constexpr int kLogHead = 9;
static size_t log_block(int itnlim) { return kLogHead + (size_t)kTraceStride * std::max(itnlim, 1); }

static void scal_info(const double *q, double *info)
{
    info[0] = q[kSFac]; info[1] = q[kSAnorm]; info[2] = q[kSAcond]; info[3] = q[kSRnorm];
    info[4] = q[kSArnorm]; info[5] = q[kSYnorm]; info[6] = q[kSBeta1]; info[7] = q[kSIstop];
    info[8] = q[kSItn];
}

// Solves (H + fac I) x = b per shot as sublinearminres/MINRES do
// (sublinearminres.f90:64-76, minresModule.f90:83-329).  b and x are
// device blocks [shot][nrec][nt].
static int minres_batch(Batch &B, const float *b, float *x, float mu, int itnlim, double rtol, double *hist_host,
                        double *scal_host, double *log_host = nullptr)
{
    const int ns = B.nshots;
    const size_t blk = B.roff[ns];       // all shots' vectors, shot s at roff[s]
    const int nb = minres_vec_blocks(), nsc = minres_scal_stride();
    MinresWork w;
    if (int rc = w.vec.alloc(8 * blk * sizeof(float))) return rc;
    std::vector<size_t> off(B.roff.begin(), B.roff.begin() + ns);
    std::vector<int> len(ns);
    for (int s = 0; s < ns; ++s) len[s] = (int)(B.roff[s + 1] - B.roff[s]);
    if (int rc = upload(w.off, off)) return rc;
    if (int rc = upload(w.len, len)) return rc;
    if (int rc = w.part0.alloc((size_t)ns * nb * sizeof(double))) return rc;
    if (int rc = w.part1.alloc((size_t)ns * nb * sizeof(double))) return rc;
    if (int rc = w.scal.alloc((size_t)ns * nsc * sizeof(double))) return rc;
    if (int rc = w.hist.alloc((size_t)ns * std::max(itnlim, 1) * sizeof(double))) return rc;
    const bool ref32 = g.minres_arith == 1;
    std::vector<int> nts(ns);
    for (int s = 0; s < ns; ++s) nts[s] = B.geo_of(s).nt;
    if (int rc = upload(w.nts, nts)) return rc;
    const int *dnts = w.nts.as<int>();
    float *base = w.vec.as<float>();
    float *v = base, *Y[3] = {base + blk, base + 2 * blk, base + 3 * blk};
    float *Wv[3] = {base + 4 * blk, base + 5 * blk, base + 6 * blk};
    float *hb = base + 7 * blk;
    const size_t *doff = w.off.as<size_t>();
    const int *dlen = w.len.as<int>();
    launch_fill(w.hist.as<float>(), (size_t)ns * std::max(itnlim, 1) * 2, 0.0f, g.stream);
    const size_t ntrace = (size_t)ns * std::max(itnlim, 1) * kTraceStride;
    if (log_host) {
        if (int rc = w.trace.alloc(ntrace * sizeof(double))) return rc;
        launch_fill(w.trace.as<float>(), ntrace * 2, 0.0f, g.stream);
    }
    double *trace = log_host ? w.trace.as<double>() : nullptr;

    // fac from ||b|| / ||H b|| (sublinearminres.f90:64-72)
    HessOut o;
    o.sis = hb; o.reverse = 0; o.variant = 0;
    if (int rc = hessian(B, b, o)) return rc;
    double *dscal = w.scal.as<double>();
    if (ref32) {
        // datanorm of b and H b in its t-outer order, then b.b as dot_product sums it
        launch_seqdot32(b, b, doff, dlen, dnts, dscal, 0, 1, false, ns, g.stream);
        launch_seqdot32(hb, hb, doff, dlen, dnts, dscal, 1, 1, false, ns, g.stream);
        launch_seqdot32(b, b, doff, dlen, dnts, dscal, 2, 0, false, ns, g.stream);
        launch_minres32_init(dscal, mu, ns, g.stream);
    } else {
        launch_dot(b, b, doff, dlen, w.part0.as<double>(), ns, g.stream);
        launch_dot(hb, hb, doff, dlen, w.part1.as<double>(), ns, g.stream);
        launch_minres_init(w.part0.as<double>(), w.part1.as<double>(), dscal, mu, ns, g.stream);
    }
    // 16-bit Krylov vectors (FP64 path only): R16[0..1] hold r1/r2, W16[0..2] w/w1/w2
    const int kf = ref32 ? (int)kSnapF32 : g.krylov_fmt;
    const bool k16 = kf != kSnapF32;
    unsigned short *Rq[2] = {nullptr, nullptr}, *Wq[3] = {nullptr, nullptr, nullptr};
    float *Rs[2] = {nullptr, nullptr}, *Ws[3] = {nullptr, nullptr, nullptr};
    const size_t *dcoff = nullptr;
    if (k16) {
        std::vector<size_t> coff(ns);
        size_t nch = 0;
        for (int s = 0; s < ns; ++s) {
            coff[s] = nch;
            nch += ((size_t)len[s] + kBlock - 1) / kBlock;
        }
        if (int rc = upload(w.coff, coff)) return rc;
        dcoff = w.coff.as<size_t>();
        if (int rc = w.q16.alloc(5 * blk * sizeof(unsigned short))) return rc;
        if (int rc = w.s16.alloc(5 * std::max<size_t>(nch, 1) * sizeof(float))) return rc;
        unsigned short *q = w.q16.as<unsigned short>();
        float *sc = w.s16.as<float>();
        for (int k = 0; k < 2; ++k) { Rq[k] = q + k * blk; Rs[k] = sc + k * nch; }
        for (int k = 0; k < 3; ++k) { Wq[k] = q + (2 + k) * blk; Ws[k] = sc + (2 + k) * nch; }
        launch_mk_init(kf, b, Rq[0], Rs[0], Rq[1], Rs[1], doff, dcoff, dlen, ns, g.stream);
        launch_mk_zero(Wq[0], Ws[0], Wq[1], Ws[1], Wq[2], Ws[2], blk, nch, g.stream);
    } else {
        // y = r1 = r2 = b; w = w1 = w2 = 0
        for (int k = 0; k < 3; ++k)
            HIPCHK(hipMemcpyAsync(Y[k], b, blk * sizeof(float), hipMemcpyDeviceToDevice, g.stream));
        for (int k = 0; k < 3; ++k) launch_fill(Wv[k], blk, 0.0f, g.stream);
    }
    launch_fill(x, blk, 0.0f, g.stream);
    int ia = 0, ib = 1;                  // 16-bit roles: r1 = R16[ia], r2 = R16[ib]
    int iy = 0, ir1 = 1, ir2 = 2;        // roles in Y
    int iw = 0, iw1 = 1, iw2 = 2;        // roles in Wv
    std::vector<double> scal((size_t)ns * nsc);
    for (int itn = 1; itn <= itnlim; ++itn) {
        HIPCHK(hipMemcpyAsync(scal.data(), w.scal.p, scal.size() * sizeof(double), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        bool any = false;
        for (int s = 0; s < ns; ++s)
            if (scal[(size_t)s * nsc + kSIstop] == 0.0 && scal[(size_t)s * nsc + kSItn] >= 0.0) any = true;
        if (!any) break;
        // y of the previous iteration lives in the r2 buffer (r2 = y, :193); the
        // free buffer Y[iy] receives the new operator output
        if (k16) {
            launch_mk_v(kf, Rq[ib], Rs[ib], v, doff, dcoff, dlen, dscal, ns, g.stream);
            HessOut h;
            h.sis = Y[0]; h.reverse = 1; h.variant = 0;
            if (int rc = hessian(B, v, h)) return rc;
            launch_mk_l1(kf, Y[0], v, Rq[ia], Rs[ia], doff, dcoff, dlen, dscal, w.part0.as<double>(), ns, g.stream);
            // the new r2 (encode(y)) goes over r1, which this iteration no longer reads
            launch_mk_l2(kf, Y[0], Rq[ib], Rs[ib], Rq[ia], Rs[ia], doff, dcoff, dlen, dscal, w.part1.as<double>(), ns,
                         g.stream);
            std::swap(ia, ib);
            launch_minres_scalars(w.part1.as<double>(), dscal, w.hist.as<double>(), trace, dlen, itnlim, rtol, ns,
                                  g.stream);
            launch_mk_x(kf, x, Wq[iw1], Ws[iw1], Wq[iw], Ws[iw], Wq[iw2], Ws[iw2], v, doff, dcoff, dlen, dscal, trace,
                        itnlim, ns, g.stream);
            const int nw = iw1, nw2 = iw, nw1 = iw2;
            iw = nw; iw2 = nw2; iw1 = nw1;
            continue;
        }
        if (ref32) launch_minres32_v(Y[ir2], v, doff, dlen, dscal, ns, g.stream);
        else launch_minres_v(Y[ir2], v, doff, dlen, dscal, ns, g.stream);
        HessOut h;
        h.sis = Y[iy]; h.reverse = 1; h.variant = 0;     // blurring: time-reversed output (:394-396)
        if (int rc = hessian(B, v, h)) return rc;
        if (ref32) {
            launch_minres32_l1(Y[iy], v, Y[ir1], doff, dlen, dscal, ns, g.stream);
            launch_seqdot32(v, Y[iy], doff, dlen, dnts, dscal, 3, 0, true, ns, g.stream);        // alfa
            launch_minres32_l2(Y[iy], Y[ir2], doff, dlen, dscal, ns, g.stream);
            launch_seqdot32(Y[iy], Y[iy], doff, dlen, dnts, dscal, 0, 0, true, ns, g.stream);    // r2.y
        } else {
            launch_minres_l1(Y[iy], v, Y[ir1], doff, dlen, dscal, w.part0.as<double>(), ns, g.stream);
            launch_minres_l2(Y[iy], Y[ir2], doff, dlen, dscal, w.part1.as<double>(), ns, g.stream);
        }
        // r1 <- r2, r2 <- y, y <- free
        const int old_r1 = ir1;
        ir1 = ir2;
        ir2 = iy;
        iy = old_r1;
        // w1 <- w2, w2 <- w, w <- new (written over old w1)
        if (ref32) {
            launch_minres32_scalars(dscal, w.hist.as<double>(), trace, dlen, itnlim, (float)rtol, ns, g.stream);
            launch_minres32_x(x, Wv[iw1], Wv[iw], Wv[iw2], v, doff, dlen, dscal, trace, itnlim, ns, g.stream);
        } else {
            launch_minres_scalars(w.part1.as<double>(), dscal, w.hist.as<double>(), trace, dlen, itnlim, rtol, ns,
                                  g.stream);
            launch_minres_x(x, Wv[iw1], Wv[iw], Wv[iw2], v, doff, dlen, dscal, trace, itnlim, ns, g.stream);
        }
        const int nw = iw1, nw2 = iw, nw1 = iw2;
        iw = nw; iw2 = nw2; iw1 = nw1;
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    if (hist_host)
        HIPCHK(hipMemcpy(hist_host, w.hist.p, (size_t)ns * std::max(itnlim, 1) * sizeof(double), hipMemcpyDeviceToHost));
    if (scal_host) HIPCHK(hipMemcpy(scal_host, w.scal.p, (size_t)ns * nsc * sizeof(double), hipMemcpyDeviceToHost));
    if (log_host) {
        std::vector<double> tr(ntrace), sc((size_t)ns * nsc);
        HIPCHK(hipMemcpy(tr.data(), w.trace.p, ntrace * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(sc.data(), w.scal.p, sc.size() * sizeof(double), hipMemcpyDeviceToHost));
        const size_t per = (size_t)std::max(itnlim, 1) * kTraceStride;
        for (int s = 0; s < ns; ++s) {
            double *blk = log_host + (size_t)s * log_block(itnlim);
            scal_info(sc.data() + (size_t)s * nsc, blk);
            std::memcpy(blk + kLogHead, tr.data() + (size_t)s * per, per * sizeof(double));
        }
    }
    return 0;
}

// Modes 1 and 2 of sub_redatum_DAW.f90 for the shots of one batch.  Trace
// blocks follow the batch's layout: shot s at roff[s] (acquisition) and
// doff[s] (datum).
struct RedatumIO {
    const float *s_dev = nullptr;
    int type_src = 0;
    const float *dobs = nullptr;
    float mu = 0.0f, rtol = 0.0f;
    int maxiter = 0;
    float *datum = nullptr, *datae = nullptr, *deltade = nullptr;
    double *hist = nullptr;
    double *log = nullptr;    // per-shot MINRES log blocks (host) or null
    int method = 1;           // 1: DAW (mode 1); 3: Symes (mode 3)
    int filter = 0;
    float dt = 0.0f, flow = 0.0f, fhig = 0.0f, llow = 0.0f, lhig = 0.0f;
};

static int redatum_run(Batch &B, const RedatumIO &io)
{
    const int ns = B.nshots;
    const size_t nv = B.roff[ns], nd = B.doff[ns];
    DevBuf d_datr, d_datum_r, d_res, d_x, d_esc, d_datum_e;
    if (int rc = d_datr.alloc(nv * 4)) return rc;
    if (int rc = d_res.alloc(nv * 4)) return rc;
    if (int rc = d_x.alloc(nv * 4)) return rc;
    if (int rc = d_esc.alloc(nv * 4)) return rc;
    if (nd > 0) {
        if (int rc = d_datum_r.alloc(nd * 4)) return rc;
        if (int rc = d_datum_e.alloc(nd * 4)) return rc;
    }
    // 1. background forward in c0: d^r and the datum gather of u^r (sub_redatum_DAW.f90:82-83).
    // The Symes variant has no reduced wavefield: its residual is the observed
    // data, time reversed (:175-179), and its extended wavefield is u^e alone (:219)
    if (io.method == 3) {
        launch_fill(d_datr.as<float>(), nv, 0.0f, g.stream);
        if (nd) launch_fill(d_datum_r.as<float>(), nd, 0.0f, g.stream);
    } else if (int rc = forward(B, io.s_dev, io.type_src, 1, d_datr.as<float>(), nd ? d_datum_r.as<float>() : nullptr,
                                nullptr)) {
        return rc;
    }
    // 2. reduced residual, time reversed (:103)
    if (B.geo.size() == 1) {
        launch_residuals(io.dobs, d_datr.as<float>(), d_res.as<float>(), B.geo[0]->nt,
                         (size_t)ns * B.geo[0]->rec.n, g.stream);
    } else {
        for (int s = 0; s < ns; ++s)
            launch_residuals(io.dobs + B.roff[s], d_datr.as<float>() + B.roff[s], d_res.as<float>() + B.roff[s],
                             B.geo_of(s).nt, (size_t)B.geo_of(s).rec.n, g.stream);
    }
    // 3. MINRES on (H + fac I) de = dr (:115)
    if (int rc = minres_batch(B, d_res.as<float>(), d_x.as<float>(), io.mu, io.maxiter, (double)io.rtol, io.hist, nullptr,
                              io.log))
        return rc;
    // 4. source extension wavefield: dataesc and its datum gather (:131-132)
    HessOut o;
    o.sis = d_esc.as<float>(); o.reverse = 0; o.variant = 1;
    o.datum = nd ? d_datum_e.as<float>() : nullptr;
    if (int rc = hessian(B, d_x.as<float>(), o)) return rc;
    // 5. data-assimilated data and the redatumed gather of u^r + u^e (:143-145, :238)
    if (io.datae) launch_add(d_datr.as<float>(), d_esc.as<float>(), io.datae, nv, g.stream);
    if (io.deltade) HIPCHK(hipMemcpyAsync(io.deltade, d_x.p, nv * 4, hipMemcpyDeviceToDevice, g.stream));
    if (nd > 0) {
        launch_add(d_datum_r.as<float>(), d_datum_e.as<float>(), io.datum, nd, g.stream);
        // mode 2 band-pass of the redatumed gathers (sub_redatum_DAW.f90:240)
        if (io.filter) {
            if (B.geo.size() == 1) {
                if (int rc = filter_device(io.datum, io.datum, ns * B.geo[0]->dat.n, B.geo[0]->nt, io.dt, io.flow,
                                           io.fhig, io.llow, io.lhig)) return rc;
            } else {
                for (int s = 0; s < ns; ++s)
                    if (int rc = filter_device(io.datum + B.doff[s], io.datum + B.doff[s], B.geo_of(s).dat.n,
                                               B.geo_of(s).nt, io.dt, io.flow, io.fhig, io.llow, io.lhig)) return rc;
            }
        }
    }
    return 0;
}

}  // namespace daw

// ====================================================================== C-ABI
using namespace daw;

extern "C" {

const char *daw_last_error(void) { return g_err.c_str(); }

int daw_version(void) { return DAW_ABI_VERSION; }

int daw_kwsinc2d(int ofs)
{
    std::lock_guard<std::mutex> lk(g.mu);
    g.hicks_ofs = ofs;
    return 0;
}

int daw_deallocatekws2d(void)
{
    std::lock_guard<std::mutex> lk(g.mu);
    g.hicks_ofs = -1000;
    g_model.reset();
    return 0;
}

int daw_hicks1d(float *out)
{
    if (!out) return fail(DAW_ERR_ARG, "null output");
    hicks1d(out);
    return 0;
}

int daw_hicks2d_tile(int ia1, int ia2, int ip, int ofs, float *out)
{
    if (!out || ia1 < 1 || ia1 > kNalpha || ia2 < 1 || ia2 > kNalpha || ip < 0 || ip > kIrad)
        return fail(DAW_ERR_ARG, "hicks2d_tile: index out of range");
    hicks2d_tile(ia1, ia2, ip, ofs, out);
    return 0;
}

int daw_subsourcekws(float z, float x, float h, int ofs, int *ix, int *ialpha, int *nfs)
{
    if (!ix || !ialpha || !nfs || !(h > 0.0f)) return fail(DAW_ERR_ARG, "subsourcekws: bad argument");
    subsourcekws(z, x, h, ofs, ix, ialpha, nfs);
    return 0;
}

int daw_subacqui(const float *z, const float *x, int n, float h, int npml, int ofs, int *tab)
{
    if ((n > 0 && (!z || !x || !tab)) || n < 0 || !(h > 0.0f)) return fail(DAW_ERR_ARG, "subacqui: bad argument");
    for (int i = 0; i < n; ++i) {
        int row[5];
        acqui_row(z[i], x[i], h, npml, ofs, row);
        for (int k = 0; k < 5; ++k) tab[i + k * n] = row[k];
    }
    return 0;
}

int daw_subresiduals(const float *dobs, const float *dcal, float *res, int nt, int nr)
{
    if (nt < 0 || nr < 0 || (nt * (size_t)nr > 0 && (!dobs || !dcal || !res))) return fail(DAW_ERR_ARG, "subresiduals: bad argument");
    std::lock_guard<std::mutex> lk(g.mu);
    if (int rc = ensure_stream()) return rc;
    const size_t n = (size_t)nt * nr;
    if (n == 0) return 0;
    DevBuf a, b, c;
    if (int rc = a.alloc(n * 4)) return rc;
    if (int rc = b.alloc(n * 4)) return rc;
    if (int rc = c.alloc(n * 4)) return rc;
    HIPCHK(hipMemcpy(a.p, dobs, n * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.p, dcal, n * 4, hipMemcpyHostToDevice));
    launch_residuals(a.as<float>(), b.as<float>(), c.as<float>(), nt, nr, g.stream);
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(res, c.p, n * 4, hipMemcpyDeviceToHost));
    return 0;
}

static int check_resample(int n1in, int n2, int n1out, float d1in, float d1out)
{
    if (n1in < 1 || n2 < 0 || n1out < 0 || !(d1in > 0.0f) || !(d1out > 0.0f))
        return fail(DAW_ERR_ARG, "subresample: n1in=%d n2=%d n1out=%d d1in=%g d1out=%g", n1in, n2, n1out, (double)d1in,
                    (double)d1out);
    if ((double)(n1out > 0 ? n1out - 1 : 0) * d1out / d1in > 2.0e9)
        return fail(DAW_ERR_ARG, "subresample: output times reach past 2e9 input samples");
    return 0;
}

// subresample on device traces (caller holds the lock)
static int time_resample_dev(const float *xin, float *xout, int n1in, int n2, int n1out, float d1in, float d1out)
{
    if (int rc = check_resample(n1in, n2, n1out, d1in, d1out)) return rc;
    if ((size_t)n2 * n1out == 0) return 0;
    if (!xin || !xout) return fail(DAW_ERR_ARG, "subresample: null trace pointer");
    if (int rc = ensure_stream()) return rc;
    std::vector<float> coef(9 * kRsNalpha);
    resample_table(coef.data());
    std::vector<int> plan;
    resample_plan(n1out, d1in, d1out, plan);
    DevBuf d_coef, d_plan;
    if (int rc = upload(d_coef, coef)) return rc;
    if (int rc = upload(d_plan, plan)) return rc;
    launch_time_resample(xin, xout, n1in, n1out, (size_t)n2, d_plan.as<int>(), d_coef.as<float>(), g.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g.stream));
    return 0;
}

int daw_subresample(const float *xin, float *xout, int n1in, int n2, int n1out, float d1in, float d1out)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (int rc = check_resample(n1in, n2, n1out, d1in, d1out)) return rc;
    const size_t nin = (size_t)n1in * n2, nout = (size_t)n1out * n2;
    if (nout == 0) return 0;
    if (!xin || !xout) return fail(DAW_ERR_ARG, "subresample: null trace pointer");
    if (int rc = ensure_stream()) return rc;
    DevBuf a, b;
    if (int rc = a.alloc(nin * 4)) return rc;
    if (int rc = b.alloc(nout * 4)) return rc;
    HIPCHK(hipMemcpy(a.p, xin, nin * 4, hipMemcpyHostToDevice));
    if (int rc = time_resample_dev(a.as<float>(), b.as<float>(), n1in, n2, n1out, d1in, d1out)) return rc;
    HIPCHK(hipMemcpy(xout, b.p, nout * 4, hipMemcpyDeviceToHost));
    return 0;
}

int daw_subresample_dev(const float *xin, float *xout, int n1in, int n2, int n1out, float d1in, float d1out)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (int rc = check_resample(n1in, n2, n1out, d1in, d1out)) return rc;
    // each output reads up to nine inputs of its trace: the two blocks must not overlap
    const char *a0 = reinterpret_cast<const char *>(xin), *b0 = reinterpret_cast<const char *>(xout);
    const char *a1 = a0 + (size_t)n2 * n1in * sizeof(float), *b1 = b0 + (size_t)n2 * n1out * sizeof(float);
    if (xin && xout && (size_t)n2 * n1out > 0 && a0 < b1 && b0 < a1)
        return fail(DAW_ERR_ARG, "subresample_dev: input and output traces overlap");
    return time_resample_dev(xin, xout, n1in, n2, n1out, d1in, d1out);
}

// ------------------------------------------------------------------ subresample without allocation
// Scratch of daw_subresample_dev_noalloc: the per-output-sample plan (it0,
// ialpha) of up to kNoAllocMaxOut output samples, then the 9 x 10 table.
// Each GPU holds its own copy of this array for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress gives
// for that device (the array's name is a host-side stand-in).
namespace {
constexpr int kNoAllocMaxOut = 65536;
constexpr int kNoAllocCoef = 2 * kNoAllocMaxOut;            // int offset of the table
constexpr int kNoAllocInts = kNoAllocCoef + 9 * kRsNalpha;
constexpr int kMaxDevices = 64;
}  // namespace
static __device__ __align__(256) int scratch[kNoAllocInts];
static int *g_scratch_dev[kMaxDevices];          // per device: device address of scratch
static int g_scratch_host[kNoAllocInts];         // staging of the plan and table (under g.mu)

// The device address of this device's scratch, looked up once per device.
static int scratch_ptr(int *&ptr)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(DAW_ERR_ARG, "subresample_dev_noalloc: device %d", dev);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<int *>(p);
    }
    ptr = g_scratch_dev[dev];
    return 0;
}

int daw_subresample_dev_noalloc(const float *xin, float *xout, int n1in, int n2, int n1out, float d1in, float d1out)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (int rc = check_resample(n1in, n2, n1out, d1in, d1out)) return rc;
    if (n1out > kNoAllocMaxOut)
        return fail(DAW_ERR_ARG, "subresample_dev_noalloc: n1out=%d exceeds the %d output samples its scratch holds",
                    n1out, kNoAllocMaxOut);
    const char *a0 = reinterpret_cast<const char *>(xin), *b0 = reinterpret_cast<const char *>(xout);
    const char *a1 = a0 + (size_t)n2 * n1in * sizeof(float), *b1 = b0 + (size_t)n2 * n1out * sizeof(float);
    if (xin && xout && (size_t)n2 * n1out > 0 && a0 < b1 && b0 < a1)
        return fail(DAW_ERR_ARG, "subresample_dev_noalloc: input and output traces overlap");
    if ((size_t)n2 * n1out == 0) return 0;
    if (!xin || !xout) return fail(DAW_ERR_ARG, "subresample_dev_noalloc: null trace pointer");
    if (int rc = ensure_stream()) return rc;
    int *ptr = nullptr;
    if (int rc = scratch_ptr(ptr)) return rc;
    resample_plan_into(n1out, d1in, d1out, g_scratch_host);
    resample_table(reinterpret_cast<float *>(g_scratch_host + kNoAllocCoef));
    HIPCHK(hipMemcpyAsync(ptr, g_scratch_host, (size_t)2 * n1out * sizeof(int), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(ptr + kNoAllocCoef, g_scratch_host + kNoAllocCoef, 9 * kRsNalpha * sizeof(int),
                          hipMemcpyHostToDevice, g.stream));
    launch_time_resample(xin, xout, n1in, n1out, (size_t)n2, ptr, reinterpret_cast<const float *>(ptr + kNoAllocCoef),
                         g.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g.stream));
    return 0;
}

int daw_submodeling(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt, const float *s,
                    int type_src, int type_rec, const int *is5, const int *ir, int nr, float *sis, float *wavefield)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (int rc = check_hicks()) return rc;
    if (int rc = check_grid(n1, n2, npml, nt, h, dt)) return rc;
    if (!vp || !s || !is5 || (nr > 0 && (!ir || !sis)) || nr < 0) return fail(DAW_ERR_ARG, "submodeling: null argument");
    if (int rc = ensure_stream()) return rc;
    const Grid G = make_grid(n1, n2, npml, h, dt, ofs);
    const DevModel *M;
    if (int rc = get_model(M, G, vp)) return rc;
    Batch B;
    if (int rc = build_batch(B, G, M, nt, 1, is5, 1, ir, nr, nullptr, 0, false)) return rc;
    DevBuf d_s, d_sis, d_wf;
    if (int rc = d_s.alloc((size_t)nt * 4)) return rc;
    HIPCHK(hipMemcpy(d_s.p, s, (size_t)nt * 4, hipMemcpyHostToDevice));
    if (int rc = d_sis.alloc((size_t)std::max(nr, 1) * nt * 4)) return rc;
    launch_fill(d_sis.as<float>(), (size_t)std::max(nr, 1) * nt, 0.0f, g.stream);
    const bool rec = type_rec == 0;
    if (wavefield && rec) {
        if (int rc = d_wf.alloc(G.plane_i() * nt * 4)) return rc;
    }
    if (int rc = forward(B, d_s.as<float>(), type_src, rec ? 1 : 0, d_sis.as<float>(), nullptr, d_wf.as<float>()))
        return rc;
    HIPCHK(hipStreamSynchronize(g.stream));
    if (nr > 0) HIPCHK(hipMemcpy(sis, d_sis.p, (size_t)nr * nt * 4, hipMemcpyDeviceToHost));
    if (wavefield) {
        if (rec) HIPCHK(hipMemcpy(wavefield, d_wf.p, G.plane_i() * nt * 4, hipMemcpyDeviceToHost));
        else std::memset(wavefield, 0, G.plane_i() * nt * 4);
    }
    return 0;
}

static int esource_common(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                          const float *res, int type_rec, int nr, const int *ir, float *sis, float *wavefieldp,
                          int variant)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (int rc = check_hicks()) return rc;
    if (int rc = check_grid(n1, n2, npml, nt, h, dt)) return rc;
    if (!vp || nr < 0 || (nr > 0 && (!ir || !res || !sis))) return fail(DAW_ERR_ARG, "esource: null argument");
    if (int rc = ensure_stream()) return rc;
    const Grid G = make_grid(n1, n2, npml, h, dt, ofs);
    const DevModel *M;
    if (int rc = get_model(M, G, vp)) return rc;
    Batch B;
    if (int rc = build_batch(B, G, M, nt, 1, nullptr, 1, ir, nr, nullptr, 0, true)) return rc;
    DevBuf d_res, d_sis, d_wf;
    const size_t n = (size_t)std::max(nr, 1) * nt;
    if (int rc = d_res.alloc(n * 4)) return rc;
    if (int rc = d_sis.alloc(n * 4)) return rc;
    launch_fill(d_res.as<float>(), n, 0.0f, g.stream);
    launch_fill(d_sis.as<float>(), n, 0.0f, g.stream);
    HIPCHK(hipStreamSynchronize(g.stream));
    if (nr > 0) HIPCHK(hipMemcpy(d_res.p, res, (size_t)nr * nt * 4, hipMemcpyHostToDevice));
    const bool rec = type_rec == 0;
    if (wavefieldp && rec)
        if (int rc = d_wf.alloc(G.plane_i() * nt * 4)) return rc;
    HessOut o;
    o.sis = d_sis.as<float>(); o.reverse = 0; o.variant = variant; o.record = rec ? 1 : 0;
    o.wf = d_wf.as<float>();
    if (int rc = hessian(B, d_res.as<float>(), o)) return rc;
    HIPCHK(hipStreamSynchronize(g.stream));
    if (nr > 0) HIPCHK(hipMemcpy(sis, d_sis.p, (size_t)nr * nt * 4, hipMemcpyDeviceToHost));
    if (wavefieldp) {
        if (rec) HIPCHK(hipMemcpy(wavefieldp, d_wf.p, G.plane_i() * nt * 4, hipMemcpyDeviceToHost));
        else std::memset(wavefieldp, 0, G.plane_i() * nt * 4);
    }
    return 0;
}

int daw_submodeling_adjoint_esource_data(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                                         const float *res, int type_rec, int nr, const int *ir, float *sis)
{
    return esource_common(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis, nullptr, 0);
}

int daw_submodeling_adjoint_esource(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                                    const float *res, int type_rec, int nr, const int *ir, float *sis,
                                    float *wavefieldp)
{
    return esource_common(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, ir, sis, wavefieldp, 1);
}

int daw_subwavefieldresample(int n1, int n2, int npml, int nt, float dt, const int *ir, int nr, float *sis,
                             const float *wavefield)
{
    (void)dt;
    std::lock_guard<std::mutex> lk(g.mu);
    if (int rc = check_hicks()) return rc;
    if (n1 < 1 || n2 < 1 || nt < 1 || nr < 0 || (nr > 0 && (!ir || !sis || !wavefield)))
        return fail(DAW_ERR_ARG, "subwavefieldresample: bad argument");
    if (nr == 0) return 0;
    if (int rc = ensure_stream()) return rc;
    const Grid G = make_grid(n1, n2, npml, 1.0f, 1.0f, 0);
    PointSet P;
    if (int rc = make_points(P, ir, nr, G, true)) return rc;
    std::vector<int> c1(nr), c2(nr);
    for (int r = 0; r < nr; ++r) { c1[r] = P.ix1[r] - npml; c2[r] = P.ix2[r] - npml; }
    DevBuf d_wf, d_c1, d_c2, d_coef, d_sis;
    if (int rc = d_wf.alloc(G.plane_i() * nt * 4)) return rc;
    HIPCHK(hipMemcpy(d_wf.p, wavefield, G.plane_i() * nt * 4, hipMemcpyHostToDevice));
    if (int rc = upload(d_c1, c1)) return rc;
    if (int rc = upload(d_c2, c2)) return rc;
    if (int rc = upload(d_coef, P.coef)) return rc;
    if (int rc = d_sis.alloc((size_t)nr * nt * 4)) return rc;
    launch_resample(d_wf.as<float>(), n1, n2, nt, d_c1.as<int>(), d_c2.as<int>(), d_coef.as<float>(), nr,
                    d_sis.as<float>(), g.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(sis, d_sis.p, (size_t)nr * nt * 4, hipMemcpyDeviceToHost));
    return 0;
}

static int sublinearminres(const float *c0, int n1, int n2, int npml, float h, int ofs, int nt, float dt, int type_rec,
                           int nrec, const int *ir, float mu, float maxiter, float error, const float *deltadr,
                           float *deltade, double *hist, double *info, double *log)
{
    if (int rc = check_hicks()) return rc;
    if (int rc = check_grid(n1, n2, npml, nt, h, dt)) return rc;
    if (!c0 || nrec < 1 || !ir || !deltadr || !deltade) return fail(DAW_ERR_ARG, "sublinearminres: null argument");
    if (type_rec != 0) return fail(DAW_ERR_ARG, "sublinearminres: type_rec must be 0 (the data operator records pressure)");
    if (int rc = ensure_stream()) return rc;
    const int itnlim = (int)maxiter;   // itnlim = grid%maxiter (REAL -> INTEGER), sublinearminres.f90:48
    const Grid G = make_grid(n1, n2, npml, h, dt, ofs);
    const DevModel *M;
    if (int rc = get_model(M, G, c0)) return rc;
    Batch B;
    if (int rc = build_batch(B, G, M, nt, 1, nullptr, 1, ir, nrec, nullptr, 0, true)) return rc;
    const size_t n = (size_t)nrec * nt;
    DevBuf d_b, d_x;
    if (int rc = d_b.alloc(n * 4)) return rc;
    if (int rc = d_x.alloc(n * 4)) return rc;
    HIPCHK(hipMemcpy(d_b.p, deltadr, n * 4, hipMemcpyHostToDevice));
    std::vector<double> scal(minres_scal_stride());
    std::vector<double> hh(std::max(itnlim, 1));
    if (int rc = minres_batch(B, d_b.as<float>(), d_x.as<float>(), mu, itnlim, (double)error, hh.data(), scal.data(),
                              log))
        return rc;
    HIPCHK(hipMemcpy(deltade, d_x.p, n * 4, hipMemcpyDeviceToHost));
    if (hist) std::memcpy(hist, hh.data(), (size_t)std::max(itnlim, 1) * sizeof(double));
    if (info) scal_info(scal.data(), info);
    return 0;
}

int daw_sublinearminres(const float *c0, int n1, int n2, int npml, float h, int ofs, int nt, float dt, int type_rec,
                        int nrec, const int *ir, float mu, float maxiter, float error, const float *deltadr,
                        float *deltade, double *hist, double *info)
{
    std::lock_guard<std::mutex> lk(g.mu);
    return sublinearminres(c0, n1, n2, npml, h, ofs, nt, dt, type_rec, nrec, ir, mu, maxiter, error, deltadr, deltade,
                           hist, info, nullptr);
}

static int write_text(const char *path, const std::string &text)
{
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(DAW_ERR_ARG, "cannot open %s for writing", path);
    const size_t w = std::fwrite(text.data(), 1, text.size(), f);
    if (std::fclose(f) != 0 || w != text.size()) return fail(DAW_ERR_ARG, "cannot write %s", path);
    return 0;
}

int daw_sublinearminres_log(const float *c0, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                            int type_rec, int nrec, const int *ir, float mu, float maxiter, float error,
                            const float *deltadr, float *deltade, const char *log_path)
{
    std::lock_guard<std::mutex> lk(g.mu);
    const int itnlim = (int)maxiter;
    std::vector<double> blk(log_block(itnlim));
    if (int rc = sublinearminres(c0, n1, n2, npml, h, ofs, nt, dt, type_rec, nrec, ir, mu, maxiter, error, deltadr,
                                 deltade, nullptr, nullptr, blk.data()))
        return rc;
    if (!log_path) return 0;
    return write_text(log_path, minres_log_text(nrec * nt, itnlim, error, blk.data()));
}

int daw_write_minres_log(const char *path, int n, int itnlim, float rtol, const double *block)
{
    if (!path || !block || n < 0) return fail(DAW_ERR_ARG, "write_minres_log: bad argument");
    return write_text(path, minres_log_text(n, itnlim, rtol, block));
}

int daw_readpar(const char *path, daw_par *par)
{
    if (!path || !par) return fail(DAW_ERR_ARG, "readpar: null argument");
    std::string err;
    if (read_par(path, par, err) != 0) return fail(DAW_ERR_ARG, "readpar: %s", err.c_str());
    return 0;
}

int daw_read_acqui(const char *path, int *nsrc, int *nrec, float *src_zx, int src_cap, float *rec_zx, int rec_cap)
{
    if (!path || !nsrc || !nrec) return fail(DAW_ERR_ARG, "read_acqui: null argument");
    std::string err;
    std::vector<float> s, r;
    int ns = 0, nr = 0;
    if (read_acqui(path, ns, nr, s, r, err) != 0) return fail(DAW_ERR_ARG, "read_acqui: %s", err.c_str());
    *nsrc = ns;
    *nrec = nr;
    if (src_zx) {
        if (src_cap < ns) return fail(DAW_ERR_ARG, "read_acqui: %d sources, room for %d", ns, src_cap);
        std::memcpy(src_zx, s.data(), s.size() * sizeof(float));
    }
    if (rec_zx) {
        if (rec_cap < nr) return fail(DAW_ERR_ARG, "read_acqui: %d receivers, room for %d", nr, rec_cap);
        std::memcpy(rec_zx, r.data(), r.size() * sizeof(float));
    }
    return 0;
}

int daw_filmartin2d(const float *din, int nr, int nt, float dt, float flow, float fhig, float llow, float lhig,
                    float *dout)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (nr < 0 || nt < 1 || (nr > 0 && (!din || !dout))) return fail(DAW_ERR_ARG, "filmartin2d: bad argument");
    std::vector<double> w;
    if (filter_design(nt, dt, flow, fhig, llow, lhig, w) < 0)
        return fail(DAW_ERR_ARG, "filmartin2d: nt=%d exceeds the 65536-sample FFT length list", nt);
    if (nr == 0) return 0;
    if (int rc = ensure_stream()) return rc;
    const size_t n = (size_t)nr * nt;
    DevBuf d;
    if (int rc = d.alloc(n * sizeof(float))) return rc;
    HIPCHK(hipMemcpy(d.p, din, n * sizeof(float), hipMemcpyHostToDevice));
    if (int rc = filter_device(d.as<float>(), d.as<float>(), nr, nt, dt, flow, fhig, llow, lhig)) return rc;
    HIPCHK(hipMemcpy(dout, d.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

// ------------------------------------------------------------------ batched redatuming
int daw_redatum_batch(const daw_redatum_args *a)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (!a) return fail(DAW_ERR_ARG, "null args");
    if (int rc = check_hicks()) return rc;
    if (int rc = check_grid(a->n1, a->n2, a->npml, a->nt, a->h, a->dt)) return rc;
    if (!a->c0 || (!a->s_dev && a->method != 3) || !a->src_tab || !a->rec_tab || a->nrec < 1 || a->nshots < 1 ||
        !a->dobs_dev)
        return fail(DAW_ERR_ARG, "redatum_batch: missing input");
    if (a->method != 0 && a->method != 1 && a->method != 3)
        return fail(DAW_ERR_ARG, "redatum_batch: method %d (1 = DAW, 3 = Symes)", a->method);
    if (a->ndat > 0 && (!a->dat_tab || !a->datum_dev)) return fail(DAW_ERR_ARG, "redatum_batch: datum receivers need dat_tab and datum_dev");
    if (int rc = ensure_stream()) return rc;
    const Grid G = make_grid(a->n1, a->n2, a->npml, a->h, a->dt, a->ofs);
    const DevModel *M;
    if (int rc = get_model(M, G, a->c0)) return rc;
    const int nt = a->nt, nrec = a->nrec, ndat = a->ndat;
    const size_t nvec = (size_t)nrec * nt, ndv = (size_t)ndat * nt;
    // HBM admission: snapshots dominate (3*n1*n2*nt floats per shot)
    const int fmt = g.hess_exact ? (int)kSnapF32 : g.snap_fmt;
    size_t per_shot = snap_bytes(G, nt, fmt) + 8 * G.plane_e() * 4 + 14 * nvec * 4 + 3 * ndv * 4 + (1u << 20);
    int maxb = a->max_batch;
    if (maxb <= 0) {
        size_t freeb = 0, total = 0;
        HIPCHK(hipMemGetInfo(&freeb, &total));
        const double budget = a->hbm_budget_bytes > 0 ? (double)a->hbm_budget_bytes : 0.85 * (double)freeb;
        maxb = (int)std::max(1.0, std::floor(budget / (double)per_shot));
        // shots per launch: no more than keep a launch's double-buffered wave
        // state within ~192 MB, so that the state written by one step is
        // still in the 256 MB last-level cache when the next step reads it,
        // and a whole multiple of the 8 XCDs, so that every XCD advances whole
        // shots (the tile order gives each XCD a contiguous run of tiles, and
        // neighbouring tiles then share its L2).  Measured at 201x801, us per
        // shot-step of back-propagation / re-injection: 16 shots 2.42 / 2.57,
        // 20: 2.65 / 2.47, 22: 2.66 / 2.43, 24: 2.41 / 2.37, 28: 2.62 / 2.42.
        // At least 8, so a launch still fills the GPU.
        const double state = 8.0 * (double)G.plane_e() * 4.0;
        int cap = std::max(8, (int)(192e6 / state));
        if (cap > 8) cap -= cap % 8;
        maxb = std::min(maxb, cap);
    }
    maxb = std::min(maxb, a->nshots);
    {   // sub-batches of nearly equal count, whole multiples of 8 shots where that fits
        const int nb = (a->nshots + maxb - 1) / maxb;
        int eq = (a->nshots + nb - 1) / nb;
        if (maxb % 8 == 0 && eq % 8 != 0 && (eq + 7) / 8 * 8 <= maxb) eq = (eq + 7) / 8 * 8;
        maxb = eq;
    }
    g.max_batch_used = maxb;
    for (int s0 = 0; s0 < a->nshots; s0 += maxb) {
        const int ns = std::min(maxb, a->nshots - s0);
        Batch B;
        std::vector<int> srows((size_t)ns * 5);
        for (int s = 0; s < ns; ++s)
            for (int k = 0; k < 5; ++k) srows[s + (size_t)k * ns] = a->src_tab[(s0 + s) + (size_t)k * a->nsrc_total];
        if (int rc = build_batch(B, G, M, nt, ns, srows.data(), ns, a->rec_tab, nrec, a->dat_tab, ndat, true)) return rc;
        if (int rc = autotune_tiles(B)) return rc;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, g.stream));
        RedatumIO io;
        io.s_dev = a->s_dev; io.type_src = a->type_src;
        io.dobs = a->dobs_dev + (size_t)s0 * nvec;
        io.mu = a->mu; io.maxiter = a->maxiter; io.rtol = a->rtol;
        io.datum = ndat > 0 ? a->datum_dev + (size_t)s0 * ndv : nullptr;
        io.datae = a->datae_dev ? a->datae_dev + (size_t)s0 * nvec : nullptr;
        io.deltade = a->deltade_dev ? a->deltade_dev + (size_t)s0 * nvec : nullptr;
        io.hist = a->hist_host ? a->hist_host + (size_t)s0 * std::max(a->maxiter, 1) : nullptr;
        io.log = a->minres_log_host ? a->minres_log_host + (size_t)s0 * log_block(a->maxiter) : nullptr;
        io.method = a->method == 3 ? 3 : 1;
        io.filter = a->filter; io.dt = a->dt; io.flow = a->flow; io.fhig = a->fhig; io.llow = a->llow; io.lhig = a->lhig;
        if (int rc = redatum_run(B, io)) return rc;
        HIPCHK(hipEventRecord(e1, g.stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        g.sweep_ms += ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    return 0;
}

int daw_redatum_uneven(const daw_uneven_args *a)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (!a) return fail(DAW_ERR_ARG, "null args");
    if (int rc = check_hicks()) return rc;
    if (!a->c0 || (!a->s_dev && a->method != 3) || !a->x0 || !a->n2s || !a->nts || !a->src_tab || !a->nrecs ||
        !a->rec_tabs || !a->dobs_dev || a->nshots < 1)
        return fail(DAW_ERR_ARG, "redatum_uneven: missing input");
    if (a->method != 0 && a->method != 1 && a->method != 3)
        return fail(DAW_ERR_ARG, "redatum_uneven: method %d (1 = DAW, 3 = Symes)", a->method);
    const int ns_all = a->nshots;
    const bool dat = a->ndats != nullptr && a->dat_tabs != nullptr && a->datum_dev != nullptr;
    // per shot: subdomain, tables and trace offsets in the caller's blocks
    std::vector<size_t> rtab(ns_all + 1, 0), dtab(ns_all + 1, 0), roff(ns_all + 1, 0), doff(ns_all + 1, 0);
    for (int s = 0; s < ns_all; ++s) {
        const int nd = dat ? a->ndats[s] : 0;
        if (a->x0[s] < 0 || a->n2s[s] < 1 || a->x0[s] + a->n2s[s] > a->n2 || a->nrecs[s] < 1 || nd < 0)
            return fail(DAW_ERR_ARG, "redatum_uneven: shot %d: subdomain [%d, %d) of %d columns, %d receivers", s,
                        a->x0[s], a->x0[s] + a->n2s[s], a->n2, a->nrecs[s]);
        if (int rc = check_grid(a->n1, a->n2s[s], a->npml, a->nts[s], a->h, a->dt)) return rc;
        rtab[s + 1] = rtab[s] + (size_t)5 * a->nrecs[s];
        dtab[s + 1] = dtab[s] + (size_t)5 * nd;
        roff[s + 1] = roff[s] + (size_t)a->nrecs[s] * a->nts[s];
        doff[s + 1] = doff[s] + (size_t)nd * a->nts[s];
    }
    if (int rc = ensure_stream()) return rc;
    size_t freeb = 0, total = 0;
    HIPCHK(hipMemGetInfo(&freeb, &total));
    const double budget = a->hbm_budget_bytes > 0 ? a->hbm_budget_bytes : 0.85 * (double)freeb;
    const int fmt = g.hess_exact ? (int)kSnapF32 : g.snap_fmt;
    g.max_batch_used = 0;
    for (int s0 = 0; s0 < ns_all;) {
        // admit shots in order while they fit the HBM budget and, unless
        // max_batch is given, keep a launch's wave state within ~192 MB (at least 8 shots)
        int ns = 0;
        double bytes = 0.0, state = 0.0;
        while (s0 + ns < ns_all) {
            const int s = s0 + ns;
            const Grid Gs = make_grid(a->n1, a->n2s[s], a->npml, a->h, a->dt, a->ofs);
            const double b = (double)snap_bytes(Gs, a->nts[s], fmt) + 8.0 * Gs.plane_e() * 4 +
                             14.0 * (roff[s + 1] - roff[s]) * 4 + 3.0 * (doff[s + 1] - doff[s]) * 4 + (1 << 20);
            const double st = 8.0 * Gs.plane_e() * 4;
            if (ns > 0 && (bytes + b > budget || (a->max_batch > 0 ? ns >= a->max_batch : (ns >= 8 && state + st > 192e6))))
                break;
            bytes += b;
            state += st;
            ++ns;
        }
        {   // balance: the remaining shots in sub-batches of nearly equal count
            const int rest = ns_all - s0, nb = (rest + ns - 1) / ns;
            ns = (rest + nb - 1) / nb;
        }
        g.max_batch_used = std::max(g.max_batch_used, ns);
        Batch B;
        B.nshots = ns;
        B.gi.resize(ns);
        B.inj_src.resize(ns);
        for (int k = 0; k < ns; ++k) {
            const int s = s0 + k;
            auto Q = std::make_unique<Geo>();
            Q->G = make_grid(a->n1, a->n2s[s], a->npml, a->h, a->dt, a->ofs);
            Q->own = std::make_unique<DevModel>();
            // the subdomain's columns x0 .. x0+n2s-1 of c0(n1,n2) are contiguous
            if (int rc = make_model(*Q->own, Q->G, a->c0 + (size_t)a->x0[s] * a->n1)) return rc;
            Q->M = Q->own.get();
            Q->nt = a->nts[s];
            if (int rc = make_geo(*Q, a->rec_tabs + rtab[s], a->nrecs[s], dat ? a->dat_tabs + dtab[s] : nullptr,
                                  dat ? a->ndats[s] : 0)) return rc;
            B.geo.push_back(std::move(Q));
            B.gi[k] = k;
            if (int rc = make_source(B, k, a->src_tab + s, ns_all)) return rc;
        }
        if (int rc = finish_batch(B, true)) return rc;
        if (int rc = autotune_tiles(B)) return rc;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, g.stream));
        RedatumIO io;
        io.s_dev = a->s_dev; io.type_src = a->type_src;
        io.dobs = a->dobs_dev + roff[s0];
        io.mu = a->mu; io.maxiter = a->maxiter; io.rtol = a->rtol;
        io.datum = dat ? a->datum_dev + doff[s0] : nullptr;
        io.datae = a->datae_dev ? a->datae_dev + roff[s0] : nullptr;
        io.deltade = a->deltade_dev ? a->deltade_dev + roff[s0] : nullptr;
        io.hist = a->hist_host ? a->hist_host + (size_t)s0 * std::max(a->maxiter, 1) : nullptr;
        io.log = a->minres_log_host ? a->minres_log_host + (size_t)s0 * log_block(a->maxiter) : nullptr;
        io.method = a->method == 3 ? 3 : 1;
        io.filter = a->filter; io.dt = a->dt; io.flow = a->flow; io.fhig = a->fhig; io.llow = a->llow; io.lhig = a->lhig;
        if (int rc = redatum_run(B, io)) return rc;
        HIPCHK(hipEventRecord(e1, g.stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        g.sweep_ms += ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        s0 += ns;
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    return 0;
}

int daw_modeling_batch(const daw_modeling_args *a)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (!a) return fail(DAW_ERR_ARG, "null args");
    if (int rc = check_hicks()) return rc;
    if (int rc = check_grid(a->n1, a->n2, a->npml, a->nt, a->h, a->dt)) return rc;
    if (!a->vp || !a->s_dev || !a->src_tab || a->nshots < 1 || a->nrec < 0 || (a->nrec > 0 && (!a->rec_tab || !a->dobs_dev)))
        return fail(DAW_ERR_ARG, "modeling_batch: missing input");
    if (a->ndat > 0 && (!a->dat_tab || !a->datum_dev)) return fail(DAW_ERR_ARG, "modeling_batch: datum receivers need dat_tab and datum_dev");
    if (int rc = ensure_stream()) return rc;
    const Grid G = make_grid(a->n1, a->n2, a->npml, a->h, a->dt, a->ofs);
    const DevModel *M;
    if (int rc = get_model(M, G, a->vp)) return rc;
    const int nt = a->nt;
    const size_t nvec = (size_t)a->nrec * nt, ndv = (size_t)a->ndat * nt;
    const int maxb = std::max(1, std::min(a->max_batch > 0 ? a->max_batch : 256, a->nshots));
    for (int s0 = 0; s0 < a->nshots; s0 += maxb) {
        const int ns = std::min(maxb, a->nshots - s0);
        std::vector<int> srows((size_t)ns * 5);
        for (int s = 0; s < ns; ++s)
            for (int k = 0; k < 5; ++k) srows[s + (size_t)k * ns] = a->src_tab[(s0 + s) + (size_t)k * a->nsrc_total];
        Batch B;
        if (int rc = build_batch(B, G, M, nt, ns, srows.data(), ns, a->rec_tab, a->nrec, a->dat_tab, a->ndat, false))
            return rc;
        if (int rc = forward(B, a->s_dev, a->type_src, a->type_rec == 0 ? 1 : 0,
                             a->nrec > 0 ? a->dobs_dev + (size_t)s0 * nvec : nullptr,
                             a->ndat > 0 ? a->datum_dev + (size_t)s0 * ndv : nullptr, nullptr))
            return rc;
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    return 0;
}

int daw_filmartin2d_dev(const float *din, int nr, int nt, float dt, float flow, float fhig, float llow, float lhig,
                        float *dout)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (nr < 0 || nt < 1 || (nr > 0 && (!din || !dout))) return fail(DAW_ERR_ARG, "filmartin2d_dev: bad argument");
    if (int rc = ensure_stream()) return rc;
    return filter_device(din, dout, nr, nt, dt, flow, fhig, llow, lhig);
}

int daw_submodeling_transpose(const float *vp, int n1, int n2, int npml, float h, int ofs, int nt, float dt,
                              const float *d, const int *is5, const int *ir, int nr, float *gsrc)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (int rc = check_hicks()) return rc;
    if (int rc = check_grid(n1, n2, npml, nt, h, dt)) return rc;
    if (!vp || !is5 || !gsrc || nr < 1 || !ir || !d) return fail(DAW_ERR_ARG, "submodeling_transpose: null argument");
    if (int rc = ensure_stream()) return rc;
    const Grid G = make_grid(n1, n2, npml, h, dt, ofs);
    const DevModel *M;
    if (int rc = get_model(M, G, vp)) return rc;
    Batch B;
    if (int rc = build_batch(B, G, M, nt, 1, nullptr, 1, ir, nr, nullptr, 0, false)) return rc;
    // the source's Hicks tile and band carry B^T (the source gradient)
    PointSet src;
    if (int rc = make_points(src, is5, 1, G, false)) return rc;
    Band sband;
    if (int rc = make_band(sband, {&src}, G)) return rc;
    DevBuf d_scoef, d_scoefT, d_d, d_g, d_band, d_items;
    if (int rc = upload_record_coef(d_scoef, d_scoefT, src.coef)) return rc;
    if (int rc = d_d.alloc((size_t)nr * nt * 4)) return rc;
    HIPCHK(hipMemcpy(d_d.p, d, (size_t)nr * nt * 4, hipMemcpyHostToDevice));
    if (int rc = d_g.alloc((size_t)nt * 4)) return rc;
    if (int rc = d_band.alloc((size_t)kBandSlots * std::max(sband.nslot, 1) * G.ldx * 4)) return rc;
    ShotDev &S = B.W.host[0];
    set_inject(S, B.geo[0]->inj_rec_c, d_d.as<float>(), nt, 0);
    S.amp_rev = 0;
    S.band_row = sband.d_row.as<int>();
    S.band_stride = std::max(sband.nslot, 1) * G.ldx;
    S.pband = d_band.as<float>();
    std::vector<RecItem> items;
    add_records(items, src, 0, 0, 0, d_scoef, d_scoefT, d_g.as<float>(), nt);
    if (int rc = upload(d_items, items)) return rc;
    if (int rc = refresh_shots(B.W)) return rc;
    if (int rc = run_adjoint_sweep(false, B.W, d_items.as<RecItem>(), (int)items.size())) return rc;
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(gsrc, d_g.p, (size_t)nt * 4, hipMemcpyDeviceToHost));
    return 0;
}

int daw_set_graph(int on)
{
    std::lock_guard<std::mutex> lk(g.mu);
    g.graph = on != 0;
    return 0;
}

int daw_set_snapshot_format(int fmt)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (fmt != kSnapF32 && fmt != kSnapBF16 && fmt != kSnapF16)
        return fail(DAW_ERR_ARG, "snapshot format %d: 0 (fp32), 1 (bf16) or 2 (fp16)", fmt);
    g.snap_fmt = fmt;
    return 0;
}

int daw_set_hessian(int exact)
{
    std::lock_guard<std::mutex> lk(g.mu);
    g.hess_exact = exact != 0;
    return 0;
}

int daw_set_krylov_format(int fmt)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (fmt < kSnapF32 || fmt > kSnapF16) return fail(DAW_ERR_ARG, "Krylov vector format %d: 0, 1 or 2", fmt);
    g.krylov_fmt = fmt;
    return 0;
}

int daw_set_record_batch(int on)
{
    std::lock_guard<std::mutex> lk(g.mu);
    g.rec_batch = on != 0;
    return 0;
}

int daw_set_minres_arith(int arith)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (arith != 0 && arith != 1) return fail(DAW_ERR_ARG, "minres arithmetic %d: 0 (FP64) or 1 (the reference's FP32)", arith);
    g.minres_arith = arith;
    return 0;
}

int daw_set_tile(int rows, int cols)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (rows == 0) { g.tz_auto = true; return 0; }
    if (cols == kMarch2Cols) {
        if (!march2_rows_ok(rows))
            return fail(DAW_ERR_ARG, "two-step column march of %d rows: 16, 24, 32, 40 or 48", rows);
        if (!g.steps_auto && g.steps == 1) return fail(DAW_ERR_ARG, "the two-step column march advances two steps per launch");
        g.tz_auto = false;
        g.tz = rows;
        g.txw = 0;
        g.march = 2;
        return 0;
    }
    if (cols == kMarchCols) {
        if (!march_rows_ok(rows))
            return fail(DAW_ERR_ARG, "column march of %d rows has no compiled kernel (8, 12, 16, 24, 32)", rows);
        if (!g.steps_auto && g.steps == 2) return fail(DAW_ERR_ARG, "the column march advances one step per launch");
        g.tz_auto = false;
        g.tz = rows;
        g.txw = 0;
        g.march = 1;
        return 0;
    }
    if (cols % 64 != 0 || !tile_shape_ok(rows, cols / 64))
        return fail(DAW_ERR_ARG, "tile %dx%d has no compiled kernel (8x128, 16x64, 16x128, 20x64, 24x64, 32x64, 32x128, 64x64; column march Rx58)", rows, cols);
    if (!g.steps_auto && g.steps == 2 && !tile2_shape_ok(rows, cols / 64))
        return fail(DAW_ERR_ARG, "tile %dx%d has no two-step kernel (16x64, 32x64, 16x128)", rows, cols);
    g.tz_auto = false;
    g.tz = rows;
    g.txw = cols / 64;
    g.march = 0;
    return 0;
}

int daw_set_steps(int steps)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (steps == 0) { g.steps_auto = true; return 0; }
    if (steps != 1 && steps != 2) return fail(DAW_ERR_ARG, "steps per launch %d: 0 (autotuned), 1 or 2", steps);
    if (steps == 2 && !g.tz_auto && (g.march == 1 || (!g.march && !tile2_shape_ok(g.tz, g.txw))))
        return fail(DAW_ERR_ARG, "tile %dx%d has no two-step kernel", g.tz, g.march ? kMarchCols : 64 * g.txw);
    if (steps == 1 && !g.tz_auto && g.march == 2)
        return fail(DAW_ERR_ARG, "the two-step column march advances two steps per launch");
    g.steps_auto = false;
    g.steps = steps;
    return 0;
}

int daw_get_launch(int mode, int *steps, int *rows, int *cols)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (mode < 0 || mode > 2) return fail(DAW_ERR_ARG, "sweep mode %d: 0 forward, 1 back-propagation, 2 re-injection", mode);
    const LaunchCfg &c = g.last_cfg[mode];
    if (steps) *steps = c.steps;
    if (rows) *rows = c.tz;
    if (cols) *cols = tile_cols(c);
    return 0;
}

int daw_get_steps(void)
{
    std::lock_guard<std::mutex> lk(g.mu);
    return g.steps;
}

int daw_get_tile(int *rows, int *cols)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (rows) *rows = g.tz;
    if (cols) *cols = g.march == 2 ? kMarch2Cols : g.march ? kMarchCols : 64 * g.txw;
    return 0;
}

int daw_set_timing(int on)
{
    std::lock_guard<std::mutex> lk(g.mu);
    g.timing = on != 0;
    return 0;
}

int daw_reset_stats(void)
{
    std::lock_guard<std::mutex> lk(g.mu);
    for (int i = 0; i < 3; ++i) { g.t_ms[i] = 0; g.t_cnt[i] = 0; g.t_rec_ms[i] = 0; }
    g.sweep_ms = 0;
    g.pts = 0;
    g.launches = 0;
    return 0;
}

int daw_get_stats(daw_stats *st)
{
    if (!st) return fail(DAW_ERR_ARG, "null stats");
    std::lock_guard<std::mutex> lk(g.mu);
    for (int i = 0; i < 3; ++i) {
        st->sweep_ms[i] = g.t_ms[i];
        st->sweep_steps[i] = g.t_cnt[i];
        st->record_ms[i] = g.t_rec_ms[i];
    }
    st->batch_ms = g.sweep_ms;
    st->point_updates = g.pts;
    st->step_launches = g.launches;
    st->max_batch_used = g.max_batch_used;
    return 0;
}

}  // extern "C"
