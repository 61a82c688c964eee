// Host-side geometry and setup (see daw_host.h).  Compiled with
// -ffp-contract=off: every single-precision expression below is evaluated as
// written, one rounding per operation, which is what makes the tables
// bit-identical to the reference's.
#include "daw_host.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace daw {

float bessi0(float x)
{
    // A&S 9.8.1 / 9.8.2 polynomial; argument algebra in REAL*4, polynomial
    // in REAL*8, result rounded to REAL*4, as in kwsinc.f90:624-631.
    if (std::fabs(x) < 3.75f) {
        const float q = x / 3.75f;
        const double y = static_cast<double>(q * q);
        const double p = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 +
                         y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
        return static_cast<float>(p);
    }
    const float ax = std::fabs(x);
    const double y = static_cast<double>(3.75f / ax);
    const float pre = std::exp(ax) / std::sqrt(ax);
    const double q = 0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 +
                     y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 +
                     y * (-0.1647633e-1 + y * 0.392377e-2)))))));
    return static_cast<float>(static_cast<double>(pre) * q);
}

void hicks1d_table(int irad, int nalpha, float beta, float *out)
{
    const float pi = 3.14159265359f;          // kwsinc.f90:108 (REAL literal)
    const float i0_beta = bessi0(beta);
    const int w = 2 * irad + 1;
    for (int ia = 1; ia <= nalpha; ++ia) {
        const float alpha = 0.5f - static_cast<float>(ia - 1) / static_cast<float>(nalpha);
        for (int n = -irad; n <= irad; ++n) {
            const float xn = static_cast<float>(n) + alpha;
            float c = 0.0f;
            if (std::fabs(xn) <= static_cast<float>(irad)) {
                if (std::fabs(xn) > 1e-15f) {
                    const float arg = beta * std::sqrt(1.0f - xn * xn / static_cast<float>(irad * irad));
                    const float window = bessi0(arg) / i0_beta;
                    c = window * std::sin(pi * xn) / (pi * xn);
                    if (std::fabs(c) < 1e-15f) c = 0.0f;
                } else {
                    c = 1.0f;
                }
            }
            out[(n + irad) + w * (ia - 1)] = c;
        }
    }
}

// bhicks0(irad = 10), kwsinc.f90:43
void hicks1d(float *out) { hicks1d_table(kIrad, kNalpha, 14.18f, out); }

// bhicks0(irad = 4), subresample_hicks.f90:103
void resample_table(float *out) { hicks1d_table(kRsRad, kRsNalpha, 6.31f, out); }

void resample_plan(int n1out, float d1in, float d1out, std::vector<int> &plan)
{
    plan.resize(2 * static_cast<size_t>(n1out));
    resample_plan_into(n1out, d1in, d1out, plan.data());
}

void resample_plan_into(int n1out, float d1in, float d1out, int *plan)
{
    for (int i1 = 1; i1 <= n1out; ++i1) {
        const float tout = static_cast<float>(i1 - 1) * d1out;
        const int it0 = static_cast<int>(std::lround(tout / d1in)) + 1;     // NINT of a REAL
        const float t0 = static_cast<float>(it0 - 1) * d1in;
        const float alpha = (t0 - tout) / d1in;
        int ia = static_cast<int>(std::lround((0.5f - alpha) * static_cast<float>(kRsNalpha))) + 1;
        if (ia > kRsNalpha) ia = kRsNalpha;
        if (ia < 1) ia = 1;       // alpha > 0.55 is out of reach of rounding; the reference has no guard
        plan[2 * (size_t)(i1 - 1)] = it0;
        plan[2 * (size_t)(i1 - 1) + 1] = ia;
    }
}

namespace {
struct Table1D {
    float c[kWin * kNalpha];
    Table1D() { hicks1d(c); }
};
const float *table1d()
{
    static const Table1D t;   // thread-safe one-time init
    return t.c;
}
}  // namespace

void hicks2d_tile(int ia1, int ia2, int ip, int ofs, float *out)
{
    const float *c1 = table1d();
    float base[kWin * kWin];
    // tensor product c(n1,n2) = c1(n2,a2) * c1(n1,a1), kwsinc.f90:268-269
    for (int n2 = 0; n2 < kWin; ++n2)
        for (int n1 = 0; n1 < kWin; ++n1)
            base[n1 + kWin * n2] = c1[n2 + kWin * (ia2 - 1)] * c1[n1 + kWin * (ia1 - 1)];
    if (ip == 0) {
        std::memcpy(out, base, sizeof base);
        return;
    }
    if (std::abs(ofs) != 1) {     // slices ip >= 1 exist only with a free surface
        std::memset(out, 0, sizeof base);
        return;
    }
    // Free surface `ip` nodes above the centre row: the surface node is set to
    // zero and the ip nodes below it become odd-mirrored differences
    // (kwsinc.f90:336-355).  Rows further above keep the plain product.
    std::memcpy(out, base, sizeof base);
    const int surf = ip;          // (-irad + ip) + irad, 0-based row of the surface node
    for (int n2 = 0; n2 < kWin; ++n2) {
        out[surf + kWin * n2] = 0.0f;
        for (int k = 1; k <= ip; ++k)
            out[surf + k + kWin * n2] = base[surf + k + kWin * n2] - base[surf - k + kWin * n2];
    }
}

static int nint(float v) { return static_cast<int>(std::lround(v)); }   // Fortran NINT

void subsourcekws(float z, float x, float h, int ofs, int ix[2], int ialpha[2], int *nfs)
{
    const float xs[2] = {z, x};
    for (int d = 0; d < 2; ++d) {
        ix[d] = nint(xs[d] / h) + 1;
        const float x0 = static_cast<float>(ix[d] - 1) * h;
        const float alpha = (x0 - xs[d]) / h;
        const int ia = nint((0.5f - alpha) * static_cast<float>(kNalpha)) + 1;
        ialpha[d] = ia < kNalpha ? ia : kNalpha;
    }
    int nf = kIrad - nint(xs[0] / h);
    if (nf < 0) nf = 0;
    if (ofs == 0) nf = 0;
    *nfs = nf;
}

void acqui_row(float z, float x, float h, int npml, int ofs, int row[5])
{
    int ix[2], ia[2], nf;
    subsourcekws(z, x, h, ofs, ix, ia, &nf);
    row[0] = ix[0] + npml;
    row[1] = ix[1] + npml;
    row[2] = ia[0];
    row[3] = ia[1];
    row[4] = ofs == 0 ? 0 : nf;
}

// xxfac = 0.12 (submodeling2.f90:35); profiles exp(-(fac*x)^2), exp(-0.5*(fac*x)^2)
static void sponge_pair(float x, float &a, float &a1)
{
    const float fac = 0.12f;
    const float t = fac * x;
    a = std::exp(-(t * t));
    a1 = std::exp(-(0.5f * (t * t)));
}

void sponge_a(int npml, int n, std::vector<float> &damp, std::vector<float> &damp1)
{
    const int ne = n + 2 * npml;
    damp.assign(ne, 1.0f);
    damp1.assign(ne, 1.0f);
    // left ramp at integer nodes x = npml-1 .. 0, mirrored on the right
    for (int i = 1; i <= npml; ++i) {
        float a, a1;
        sponge_pair(static_cast<float>(npml - i), a, a1);
        damp[i - 1] = a;
        damp1[i - 1] = a1;
        damp[npml + n + (npml - i + 1) - 1] = a;
        damp1[npml + n + (npml - i + 1) - 1] = a1;
    }
}

void sponge_b(int npml, int n, std::vector<float> &damp, std::vector<float> &damp1)
{
    const int ne = n + 2 * npml;
    damp.assign(ne, 1.0f);
    damp1.assign(ne, 1.0f);
    // half-node ramps: left at x = npml-1.5 .. 0.5 (npml-1 nodes), right at 0.5 .. npml-0.5
    for (int i = 1; i <= npml - 1; ++i) {
        float a, a1;
        sponge_pair(static_cast<float>(npml - 1 - i) + 0.5f, a, a1);
        damp[i - 1] = a;
        damp1[i - 1] = a1;
    }
    for (int i = 1; i <= npml; ++i) {
        float a, a1;
        sponge_pair(static_cast<float>(i - 1) + 0.5f, a, a1);
        damp[npml + n + i - 1] = a;
        damp1[npml + n + i - 1] = a1;
    }
}

static void hanning_taper(float *han, int nf)
{
    // sphanning(han, nf, nord = 2), minresModule.f90:619-632; han 0-based, nf+1 values
    const float pi = 3.14159265f;
    for (int j = 1; j <= nf + 1; ++j) {
        const float a = 2.0f * pi * static_cast<float>(j - 1) / static_cast<float>(2 * nf);
        const float c = (std::cos(a) + 1.0f) / 2.0f + 0.000001f;
        han[j - 1] = std::exp(std::log(c) * 2.0f);
    }
}

int filter_design(int nt, float dt, float flow, float fhig, float llow, float lhig, std::vector<double> &w)
{
    static const int lengths[14] = {8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536};
    int nto = -1;
    for (int L : lengths)
        if (nt <= L) { nto = L; break; }
    if (nto < 0) return -1;
    const float fe = 1.0f / dt;
    const float fmax = fe / 2.0f;
    if (fhig > fmax) { fhig = fmax; lhig = 0.0f; }
    const int conv = static_cast<int>(static_cast<float>(nto) / fe);   // integer, truncated
    const int iflow = static_cast<int>(flow * static_cast<float>(conv)) + 1;
    const int ifhig = static_cast<int>(fhig * static_cast<float>(conv)) + 1;
    int illow = static_cast<int>(llow * static_cast<float>(conv)) + 1;
    int ilhig = static_cast<int>(lhig * static_cast<float>(conv)) + 1;
    if (iflow - illow < 1 && iflow > 0) illow = iflow - 1;
    if (ifhig + ilhig > nto) ilhig = nto - ifhig;
    std::vector<float> f(nto + 2, 0.0f);   // 1-based fmartin(1..nto+1)
    std::vector<float> han((illow > ilhig ? illow : ilhig) + 2, 0.0f);
    if (illow > 0) {
        hanning_taper(han.data(), illow);
        for (int j = 1; j <= illow; ++j) f[iflow - j + 1] = han[j - 1];
    }
    if (ilhig > 0) {
        hanning_taper(han.data(), ilhig);
        for (int j = 1; j <= ilhig; ++j) f[ifhig + j - 1] = han[j - 1];
    }
    for (int j = iflow; j <= ifhig; ++j) f[j] = 1.0f;
    w.assign(nto + 1, 0.0);
    for (int j = 1; j <= nto + 1; ++j) w[j - 1] = static_cast<double>(f[j] + 0.0001f);
    return nto;
}

}  // namespace daw
