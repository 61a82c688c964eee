// Host-side geometry and setup for the DAW redatuming library.
//
// Everything here runs once per model / acquisition on the host in IEEE
// single precision with no FP contraction, so its results are bit-identical
// to the reference's Fortran (which is compiled for baseline x86-64, no FMA).
// The device only ever sees the finished tables.
#pragma once
#include <cstdint>
#include <vector>

namespace daw {

constexpr int kIrad = 10;          // Hicks half-width, kwsinc.f90:25
constexpr int kWin = 2 * kIrad + 1; // 21-point support
constexpr int kNalpha = 100;       // kwsinc.f90:26

// Modified Bessel I0 in the reference's mixed precision (kwsinc.f90:612-635).
float bessi0(float x);

// 1-D Kaiser-windowed-sinc monopole table (kwsinc.f90:94-179);
// out[(n+10) + 21*(ia-1)], n = -10..10, ia = 1..100.
void hicks1d(float *out);

// Kaiser-windowed-sinc monopole table of half-width irad tabulated at nalpha
// sub-sample shifts with Kaiser beta (the loop of kwsinc.f90:115-177 and of
// subhicks1d, subresample_hicks.f90:110-155); out[(n+irad) + (2*irad+1)*(ia-1)].
// Entries outside the window's support are 0.
void hicks1d_table(int irad, int nalpha, float beta, float *out);

// The 9x10 table of the time resampler: irad = 4, nalpha = 10, beta =
// bhicks0(4) = 6.31 (subresample_hicks.f90:40-41, 103-106).
constexpr int kRsRad = 4, kRsNalpha = 10;
void resample_table(float *out);

// Per output sample i1 (0-based) of subresample: it0 (1-based nearest input
// sample) and ialpha, computed as subresample_hicks.f90:59-66 computes them.
void resample_plan(int n1out, float d1in, float d1out, std::vector<int> &plan);
// the same into plan[2*n1out] (no allocation)
void resample_plan_into(int n1out, float d1in, float d1out, int *plan);

// One 21x21 tile of the 2-D monopole table for (ialpha1, ialpha2, ip) as
// kwsinc2d(ofs) leaves it (kwsinc.f90:212-378); out[(n1+10) + 21*(n2+10)].
void hicks2d_tile(int ia1, int ia2, int ip, int ofs, float *out);

// Grid index / sub-cell shift / nodes-above-free-surface for a point
// (kwsinc.f90:582-606).
void subsourcekws(float z, float x, float h, int ofs, int ix[2], int ialpha[2], int *nfs);

// One row of the src_tab/rec_tab contract of subacqui.f90:33-58.
void acqui_row(float z, float x, float h, int npml, int ofs, int row[5]);

// Split-sponge profiles of cspongea/cspongeb (submodeling2.f90:1331-1393),
// 0-based arrays of length n + 2*npml.
void sponge_a(int npml, int n, std::vector<float> &damp, std::vector<float> &damp1);
void sponge_b(int npml, int n, std::vector<float> &damp, std::vector<float> &damp1);

// Band-pass weights of filmartin2d (minresModule.f90:486-586): returns the
// FFT length nto (power of two >= nt from the reference's list, -1 if nt is
// too large) and fills w[0..nto] with dmartin(1..nto+1).
int filter_design(int nt, float dt, float flow, float fhig, float llow, float lhig, std::vector<double> &w);

}  // namespace daw
