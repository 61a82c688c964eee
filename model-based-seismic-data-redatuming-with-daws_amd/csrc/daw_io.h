// On-disk formats of the reference driver (host side): the parameter file,
// the acquisition file, per-shot file names and the MINRES log.
#pragma once
#include <string>
#include <vector>

#include "daw.h"

namespace daw {

constexpr int kNameLen = 160;        // CHARACTER(LEN=160) file names, main_redatum_DAW.f90:55
constexpr int kLogHeadDoubles = 9;   // fac, Anorm, Acond, rnorm, Arnorm, ynorm, beta1, istop, itn
constexpr int kTraceDoubles = 5;     // per iteration: x(1), qrnorm/beta1, Anorm, Acond, printed

// subreadpar (subreadpar.f90:9-21) on the file at path; 0 or -1 with err set.
int read_par(const char *path, daw_par *p, std::string &err);

// Acquisition file: "nsrc nrec", then z x per source, then z x per receiver
// (main_redatum_DAW.f90:120-123, subacqui.f90:29-57).  zx = [z0 x0 z1 x1 ...].
int read_acqui(const char *path, int &nsrc, int &nrec, std::vector<float> &src_zx, std::vector<float> &rec_zx,
               std::string &err);

// subname (subname.f90:2-20): five digits, zero padded; empty if out of range.
std::string shot_name(int i);

// Text of MINRES_<shot>.txt for one solve: the heading, the printed iteration
// lines and the final status, as minresModule.f90 writes them.  blk holds
// kLogHeadDoubles values and then kTraceDoubles per iteration (itnlim of them).
std::string minres_log_text(int n, int itnlim, float rtol, const double *blk);

}  // namespace daw
