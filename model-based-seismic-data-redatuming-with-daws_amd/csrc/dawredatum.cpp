// dawredatum -- the reference's driver (main_redatum_DAW.f90 +
// sub_redatum_DAW.f90) over libdaw, for hosts without the Fortran program.
//
// In the working directory it reads redatum.par (subreadpar.f90:9-21), the
// wavelet, the two models and the acquisition file named there
// (main_redatum_DAW.f90:98-123), runs the mode the par file selects for every
// shot and writes the files that mode writes, with the same names and layouts
// (sub_redatum_DAW.f90:53-245):
//   mode 0  dataobs_<shot>.bin (nt,nrec); oio=1: wavefield_obs<shot>.bin (n1,n2,nt)
//   mode 1  reads dataobs_org_<shot>.bin; oio=1: datar_, wavefield_ini, deltadr_,
//           deltade_ (both in forward time), MINRES_<shot>.txt, dataesc_, datae_,
//           wavefield_exd
//   mode 2  reads wavefield_exd<shot>.bin; writes dataobs_redatum_<shot>.bin
//   mode 3  (Symes) reads dataobs_org_; oio=1: deltadr_, deltade_, MINRES_,
//           dataesc_, wavefield_exd (= the source-extension wavefield alone)
// All files are unformatted stream REAL*4, first index fastest.
//
// Options beyond the reference:
//   --datum FILE  modes 1 and 3: also do mode 2 in the same run, for the
//                 receivers of acquisition file FILE (its sources are not
//                 used), without writing or reading any wavefield: all shots
//                 run through the batched entry point and each shot's
//                 band-passed datum gather goes to dataobs_redatum_<shot>.bin.
//                 With oio=1, deltade_, datae_ (mode 1) / dataesc_ (mode 3) and
//                 MINRES_ files are written too.
//   --batch N     shots per batched call (default 64); mode 0 with oio=0 and
//                 --datum runs are batched.
//   --dto         wherever dataobs_redatum_<shot>.bin is written (mode 2,
//                 --datum), also write dataobs_redatum_dto_<shot>.bin: the
//                 gathers resampled from dt to dto (nto samples per trace,
//                 sub_redatum_DAW.f90:24) with subresample
//                 (subresample_hicks.f90:25-83), which the reference ships
//                 but does not call.
//   --par FILE    parameter file (default redatum.par).
//   --dry-run     read and check every input (the par file, wavelet, models,
//                 acquisition files, and every shot's dataobs_org_ or
//                 wavefield_exd file the mode will read), print the geometry
//                 tables, stop before any GPU work.  Every run makes the same
//                 checks before its first shot.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "daw.h"

namespace {

[[noreturn]] void die(const std::string &msg)
{
    std::fprintf(stderr, "dawredatum: %s\n", msg.c_str());
    std::exit(1);
}

void check(int rc, const char *what)
{
    if (rc != 0) die(std::string(what) + ": " + daw_last_error());
}

void hip(hipError_t e, const char *what)
{
    if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

// subname.f90:2-20
std::string shot_name(int i)
{
    if (i < 0 || i >= 100000) die("i too big (subroutine subname)");
    char b[16];
    std::snprintf(b, sizeof b, "%05d", i);
    return b;
}

// First record of a direct-access file of record length n*4 bytes.
std::vector<float> read_record(const std::string &path, size_t n)
{
    std::vector<float> v(n);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) die("cannot open " + path);
    const size_t got = std::fread(v.data(), sizeof(float), n, f);
    std::fclose(f);
    if (got != n) die(path + ": expected " + std::to_string(n) + " REAL*4 values, found " + std::to_string(got));
    return v;
}

void write_stream(const std::string &path, const float *p, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) die("cannot open " + path + " for writing");
    const size_t w = std::fwrite(p, sizeof(float), n, f);
    if (std::fclose(f) != 0 || w != n) die("cannot write " + path);
}

// deltadr/deltade are kept time-reversed; their files hold them in forward
// time, trace after trace (sub_redatum_DAW.f90:106-112).
void write_unreversed(const std::string &path, const float *rev, int nt, int nr)
{
    std::vector<float> fwd((size_t)nt * nr);
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nt; ++j) fwd[(size_t)i * nt + j] = rev[(size_t)i * nt + (nt - 1 - j)];
    write_stream(path, fwd.data(), fwd.size());
}

struct Acq {
    int nsrc = 0, nrec = 0;
    std::vector<int> src_tab, rec_tab;   // (n,5) Fortran order
};

Acq read_acquisition(const char *path, const daw_par &p)
{
    Acq a;
    check(daw_read_acqui(path, &a.nsrc, &a.nrec, nullptr, 0, nullptr, 0), "acquisition");
    std::vector<float> s((size_t)2 * a.nsrc), r((size_t)2 * a.nrec);
    check(daw_read_acqui(path, &a.nsrc, &a.nrec, s.data(), a.nsrc, r.data(), a.nrec), "acquisition");
    auto table = [&](const std::vector<float> &zx, int n, std::vector<int> &tab) {
        std::vector<float> z(n), x(n);
        for (int i = 0; i < n; ++i) { z[i] = zx[(size_t)2 * i]; x[i] = zx[(size_t)2 * i + 1]; }
        tab.assign((size_t)5 * n, 0);
        if (n > 0) check(daw_subacqui(z.data(), x.data(), n, p.h, p.npml, p.ofs, tab.data()), "subacqui");
    };
    table(s, a.nsrc, a.src_tab);
    table(r, a.nrec, a.rec_tab);
    return a;
}

std::vector<int> row(const std::vector<int> &tab, int n, int i)
{
    std::vector<int> r(5);
    for (int k = 0; k < 5; ++k) r[k] = tab[(size_t)i + (size_t)k * n];
    return r;
}

// src_tab rows s0 .. s0+ns-1 as an (ns,5) table
std::vector<int> rows(const std::vector<int> &tab, int n, int s0, int ns)
{
    std::vector<int> r((size_t)5 * ns);
    for (int s = 0; s < ns; ++s)
        for (int k = 0; k < 5; ++k) r[(size_t)s + (size_t)k * ns] = tab[(size_t)(s0 + s) + (size_t)k * n];
    return r;
}

template <class T> struct Dev {
    T *p = nullptr;
    explicit Dev(size_t n) { hip(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc"); }
    ~Dev() { if (p) (void)hipFree(p); }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
};

struct Run {
    daw_par p{};
    std::vector<float> s, ctrue, c0;
    Acq acq;
    int batch = 64;
    bool quiet = false;
    bool dto = false;
    // sub_redatum_DAW.f90:24: nto = nint(float(nt-1)*dt/dto)+1
    int nto() const { return (int)std::lround((double)((float)(p.nt - 1) * p.dt / p.dto)) + 1; }
    size_t plane() const { return (size_t)p.n1 * p.n2; }
    size_t ntr() const { return (size_t)p.nt * acq.nrec; }
    std::string bin(const char *stem, int isrc) const { return stem + shot_name(isrc) + ".bin"; }
};

// ---------------------------------------------------------------- mode 0
void mode0(Run &R)
{
    const daw_par &p = R.p;
    const int nr = R.acq.nrec;
    if (p.oio == 1) {               // wavefield files: one shot at a time
        std::vector<float> sis(R.ntr()), wf(R.plane() * p.nt);
        for (int is = 1; is <= R.acq.nsrc; ++is) {
            const std::vector<int> src = row(R.acq.src_tab, R.acq.nsrc, is - 1);
            check(daw_submodeling(R.ctrue.data(), p.n1, p.n2, p.npml, p.h, p.ofs, p.nt, p.dt, R.s.data(), p.type_src,
                                  p.type_rec, src.data(), R.acq.rec_tab.data(), nr, sis.data(), wf.data()),
                  "submodeling");
            write_stream(R.bin("dataobs_", is), sis.data(), sis.size());
            write_stream(R.bin("wavefield_obs", is), wf.data(), wf.size());
        }
        return;
    }
    Dev<float> s_dev(p.nt);
    hip(hipMemcpy(s_dev.p, R.s.data(), p.nt * sizeof(float), hipMemcpyHostToDevice), "copy wavelet");
    const int B = std::min(R.batch, R.acq.nsrc);
    Dev<float> d_obs((size_t)B * R.ntr());
    std::vector<float> h_obs((size_t)B * R.ntr());
    for (int s0 = 0; s0 < R.acq.nsrc; s0 += B) {
        const int ns = std::min(B, R.acq.nsrc - s0);
        const std::vector<int> src = rows(R.acq.src_tab, R.acq.nsrc, s0, ns);
        daw_modeling_args a{};
        a.vp = R.ctrue.data();
        a.n1 = p.n1; a.n2 = p.n2; a.npml = p.npml; a.h = p.h; a.dt = p.dt;
        a.ofs = p.ofs; a.nt = p.nt; a.type_src = p.type_src; a.type_rec = p.type_rec;
        a.s_dev = s_dev.p;
        a.nshots = ns; a.nsrc_total = ns; a.src_tab = src.data();
        a.nrec = nr; a.rec_tab = R.acq.rec_tab.data();
        a.dobs_dev = d_obs.p;
        check(daw_modeling_batch(&a), "modeling_batch");
        hip(hipMemcpy(h_obs.data(), d_obs.p, (size_t)ns * R.ntr() * sizeof(float), hipMemcpyDeviceToHost), "copy");
        for (int s = 0; s < ns; ++s)
            write_stream(R.bin("dataobs_", s0 + s + 1), h_obs.data() + (size_t)s * R.ntr(), R.ntr());
    }
}

// ---------------------------------------------------------------- modes 1 and 3, one shot at a time
void mode13_plain(Run &R)
{
    const daw_par &p = R.p;
    const int nr = R.acq.nrec, nt = p.nt;
    const bool out = p.oio == 1, symes = p.mode == 3;
    const size_t nwf = R.plane() * nt;
    std::vector<float> datar(R.ntr(), 0.0f), dobs, deltadr(R.ntr()), deltade(R.ntr()), dataesc(R.ntr());
    std::vector<float> wfr(out && !symes ? nwf : 0), wfe(out ? nwf : 0);
    for (int is = 1; is <= R.acq.nsrc; ++is) {
        const std::string nm = shot_name(is);
        if (!symes) {
            // u^r and d^r in the background model (:82-83)
            const std::vector<int> src = row(R.acq.src_tab, R.acq.nsrc, is - 1);
            check(daw_submodeling(R.c0.data(), p.n1, p.n2, p.npml, p.h, p.ofs, nt, p.dt, R.s.data(), p.type_src,
                                  p.type_rec, src.data(), R.acq.rec_tab.data(), nr, datar.data(),
                                  out ? wfr.data() : nullptr),
                  "submodeling");
            if (out) {
                write_stream(R.bin("datar_", is), datar.data(), datar.size());
                write_stream(R.bin("wavefield_ini", is), wfr.data(), wfr.size());
            }
        }
        dobs = read_record(R.bin("dataobs_org_", is), R.ntr());
        if (symes) {                 // delta d^r = d^obs time-reversed (:175-179)
            for (int i = 0; i < nr; ++i)
                for (int j = 0; j < nt; ++j) deltadr[(size_t)i * nt + (nt - 1 - j)] = dobs[(size_t)i * nt + j];
        } else {                     // subresiduals (:103)
            check(daw_subresiduals(dobs.data(), datar.data(), deltadr.data(), nt, nr), "subresiduals");
        }
        if (out) write_unreversed(R.bin("deltadr_", is), deltadr.data(), nt, nr);
        // MINRES (:115 / :194), with its log when oio = 1 (sublinearminres.f90:51-61)
        const std::string log = "MINRES_" + nm + ".txt";
        check(daw_sublinearminres_log(R.c0.data(), p.n1, p.n2, p.npml, p.h, p.ofs, nt, p.dt, p.type_rec, nr,
                                      R.acq.rec_tab.data(), p.mu, p.maxiter, p.error, deltadr.data(), deltade.data(),
                                      out ? log.c_str() : nullptr),
              "sublinearminres");
        if (out) write_unreversed(R.bin("deltade_", is), deltade.data(), nt, nr);
        // source extension (:131-132 / :207-208); its band-passed copy dataescf
        // (:134 / :210) feeds no output and is not formed
        check(daw_submodeling_adjoint_esource(R.c0.data(), p.n1, p.n2, p.npml, p.h, p.ofs, nt, p.dt, deltade.data(),
                                              p.type_rec, nr, R.acq.rec_tab.data(), dataesc.data(),
                                              out ? wfe.data() : nullptr),
              "submodeling_adjoint_esource");
        if (!out) continue;
        write_stream(R.bin("dataesc_", is), dataesc.data(), dataesc.size());
        if (symes) {                 // wavefieldr = wavefieldesc (:219)
            write_stream(R.bin("wavefield_exd", is), wfe.data(), wfe.size());
        } else {                     // u^r + u^e and d^r + d^e (:142-145)
            std::vector<float> datae(R.ntr());
            for (size_t k = 0; k < datae.size(); ++k) datae[k] = datar[k] + dataesc[k];
            for (size_t k = 0; k < nwf; ++k) wfr[k] = wfr[k] + wfe[k];
            write_stream(R.bin("datae_", is), datae.data(), datae.size());
            write_stream(R.bin("wavefield_exd", is), wfr.data(), wfr.size());
        }
    }
}

// ---------------------------------------------------------------- modes 1/3 + 2 fused, batched
void mode13_fused(Run &R, const Acq &datum)
{
    const daw_par &p = R.p;
    const int nr = R.acq.nrec, nd = datum.nrec, nt = p.nt;
    const bool out = p.oio == 1, symes = p.mode == 3;
    const int itnlim = (int)p.maxiter;
    const size_t nlog = 9 + (size_t)5 * std::max(itnlim, 1);
    const int B = std::min(R.batch, R.acq.nsrc);
    Dev<float> s_dev(nt), d_obs((size_t)B * R.ntr()), d_datum((size_t)B * nd * nt);
    Dev<float> d_de(out ? (size_t)B * R.ntr() : 0), d_e(out ? (size_t)B * R.ntr() : 0);
    hip(hipMemcpy(s_dev.p, R.s.data(), nt * sizeof(float), hipMemcpyHostToDevice), "copy wavelet");
    std::vector<float> h_obs((size_t)B * R.ntr()), h_datum((size_t)B * nd * nt), h_v;
    std::vector<double> logs((size_t)B * nlog);
    for (int s0 = 0; s0 < R.acq.nsrc; s0 += B) {
        const int ns = std::min(B, R.acq.nsrc - s0);
        for (int s = 0; s < ns; ++s) {
            const std::vector<float> d = read_record(R.bin("dataobs_org_", s0 + s + 1), R.ntr());
            std::copy(d.begin(), d.end(), h_obs.begin() + (size_t)s * R.ntr());
        }
        hip(hipMemcpy(d_obs.p, h_obs.data(), (size_t)ns * R.ntr() * sizeof(float), hipMemcpyHostToDevice), "copy");
        const std::vector<int> src = rows(R.acq.src_tab, R.acq.nsrc, s0, ns);
        daw_redatum_args a{};
        a.c0 = R.c0.data();
        a.n1 = p.n1; a.n2 = p.n2; a.npml = p.npml; a.h = p.h; a.dt = p.dt;
        a.ofs = p.ofs; a.nt = nt; a.type_src = p.type_src;
        a.s_dev = s_dev.p;
        a.nshots = ns; a.nsrc_total = ns; a.src_tab = src.data();
        a.nrec = nr; a.rec_tab = R.acq.rec_tab.data();
        a.ndat = nd; a.dat_tab = datum.rec_tab.data();
        a.dobs_dev = d_obs.p;
        a.mu = p.mu; a.maxiter = itnlim; a.rtol = p.error;
        a.datum_dev = d_datum.p;
        a.datae_dev = out ? d_e.p : nullptr;
        a.deltade_dev = out ? d_de.p : nullptr;
        a.minres_log_host = out ? logs.data() : nullptr;
        a.filter = 1;                // mode 2's band-pass (:240)
        a.flow = p.flow; a.fhig = p.fhig; a.llow = p.llow; a.lhig = p.lhig;
        a.method = symes ? 3 : 1;
        check(daw_redatum_batch(&a), "redatum_batch");
        hip(hipMemcpy(h_datum.data(), d_datum.p, (size_t)ns * nd * nt * sizeof(float), hipMemcpyDeviceToHost), "copy");
        for (int s = 0; s < ns; ++s)
            write_stream(R.bin("dataobs_redatum_", s0 + s + 1), h_datum.data() + (size_t)s * nd * nt, (size_t)nd * nt);
        if (R.dto) {                 // on the device gathers, all shots of the batch in one call
            const int no = R.nto();
            Dev<float> d_dto((size_t)ns * nd * no);
            check(daw_subresample_dev(d_datum.p, d_dto.p, nt, ns * nd, no, p.dt, p.dto), "subresample");
            std::vector<float> h_dto((size_t)ns * nd * no);
            hip(hipMemcpy(h_dto.data(), d_dto.p, h_dto.size() * sizeof(float), hipMemcpyDeviceToHost), "copy");
            for (int s = 0; s < ns; ++s)
                write_stream(R.bin("dataobs_redatum_dto_", s0 + s + 1), h_dto.data() + (size_t)s * nd * no,
                             (size_t)nd * no);
        }
        if (!out) continue;
        h_v.resize((size_t)ns * R.ntr());
        hip(hipMemcpy(h_v.data(), d_de.p, h_v.size() * sizeof(float), hipMemcpyDeviceToHost), "copy");
        for (int s = 0; s < ns; ++s) {
            const int is = s0 + s + 1;
            write_unreversed(R.bin("deltade_", is), h_v.data() + (size_t)s * R.ntr(), nt, nr);
            check(daw_write_minres_log(("MINRES_" + shot_name(is) + ".txt").c_str(), nr * nt, itnlim, p.error,
                                       logs.data() + (size_t)s * nlog),
                  "MINRES log");
        }
        hip(hipMemcpy(h_v.data(), d_e.p, h_v.size() * sizeof(float), hipMemcpyDeviceToHost), "copy");
        for (int s = 0; s < ns; ++s)         // Symes: datae = dataesc (no reduced data)
            write_stream(R.bin(symes ? "dataesc_" : "datae_", s0 + s + 1), h_v.data() + (size_t)s * R.ntr(), R.ntr());
    }
}

// ---------------------------------------------------------------- mode 2
void mode2(Run &R)
{
    const daw_par &p = R.p;
    const int nr = R.acq.nrec;
    std::vector<float> sis(R.ntr()), sisf(R.ntr());
    for (int is = 1; is <= R.acq.nsrc; ++is) {
        const std::vector<float> wf = read_record(R.bin("wavefield_exd", is), R.plane() * p.nt);
        check(daw_subwavefieldresample(p.n1, p.n2, p.npml, p.nt, p.dt, R.acq.rec_tab.data(), nr, sis.data(), wf.data()),
              "subwavefieldresample");
        check(daw_filmartin2d(sis.data(), nr, p.nt, p.dt, p.flow, p.fhig, p.llow, p.lhig, sisf.data()), "filmartin2d");
        write_stream(R.bin("dataobs_redatum_", is), sisf.data(), sisf.size());
        if (R.dto) {
            std::vector<float> d((size_t)nr * R.nto());
            check(daw_subresample(sisf.data(), d.data(), p.nt, nr, R.nto(), p.dt, p.dto), "subresample");
            write_stream(R.bin("dataobs_redatum_dto_", is), d.data(), d.size());
        }
    }
}

// Every per-shot input file the mode reads, present and at least as long as
// the record it needs, before any GPU work: a missing file fails the run at
// once, not after the shots before it.  The reference finds out at the shot.
void preflight(const Run &R)
{
    const char *stem = nullptr;
    size_t need = 0;
    if (R.p.mode == 1 || R.p.mode == 3) { stem = "dataobs_org_"; need = R.ntr(); }
    if (R.p.mode == 2) { stem = "wavefield_exd"; need = R.plane() * R.p.nt; }
    if (!stem) return;
    for (int is = 1; is <= R.acq.nsrc; ++is) {
        const std::string path = R.bin(stem, is);
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) die("cannot open " + path);
        const int ok = std::fseek(f, 0, SEEK_END);
        const long len = std::ftell(f);
        std::fclose(f);
        if (ok != 0 || len < 0 || (size_t)len < need * sizeof(float))
            die(path + ": expected " + std::to_string(need) + " REAL*4 values, found " +
                std::to_string(len < 0 ? 0 : len / 4));
    }
}

void usage()
{
    std::fprintf(stderr,
                 "usage: dawredatum [--par FILE] [--datum ACQUISITION_FILE] [--batch N] [--dto] [--dry-run] [--quiet]\n"
                 "Runs the mode that redatum.par selects in the current directory (see the header of\n"
                 "dawredatum.cpp for the files each mode reads and writes).\n");
}

}  // namespace

int main(int argc, char **argv)
{
    Run R;
    std::string par = "redatum.par", datum_file;
    bool dry = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { usage(); die("missing value after " + a); }
            return argv[++i];
        };
        if (a == "--par") par = next();
        else if (a == "--datum") datum_file = next();
        else if (a == "--batch") R.batch = std::atoi(next().c_str());
        else if (a == "--dry-run") dry = true;
        else if (a == "--quiet") R.quiet = true;
        else if (a == "--dto") R.dto = true;
        else if (a == "-h" || a == "--help") { usage(); return 0; }
        else { usage(); die("unknown argument " + a); }
    }
    if (R.batch < 1) die("--batch must be at least 1");
    const auto t0 = std::chrono::steady_clock::now();
    daw_par &p = R.p;
    check(daw_readpar(par.c_str(), &p), "parameter file");
    if (p.n1 < 1 || p.n2 < 1 || p.nt < 1 || !(p.h > 0) || !(p.dt > 0)) die("bad grid in " + par);
    if (R.dto && !(p.dto > 0)) die("--dto needs dto > 0 in " + par);
    if (!R.quiet) {
        std::printf(" Run mode: %d\n", p.mode);
        std::printf(" Space & time grid dimension and interval (n1 n2 h nt dt) %d %d %g %d %g\n", p.n1, p.n2, p.h, p.nt,
                    p.dt);
        std::printf(" frequency band %g %g %g %g\n Penalty term %g\n", p.flow, p.fhig, p.llow, p.lhig, p.mu);
    }
    check(daw_kwsinc2d(p.ofs), "kwsinc2d");
    // wavelet and models (main_redatum_DAW.f90:98-118): direct access, record 1;
    // the wavelet is divided by h**2
    R.s = read_record(p.name_src, p.nt);
    const float h2 = p.h * p.h;
    for (float &v : R.s) v = v / h2;
    R.ctrue = read_record(p.name_true, R.plane());
    R.c0 = read_record(p.name_init, R.plane());
    R.acq = read_acquisition(p.name_acqui, p);
    Acq datum;
    if (!datum_file.empty()) {
        if (p.mode != 1 && p.mode != 3) die("--datum applies to modes 1 and 3");
        datum = read_acquisition(datum_file.c_str(), p);
        if (datum.nrec < 1) die(datum_file + ": no receivers");
    }
    if (!R.quiet) {
        std::printf(" Number of sources and receivers %d %d\n", R.acq.nsrc, R.acq.nrec);
        // sub_redatum_DAW.f90:24
        if (p.dto > 0) std::printf(" Niquist samping number %d\n", R.nto());
    }
    if (dry) {
        auto dump = [](const char *name, const std::vector<int> &t, int n) {
            std::printf("%s %d\n", name, n);
            for (int i = 0; i < n; ++i)
                std::printf("%d %d %d %d %d\n", t[i], t[i + (size_t)n], t[i + 2 * (size_t)n], t[i + 3 * (size_t)n],
                            t[i + 4 * (size_t)n]);
        };
        dump("src_tab", R.acq.src_tab, R.acq.nsrc);
        dump("rec_tab", R.acq.rec_tab, R.acq.nrec);
        if (!datum_file.empty()) dump("dat_tab", datum.rec_tab, datum.nrec);
    }
    preflight(R);
    if (dry) return 0;
    switch (p.mode) {
    case 0: mode0(R); break;
    case 1:
    case 3:
        if (!datum_file.empty()) mode13_fused(R, datum);
        else mode13_plain(R);
        break;
    case 2: mode2(R); break;
    default:
        std::fprintf(stderr, "dawredatum: mode %d selects no work (sub_redatum_DAW.f90:53-245)\n", p.mode);
    }
    check(daw_deallocatekws2d(), "deallocatekws2d");
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!R.quiet) std::printf(" ELAPSED TIME = %.3f\n", el);
    return 0;
}
