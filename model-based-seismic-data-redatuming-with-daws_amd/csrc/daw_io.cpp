// On-disk formats of the reference driver, host side:
//  - redatum.par as subreadpar reads it (subreadpar.f90:9-21), with Fortran
//    list-directed input semantics;
//  - the acquisition file (main_redatum_DAW.f90:120-123, subacqui.f90:29-57);
//  - per-shot file names (subname.f90:2-20);
//  - the MINRES log MINRES_<shot>.txt (sublinearminres.f90:51-61,
//    minresModule.f90:66-67, 308-325, 337-342, formats :346-358).
#include "daw_io.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace daw {

// ------------------------------------------------------------ list-directed input
//
// A READ(unit,*) statement takes values from as many records as it needs;
// values are separated by blanks or a comma; a slash ends the statement
// (remaining items keep their values); "r*c" repeats c r times and "r*" or
// an empty field between commas is a null value (item unchanged).  Each READ
// starts on a new record and drops what is left of the last one it read.
class ListReader {
public:
    explicit ListReader(std::istream &in) : in_(in) {}
    int line() const { return line_; }

    // Tokens for one READ of n items; a null token is returned as "\x01".
    bool read(int n, std::vector<std::string> &out, std::string &err)
    {
        out.clear();
        pending_.clear();
        bool slash = false;
        while ((int)out.size() < n) {
            if (pending_.empty()) {
                if (slash) break;
                std::string rec;
                if (!std::getline(in_, rec)) {
                    err = "end of file";
                    return false;
                }
                ++line_;
                split(rec, pending_, slash);
                continue;
            }
            out.push_back(pending_.front());
            pending_.erase(pending_.begin());
        }
        while ((int)out.size() < n) out.push_back("\x01");
        out.resize(n);
        return true;
    }

private:
    static void push_value(const std::string &tok, std::vector<std::string> &v)
    {
        const size_t star = tok.find('*');
        if (star != std::string::npos && star > 0 &&
            tok.find_first_not_of("0123456789") == star) {
            const int r = std::atoi(tok.substr(0, star).c_str());
            const std::string c = tok.substr(star + 1);
            for (int k = 0; k < r; ++k) v.push_back(c.empty() ? "\x01" : c);
            return;
        }
        v.push_back(tok);
    }

    static void split(const std::string &rec, std::vector<std::string> &v, bool &slash)
    {
        size_t i = 0;
        const size_t n = rec.size();
        bool comma = false;         // the last separator consumed was a comma
        while (i < n) {
            const char ch = rec[i];
            if (ch == ' ' || ch == '\t' || ch == '\r') { ++i; continue; }
            if (ch == '/') { slash = true; return; }
            if (ch == ',') {        // two commas with nothing between: a null value
                if (comma) v.push_back("\x01");
                comma = true;
                ++i;
                continue;
            }
            std::string tok;
            if (ch == '\'' || ch == '"') {          // delimited character constant
                const char q = ch;
                ++i;
                while (i < n) {
                    if (rec[i] == q) {
                        if (i + 1 < n && rec[i + 1] == q) { tok += q; i += 2; continue; }
                        ++i;
                        break;
                    }
                    tok += rec[i++];
                }
            } else {
                while (i < n && rec[i] != ' ' && rec[i] != '\t' && rec[i] != ',' && rec[i] != '/' &&
                       rec[i] != '\r')
                    tok += rec[i++];
            }
            push_value(tok, v);
            comma = false;
            // a value is ended by blanks, optionally followed by one comma
            while (i < n && (rec[i] == ' ' || rec[i] == '\t' || rec[i] == '\r')) ++i;
            if (i < n && rec[i] == ',') { ++i; comma = true; }
        }
    }

    std::istream &in_;
    std::vector<std::string> pending_;
    int line_ = 0;
};

static bool is_null(const std::string &t) { return t == "\x01"; }

static bool parse_int(const std::string &t, int &v)
{
    if (is_null(t)) return true;
    const char *s = t.c_str();
    char *end = nullptr;
    const long x = std::strtol(s, &end, 10);
    if (end == s || *end != '\0') return false;
    v = (int)x;
    return true;
}

// A real as list-directed input takes it: F, E or D forms, optional exponent
// letter ("1.0-4" is 1.0e-4).
static bool parse_real(const std::string &t, float &v)
{
    if (is_null(t)) return true;
    std::string s = t;
    for (char &c : s)
        if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'e';
    size_t k = 1;
    while (k < s.size() && (std::isdigit((unsigned char)s[k]) || s[k] == '.')) ++k;
    if (k < s.size() && (s[k] == '+' || s[k] == '-') && s.find_first_of("eE") == std::string::npos)
        s.insert(k, "e");
    const char *p = s.c_str();
    char *end = nullptr;
    const double x = std::strtod(p, &end);
    if (end == p || *end != '\0') return false;
    v = (float)x;
    return true;
}

static void set_name(char *dst, const std::string &t)
{
    if (is_null(t)) return;
    std::snprintf(dst, kNameLen, "%s", t.c_str());
}

int read_par(const char *path, daw_par *p, std::string &err)
{
    std::ifstream in(path);
    if (!in) {
        err = std::string("cannot open ") + path;
        return -1;
    }
    ListReader rd(in);
    std::vector<std::string> t;
    auto stmt = [&](int n, const char *what) {
        if (!rd.read(n, t, err)) {
            err = std::string(path) + ": " + err + " while reading " + what;
            return false;
        }
        return true;
    };
    auto bad = [&](const char *what) {
        err = std::string(path) + " line " + std::to_string(rd.line()) + ": bad value in " + what;
        return -1;
    };
    // subreadpar.f90:10-20, one READ per line
    if (!stmt(1, "mode")) return -1;
    if (!parse_int(t[0], p->mode)) return bad("mode");
    if (!stmt(2, "name_true name_init")) return -1;
    set_name(p->name_true, t[0]);
    set_name(p->name_init, t[1]);
    if (!stmt(1, "name_acqui")) return -1;
    set_name(p->name_acqui, t[0]);
    if (!stmt(1, "name_src")) return -1;
    set_name(p->name_src, t[0]);
    if (!stmt(4, "flow fhig llow lhig")) return -1;
    if (!parse_real(t[0], p->flow) || !parse_real(t[1], p->fhig) || !parse_real(t[2], p->llow) ||
        !parse_real(t[3], p->lhig))
        return bad("flow fhig llow lhig");
    if (!stmt(7, "n1 n2 npml h nt dt dto")) return -1;
    if (!parse_int(t[0], p->n1) || !parse_int(t[1], p->n2) || !parse_int(t[2], p->npml) ||
        !parse_real(t[3], p->h) || !parse_int(t[4], p->nt) || !parse_real(t[5], p->dt) || !parse_real(t[6], p->dto))
        return bad("n1 n2 npml h nt dt dto");
    if (!stmt(2, "type_src type_rec")) return -1;
    if (!parse_int(t[0], p->type_src) || !parse_int(t[1], p->type_rec)) return bad("type_src type_rec");
    if (!stmt(1, "ofs")) return -1;
    if (!parse_int(t[0], p->ofs)) return bad("ofs");
    if (!stmt(1, "mu")) return -1;
    if (!parse_real(t[0], p->mu)) return bad("mu");
    if (!stmt(1, "oio")) return -1;
    if (!parse_int(t[0], p->oio)) return bad("oio");
    if (!stmt(2, "error maxiter")) return -1;
    if (!parse_real(t[0], p->error) || !parse_real(t[1], p->maxiter)) return bad("error maxiter");
    return 0;
}

int read_acqui(const char *path, int &nsrc, int &nrec, std::vector<float> &src_zx, std::vector<float> &rec_zx,
               std::string &err)
{
    std::ifstream in(path);
    if (!in) {
        err = std::string("cannot open ") + path;
        return -1;
    }
    ListReader rd(in);
    std::vector<std::string> t;
    auto fail_at = [&](const char *what) {
        err = std::string(path) + " line " + std::to_string(rd.line()) + ": " + what;
        return -1;
    };
    // counts (main_redatum_DAW.f90:121)
    if (!rd.read(2, t, err)) return fail_at("missing nsrc nrec");
    nsrc = nrec = 0;
    if (!parse_int(t[0], nsrc) || !parse_int(t[1], nrec) || nsrc < 0 || nrec < 0) return fail_at("bad nsrc nrec");
    // then one "z x" per source and per receiver (subacqui.f90:32-33, 48-49).
    // subacqui opens the file afresh and skips one record (:30-31), so the
    // points start on the second line even if the counts ran onto it
    std::ifstream in2(path);
    std::string first;
    if (!in2 || !std::getline(in2, first)) return fail_at("cannot reread");
    ListReader rp(in2);
    src_zx.assign((size_t)2 * nsrc, 0.0f);
    rec_zx.assign((size_t)2 * nrec, 0.0f);
    for (int i = 0; i < nsrc + nrec; ++i) {
        float *zx = i < nsrc ? &src_zx[(size_t)2 * i] : &rec_zx[(size_t)2 * (i - nsrc)];
        if (!rp.read(2, t, err)) {
            err = std::string(path) + " line " + std::to_string(rp.line() + 1) + ": " +
                  (i < nsrc ? "missing source coordinates" : "missing receiver coordinates");
            return -1;
        }
        if (!parse_real(t[0], zx[0]) || !parse_real(t[1], zx[1])) {
            err = std::string(path) + " line " + std::to_string(rp.line() + 1) + ": bad coordinate";
            return -1;
        }
    }
    return 0;
}

std::string shot_name(int i)
{
    if (i < 0 || i >= 100000) return std::string();
    char b[16];
    std::snprintf(b, sizeof b, "%05d", i);
    return b;
}

// ------------------------------------------------------------ Fortran edit descriptors

// Iw
static std::string f_int(long v, int w)
{
    char b[64];
    std::snprintf(b, sizeof b, "%*ld", w, v);
    std::string s = b;
    if ((int)s.size() > w) return std::string(w, '*');
    return s;
}

// 1PEw.d: one digit before the point, d after, exponent E+xx (two digits),
// or +xxx without the letter for |exponent| >= 100.
static std::string f_e1p(double v, int w, int d)
{
    std::string s;
    if (std::isnan(v)) {
        s = "NaN";
    } else if (std::isinf(v)) {
        s = v < 0 ? "-Infinity" : "Infinity";
        if ((int)s.size() > w) s = v < 0 ? "-Inf" : "Inf";
    } else {
        char b[96];
        std::snprintf(b, sizeof b, "%.*E", d, v);
        s = b;
        const size_t e = s.find('E');
        if (e != std::string::npos && s.size() - e - 2 >= 3) s.erase(e, 1);
    }
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// REAL values print as the reference's single-precision variables hold them
static double as_real(double v) { return (double)(float)v; }

static const char *const kEnter = " Enter MINRES.  ";
static const char *const kExit = " Exit  MINRES.  ";

// The stop messages, indexed by istop + 1 (istop = -1 .. 8), each padded to
// the common length of the message table (51 characters).
static std::string stop_message(int istop)
{
    static const char *const m[] = {
        "beta2 = 0.  If M = I, b and x are eigenvectors of A",
        "beta1 = 0.  The exact solution is  x = 0",
        "Requested accuracy achieved, as determined by rtol",
        "Reasonable accuracy achieved, given eps",
        "x has converged to an eigenvector",
        "Acond has exceeded 0.1/eps",
        "The iteration limit was reached",
        "Aprod  does not define a symmetric matrix",
        "Msolve does not define a symmetric matrix",
        "Msolve does not define a pos-def preconditioner",
    };
    std::string s = (istop >= -1 && istop <= 8) ? m[istop + 1] : "";
    s.resize(51, ' ');
    return s;
}

std::string minres_log_text(int n, int itnlim, float rtol, const double *blk)
{
    const double fac = blk[0], Anorm = blk[1], Acond = blk[2], rnorm = blk[3], Arnorm = blk[4], ynorm = blk[5];
    const int istop = (int)blk[7], itn = (int)blk[8];
    const double *tr = blk + kLogHeadDoubles;
    std::ostringstream o;
    const std::string sp5(5, ' ');
    // heading (format 1000); checkA and precon are .false. (sublinearminres.f90:42-47, 74)
    o << "\n\n" << kEnter << sp5 << "Solution of symmetric   Ax = b\n";
    o << " n      =" << f_int(n, 7) << sp5 << "checkA =" << "   F" << std::string(12, ' ') << "precon =" << "   F\n";
    o << " itnlim =" << f_int(itnlim, 7) << sp5 << "rtol   =" << f_e1p((double)rtol, 11, 2) << sp5
      << "shift  =" << f_e1p(as_real(-fac), 23, 14) << "\n";
    // iteration lines (formats 1200, 1300, 1500)
    for (int k = 1; k <= itn && k <= std::max(itnlim, 1); ++k) {
        const double *r = tr + (size_t)(k - 1) * kTraceDoubles;
        if (r[4] == 0.0) continue;
        if (k == 1) o << "\n\n" << sp5 << "itn" << std::string(8, ' ') << "x(1)" << std::string(10, ' ') << "norm(r)"
                      << "   norm(A)   cond(A)\n";
        o << f_int(k, 8) << f_e1p(as_real(r[0]), 19, 10) << f_e1p(as_real(r[1]), 10, 2)
          << f_e1p(as_real(r[2]), 10, 2) << f_e1p(as_real(r[3]), 10, 2) << "\n";
        if (k % 10 == 0) o << "\n";    // format 1500 is 1x: a trailing skip writes nothing
    }
    // final status (formats 2000, 3000)
    o << "\n" << kExit << sp5 << "istop =" << f_int(istop, 3) << std::string(14, ' ') << "itn   =" << f_int(itn, 8)
      << "\n";
    o << kExit << sp5 << "Anorm =" << f_e1p(as_real(Anorm), 12, 4) << sp5 << "Acond =" << f_e1p(as_real(Acond), 12, 4)
      << "\n";
    o << kExit << sp5 << "rnorm =" << f_e1p(as_real(rnorm), 12, 4) << sp5 << "ynorm =" << f_e1p(as_real(ynorm), 12, 4)
      << sp5 << "Arnorml=" << f_e1p(as_real(Arnorm), 12, 4) << "\n";
    o << kExit << sp5 << stop_message(istop) << "\n";
    return o.str();
}

}  // namespace daw
