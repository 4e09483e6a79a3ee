// crt_inv.hip -- a CRTomo tomodir without the Fortran host: crtomo.cfg, the measurements
// with their error model, the Gauss-Newton inversion loop of the CRTomo program and its
// output files, every heavy step a call of this library (forward, Jacobian,
// regularisation, starting lambda, update, uncertainty).
//
//   crtomo.cfg     src/rall.f90:141-593 (defaults, comment lines, the optional lines and
//                  their fallbacks, validation, the switches derived from mswitch)
//   crt.lamnull    src/rall.f90:735-790 (presets of the starting lambda, rewritten)
//   volt.dat       src/rdati.f90 (measurements and the error model: wmatdr, wmatdp,
//                  wmatd_cri; individual errors)
//   start model    src/bsigm0.f90 (homogeneous, a prior/start file, or the weighted data
//                  mean with the geometric factors)
//   loop           src/inv.f90:330-956 (potentials, polarity check, misfit, control
//                  output, convergence tests and the final phase improvement, the lambda
//                  search with its step-length control and parabolic fit, update)
//   outputs        inv.ctr (src/kont1.f90, src/kont2.f90), eps.ctr (src/inv.f90:272-289,
//                  src/dmisft.f90), cjg.ctr, run.ctr, rhoNN.mag/.modl/.pha/.sig and
//                  voltNN.dat (src/wout.f90), coverage.mag (src/bbsens.f90), the .diag
//                  files of the uncertainty (src/bmcm_mod.f90)
// Fixed-format lines (inv.ctr, eps.ctr, .modl) follow the reference's FORMAT statements;
// list-directed lines follow gfortran's layout (integers I11, REAL*4 1PG16.9E2, REAL*8
// 1PG25.17E3, each after one blank).
#include "crt_common.hpp"
#include "../../include/crtomo_hip.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <sys/stat.h>
#include <vector>

// a library call of the driver: a failure is reported with the call that made it
#define CRT_TRY(call)                                                                     \
    do {                                                                                  \
        const int e_ = (call);                                                            \
        if (e_) {                                                                         \
            fprintf(stderr, "crt_tomodir_crtomo: %s returned %d\n", #call, e_);           \
            return e_;                                                                    \
        }                                                                                 \
    } while (0)

namespace {

using cplx = std::complex<double>;
const double PI = std::acos(-1.0);

// ------------------------------------------------------------ Fortran output edits
std::string rjust(const std::string &s, int w) {
    if ((int)s.size() >= w) return s;
    return std::string(w - s.size(), ' ') + s;
}

// a non-finite value as gfortran writes it: NaN, Infinity / Inf, right-justified
std::string nonfinite(double x, int w) {
    std::string s = std::isnan(x) ? "NaN" : (w >= 8 + (x < 0) ? "Infinity" : "Inf");
    if (!std::isnan(x) && x < 0) s = "-" + s;
    return (int)s.size() > w ? std::string(w, '*') : rjust(s, w);
}

// Fw.d: the leading zero of |x| < 1 is dropped only where the field is too narrow
std::string fF(double x, int w, int d) {
    if (!std::isfinite(x)) return nonfinite(x, w);
    char b[512];
    snprintf(b, sizeof b, "%.*f", d, x);
    std::string s(b);
    if (d == 0) s += ".";  // Fortran prints the point of an Fw.0 field
    if (s.find_first_not_of("-0.") == std::string::npos && s[0] == '-') s = s.substr(1);
    if ((int)s.size() > w) {
        if (s.compare(0, 2, "0.") == 0) s = s.substr(1);
        else if (s.compare(0, 3, "-0.") == 0) s = "-" + s.substr(2);
    }
    if ((int)s.size() > w) return std::string(w, '*');
    return rjust(s, w);
}

// digits and decimal exponent of x rounded to d significant digits: x = 0.D * 10^e
void sig_digits(double x, int d, std::string &dig, int &e) {
    char b[128];
    snprintf(b, sizeof b, "%.*e", d - 1, std::fabs(x));
    dig.clear();
    const char *p = b;
    for (; *p && *p != 'e'; ++p)
        if (*p >= '0' && *p <= '9') dig.push_back(*p);
    e = atoi(p + 1) + 1;
}

// Ew.dEe (0.DDDE+XX); scale 1 (1P) gives D.DDDE+XX with d digits after the point
std::string fE(double x, int w, int d, int ee = 2, bool p1 = false) {
    if (!std::isfinite(x)) return nonfinite(x, w);
    std::string dig;
    int e = 0;
    if (x == 0.0) {
        dig.assign(p1 ? d + 1 : d, '0');
        e = p1 ? 0 : 0;
    } else {
        sig_digits(x, p1 ? d + 1 : d, dig, e);
        if (p1) e -= 1;
    }
    std::string s = x < 0 ? "-" : "";
    if (p1) s += dig.substr(0, 1) + "." + dig.substr(1);
    else s += "0." + dig;
    char eb[16];
    snprintf(eb, sizeof eb, "%0*d", ee, std::abs(e));
    s += std::string("E") + (e < 0 ? "-" : "+") + eb;
    if ((int)s.size() > w && s.compare(0, 2, "0.") == 0) s = s.substr(1);
    if ((int)s.size() > w) return std::string(w, '*');
    return rjust(s, w);
}

// Gw.dEe: F form with e+2 trailing blanks where 0.1 <= |x| (rounded) < 10^d, else E
std::string fG(double x, int w, int d, int ee = 2, bool p1 = false) {
    const int n = ee + 2;
    if (!std::isfinite(x)) return nonfinite(x, w);
    if (x == 0.0) return fF(0.0, w - n, d - 1) + std::string(n, ' ');
    std::string dig;
    int e = 0;
    sig_digits(x, d, dig, e);
    if (e >= 0 && e <= d) return fF(x, w - n, d - e) + std::string(n, ' ');
    return fE(x, w, d - (p1 ? 1 : 0), ee, p1);
}

std::string fI(long long v, int w) { return rjust(std::to_string(v), w); }
std::string fL(bool v) { return v ? "T" : "F"; }

// gfortran list-directed items (each with its leading blank)
std::string ldI(long long v) { return " " + fI(v, 11); }
std::string ldR4(double v) { return " " + fG((double)(float)v, 16, 9, 2, true); }
std::string ldR8(double v) { return " " + fG(v, 25, 17, 3, true); }
std::string ldA(const std::string &s) { return " " + s; }
std::string ldL(bool v) { return std::string(" ") + (v ? "T" : "F"); }

// a line built by column (Fortran Tn positioning, 1-based)
struct Line {
    std::string s;
    Line &at(int col, const std::string &v) {
        const size_t c = col - 1;
        if (s.size() < c + v.size()) s.resize(c + v.size(), ' ');
        s.replace(c, v.size(), v);
        return *this;
    }
};

struct Out {
    FILE *f = nullptr;
    bool open(const std::string &p, const char *mode) { f = fopen(p.c_str(), mode); return f != nullptr; }
    void line(const std::string &s) { if (f) { fputs(s.c_str(), f); fputc('\n', f); } }
    void close() { if (f) fclose(f); f = nullptr; }
    ~Out() { close(); }
};

// ------------------------------------------------------------ crtomo.cfg
std::string ltrim(const std::string &s) {
    size_t i = s.find_first_not_of(" \t\r");
    return i == std::string::npos ? std::string() : s.substr(i);
}

std::string rtrim(const std::string &s) {
    size_t i = s.find_last_not_of(" \t\r");
    return i == std::string::npos ? std::string() : s.substr(0, i + 1);
}

// src/pathmod.f90 clear_string: leading blanks removed, cut at the first blank
std::string clear_string(const std::string &s) {
    std::string t = ltrim(s);
    size_t b = t.find_first_of(" \t");
    return b == std::string::npos ? rtrim(t) : t.substr(0, b);
}

struct CfgReader {
    std::vector<std::string> lines;
    size_t p = 0;
    bool load(const std::string &path) {
        std::ifstream f(path);
        if (!f) return false;
        std::string l;
        while (std::getline(f, l)) {
            if (!l.empty() && l.back() == '\r') l.pop_back();
            lines.push_back(l);
        }
        return true;
    }
    // src/get_unit.f90 read_comments: lines whose first non-blank character is '#'
    void comments() {
        while (p < lines.size()) {
            std::string t = ltrim(lines[p]);
            if (!t.empty() && t[0] == '#') ++p;
            else break;
        }
    }
    // '(a)' into CHARACTER(80); false at the end of the file
    bool read_a(std::string &o) {
        if (p >= lines.size()) return false;
        o = lines[p++].substr(0, 80);
        return true;
    }
    // list-directed read of items of the given kinds ('i' integer, 'r' real, 'l'
    // logical): 0 read, 1 end of file, 2 conversion error (at the first item that does
    // not convert; its record is consumed); blank records are skipped, the rest of the
    // last record is dropped
    int read_list(std::vector<std::string> &tok, const char *kinds);
};

bool to_int(const std::string &t, long long &v);
bool to_real(const std::string &t, double &v);
bool to_logical(const std::string &t, bool &v);

int CfgReader::read_list(std::vector<std::string> &tok, const char *kinds) {
    tok.clear();
    const size_t n = std::strlen(kinds);
    while (tok.size() < n) {
        if (p >= lines.size()) return 1;
        std::string r = lines[p++];
        for (char &c : r)
            if (c == ',' || c == '\t') c = ' ';
        size_t q = 0;
        while (tok.size() < n) {
            q = r.find_first_not_of(' ', q);
            if (q == std::string::npos) break;
            size_t e = r.find(' ', q);
            std::string t = r.substr(q, e == std::string::npos ? std::string::npos : e - q);
            if (t[0] == '/') return tok.empty() ? 2 : 0;
            long long iv;
            double dv;
            bool lv;
            const char k = kinds[tok.size()];
            if ((k == 'i' && !to_int(t, iv)) || (k == 'r' && !to_real(t, dv)) || (k == 'l' && !to_logical(t, lv)))
                return 2;
            tok.push_back(t);
            q = e == std::string::npos ? r.size() : e;
        }
    }
    return 0;
}

bool to_int(const std::string &t, long long &v) {
    size_t i = 0;
    if (t.empty()) return false;
    if (t[0] == '+' || t[0] == '-') i = 1;
    if (i >= t.size()) return false;
    for (size_t k = i; k < t.size(); ++k)
        if (t[k] < '0' || t[k] > '9') return false;
    v = std::stoll(t);
    return true;
}

bool to_real(const std::string &t, double &v) {
    std::string s = t;
    for (char &c : s)
        if (c == 'd' || c == 'D') c = 'E';
    char *end = nullptr;
    v = strtod(s.c_str(), &end);
    return end && *end == '\0' && end != s.c_str();
}

bool to_logical(const std::string &t, bool &v) {
    size_t i = t[0] == '.' ? 1 : 0;
    if (i >= t.size()) return false;
    const char c = t[i];
    if (c == 'T' || c == 't') v = true;
    else if (c == 'F' || c == 'f') v = false;
    else return false;
    return true;
}

// formatted I2 input: the record's first two characters, blanks ignored (0 if blank)
int read_i2(const std::string &rec, long long &v) {
    std::string f = rec.substr(0, std::min<size_t>(2, rec.size()));
    std::string d;
    for (char c : f)
        if (c != ' ') d.push_back(c);
    if (d.empty()) { v = 0; return 0; }
    return to_int(d, v) ? 0 : 2;
}

bool file_exists(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

std::string join(const std::string &dir, const std::string &p) {
    if (p.empty() || p[0] == '/' || dir.empty()) return p;
    return dir + "/" + p;
}

void put_str(char *dst, const std::string &s) {
    std::strncpy(dst, s.c_str(), 80);
    dst[80] = '\0';
}

// src/rall.f90:141-593 up to the file reads; the prior / start model file is looked up
// relative to dir, crt.noisemod in dir
int read_crtomo_cfg(const std::string &path, crt_crtomo_cfg &c) {
    std::memset(&c, 0, sizeof c);
    // defaults (src/rall.f90:145-235)
    c.ltri = 0;
    c.betamgs = 1.0;
    c.lamfix = 0.0;
    c.llamf = 0;
    c.iseed = 1;
    CfgReader r;
    if (!r.load(path)) return 1;
    const std::string dir = path.find('/') == std::string::npos ? "." : path.substr(0, path.rfind('/'));
    std::vector<std::string> t;
    std::string s;
    long long iv = 0;
    double dv = 0.0;
    bool lv = false;
    auto rd_int = [&](int &o) -> int {
        r.comments();
        int e = r.read_list(t, "i");
        if (e == 1) return 2;
        if (e || !to_int(t[0], iv)) return 3;
        o = (int)iv;
        return 0;
    };
    auto rd_real = [&](double &o) -> int {
        r.comments();
        int e = r.read_list(t, "r");
        if (e == 1) return 2;
        if (e || !to_real(t[0], dv)) return 3;
        o = dv;
        return 0;
    };
    auto rd_log = [&](int &o) -> int {
        r.comments();
        int e = r.read_list(t, "l");
        if (e == 1) return 2;
        if (e || !to_logical(t[0], lv)) return 3;
        o = lv ? 1 : 0;
        return 0;
    };
    auto rd_path = [&](char *o) -> int {
        r.comments();
        if (!r.read_a(s)) return 2;
        put_str(o, clear_string(s));
        return 0;
    };
    int e;
    // mswitch: an unreadable first line (e.g. ***FILES***) leaves it 0
    r.comments();
    e = r.read_list(t, "i");
    if (e == 1) return 2;
    if (e == 0 && to_int(t[0], iv)) c.mswitch = (int)iv;
    if ((e = rd_path(c.elem)) || (e = rd_path(c.elec)) || (e = rd_path(c.volt)) || (e = rd_path(c.ramd))) return e;
    if ((e = rd_log(c.ldiff))) return e;
    if ((e = rd_path(c.dd0)) || (e = rd_path(c.dm0)) || (e = rd_path(c.dfm0))) return e;
    if (c.dm0[0]) {
        c.lstart = file_exists(join(dir, c.dm0)) ? 1 : 0;
        if (c.lstart) put_str(c.dstart, c.dm0);
        else c.dm0[0] = '\0';
    }
    if (c.lstart && c.ldiff && !c.dd0[0] && !c.dfm0[0]) {
        c.lprior = 1;
        c.ldiff = 0;
    }
    // seed and variance of the prior's noise: a line that does not read is skipped
    r.comments();
    e = r.read_list(t, "ir");
    if (e == 1) return 2;
    if (e == 0 && to_int(t[0], iv) && to_real(t[1], dv)) {
        c.iseedpri = (int)iv;
        c.modl_stdn = dv;
        c.lnsepri = c.lstart;
    }
    if ((e = rd_int(c.nx)) || (e = rd_int(c.nz)) || (e = rd_real(c.alfx)) || (e = rd_real(c.alfz))) return e;
    if ((e = rd_int(c.itmax)) || (e = rd_log(c.ldc)) || (e = rd_log(c.lrobust)) || (e = rd_log(c.lfphai))) return e;
    if ((e = rd_real(c.stabw0)) || (e = rd_real(c.stabm0)) || (e = rd_real(c.stabpA1)) || (e = rd_real(c.stabpB)) ||
        (e = rd_real(c.stabpA2)) || (e = rd_real(c.stabp0)))
        return e;
    if ((e = rd_log(c.lrho0)) || (e = rd_real(c.bet0)) || (e = rd_real(c.pha0)) || (e = rd_log(c.lagain))) return e;
    if ((e = rd_int(c.swrtr)) || (e = rd_log(c.lsink)) || (e = rd_int(c.nsink)) || (e = rd_log(c.lrandb2))) return e;
    r.comments();
    if (!r.read_a(s)) return 2;
    put_str(c.drandb, clear_string(s));
    // ltri as I2; an unreadable or missing line keeps 0 and is read again
    r.comments();
    if (r.p < r.lines.size() && read_i2(r.lines[r.p], iv) == 0) {
        ++r.p;
        c.ltri = (int)iv;
        if (c.ltri & 32) {
            c.llamf = 1;
            r.comments();
            const size_t p0 = r.p;
            if (r.read_list(t, "r") == 0 && to_real(t[0], dv)) c.lamfix = dv;
            else {
                c.lamfix = 1.0;
                r.p = p0;
            }
            c.ltri -= 32;
            if (c.ltri & 64) {
                c.llamf += 2;
                c.ltri -= 64;
            }
        }
        if (c.ltri > 15) return 3;  // the reference stops here
    }
    if (c.ltri > 4 && c.ltri < 15) {
        r.comments();
        const size_t p0 = r.p;
        if (r.read_list(t, "r") == 0 && to_real(t[0], dv)) c.betamgs = dv;
        else {
            c.betamgs = 0.1;
            r.p = p0;
        }
    }
    c.lffhom = c.stabp0 < 0.0;
    if (c.lffhom) c.stabp0 = -c.stabp0;
    c.lnse2 = file_exists(join(dir, "crt.noisemod")) ? 1 : 0;
    c.lnse = c.stabw0 < 0.0;
    if (c.lnse) {
        c.stabw0 = -c.stabw0;
        c.lnse2 = 0;
        r.comments();
        const size_t p0 = r.p;
        if (r.read_list(t, "i") == 0 && to_int(t[0], iv)) c.iseed = (int)iv;
        else {
            c.iseed = 1;
            r.p = p0;
        }
    }
    if (c.lnse2) c.lnse = 1;
    if ((c.nx <= 0 || c.nz <= 0) && c.ltri == 0) c.ltri = 1;
    // validation (src/rall.f90:534-579; nrmsdm is fixed at 1)
    const double pi = PI;
    if (c.ltri == 0 && (c.nx < 2 || c.nz < 2)) return 89;
    if (c.itmax < 0 || c.itmax >= 100) return 61;
    if (c.stabw0 <= 0.0 || c.stabm0 < 0.0) return 104;
    if (!c.ldc && c.lfphai &&
        ((c.stabp0 < 0.0 || c.stabpA2 < 0.0) || (c.stabp0 == 0.0 && c.stabpA2 == 0.0 && c.stabpA1 == 0.0)))
        return 105;
    if (c.lrho0 && (c.bet0 <= 0.0 || (!c.ldc && std::fabs(c.pha0) > 1e3 * pi))) return 91;
    // switches (src/rall.f90:583-640)
    c.lelerr = !c.lfphai && !c.ldc;
    c.lcov1 = (c.mswitch >> 1) & 1;
    c.lres = (c.mswitch >> 2) & 1;
    c.lcov2 = (c.mswitch >> 3) & 1;
    c.lgauss = (c.mswitch >> 4) & 1;
    c.lelerr = ((c.mswitch >> 5) & 1) || c.lelerr;
    c.lphi0 = (c.mswitch >> 7) & 1;
    c.lsytop = (c.mswitch >> 8) & 1;
    c.lvario = (c.mswitch >> 9) & 1;
    c.lverb = (c.mswitch >> 10) & 1;
    c.lres = c.lres || c.lcov2;
    c.lcov1 = c.lres || c.lcov1;
    c.lsens = 1;
    if (c.lstart || c.ldiff) c.lrho0 = 0;
    if (c.ldiff) c.ldc = 1;
    if (c.ldc) c.lfphai = 0;
    if (c.ltri == 15) c.lvario = 1;
    return 0;
}

// ------------------------------------------------------------ grid pieces
struct Grid {
    int sanz = 0, typanz = 0, mb = 0, maxn = 0, elanz = 0, relanz = 0;
    std::vector<int> typ, nelanz, selanz, nrel, rnr;
    std::vector<double> x, y, espx, espy;
};

struct Tok {
    std::vector<std::string> t;
    size_t p = 0;
    bool load(const std::string &path) {
        std::ifstream f(path);
        if (!f) return false;
        std::string w;
        while (f >> w) t.push_back(w);
        return true;
    }
    bool more() const { return p < t.size(); }
    long long i() { return std::stoll(t[p++]); }
    double d() {
        std::string s = t[p++];
        for (char &ch : s)
            if (ch == 'D' || ch == 'd') ch = 'E';
        return std::stod(s);
    }
};

// src/relem.f90:60-151 with the boundary-element owners of :197-229 and the centroids of
// :116-142
int read_grid(const std::string &path, Grid &g) {
    Tok k;
    if (!k.load(path)) return 1;
    try {
        g.sanz = (int)k.i();
        g.typanz = (int)k.i();
        g.mb = (int)k.i();
        if (g.sanz <= 0 || g.typanz <= 0) return 3;
        for (int t = 0; t < g.typanz; ++t) {
            g.typ.push_back((int)k.i());
            g.nelanz.push_back((int)k.i());
            g.selanz.push_back((int)k.i());
            g.maxn = std::max(g.maxn, g.selanz.back());
            if (g.typ.back() > 10) g.relanz += g.nelanz.back();
            else g.elanz += g.nelanz.back();
        }
        std::vector<int> snr(g.sanz);
        std::vector<double> sx(g.sanz), sy(g.sanz);
        for (int i = 0; i < g.sanz; ++i) {
            snr[i] = (int)k.i();
            sx[i] = k.d();
            sy[i] = k.d();
        }
        g.x.resize(g.sanz);
        g.y.resize(g.sanz);
        for (int i = 0; i < g.sanz; ++i) {
            if (snr[i] < 1 || snr[i] > g.sanz) return 3;
            g.x[i] = sx[snr[i] - 1];
            g.y[i] = sy[snr[i] - 1];
        }
        const int neltot = g.elanz + g.relanz;
        g.nrel.assign((size_t)neltot * g.maxn, 0);
        int iel = 0;
        for (int t = 0; t < g.typanz; ++t)
            for (int e = 0; e < g.nelanz[t]; ++e, ++iel)
                for (int q = 0; q < g.selanz[t]; ++q) g.nrel[(size_t)iel * g.maxn + q] = (int)k.i();
        for (int r = 0; r < g.relanz; ++r) g.rnr.push_back((int)k.i());
    } catch (...) {
        return 3;
    }
    const int s0 = g.selanz[0];
    std::vector<std::pair<long long, int>> owner;
    for (int e = 0; e < g.elanz; ++e)
        for (int a = 0; a < s0; ++a)
            for (int b = a + 1; b < s0; ++b) {
                long long n1 = g.nrel[(size_t)e * g.maxn + a], n2 = g.nrel[(size_t)e * g.maxn + b];
                if (n1 > n2) std::swap(n1, n2);
                owner.push_back({(n1 << 32) | n2, e + 1});
            }
    std::stable_sort(owner.begin(), owner.end(),
                     [](const std::pair<long long, int> &p, const std::pair<long long, int> &q) { return p.first < q.first; });
    for (int r = 0; r < g.relanz; ++r) {
        long long n1 = g.nrel[(size_t)(g.elanz + r) * g.maxn], n2 = g.nrel[(size_t)(g.elanz + r) * g.maxn + 1];
        if (n1 > n2) std::swap(n1, n2);
        const long long key = (n1 << 32) | n2;
        auto it = std::upper_bound(owner.begin(), owner.end(), key,
                                   [](long long v, const std::pair<long long, int> &p) { return v < p.first; });
        if (it != owner.begin() && (it - 1)->first == key) g.rnr[r] = (it - 1)->second;
    }
    // element centroids: the mean of the node coordinates in the reference's sum order
    g.espx.assign(g.elanz, 0.0);
    g.espy.assign(g.elanz, 0.0);
    for (int t = 0, iel = 0, ia = 0; t < g.typanz; ++t)
        for (int q = 0; q < g.nelanz[t]; ++q, ++iel) {
            if (g.typ[t] > 10) continue;
            for (int j = 0; j < g.selanz[t]; ++j) {
                const int nd = g.nrel[(size_t)iel * g.maxn + j] - 1;
                g.espx[ia] += g.x[nd];
                g.espy[ia] += g.y[nd];
            }
            g.espx[ia] /= g.selanz[t];
            g.espy[ia] /= g.selanz[t];
            ++ia;
        }
    return 0;
}

// ------------------------------------------------------------ measurements (rdati)
struct Data {
    int nanz = 0;
    bool lindiv = false;
    std::vector<int> strnr, vnr, wdfak;
    std::vector<cplx> dat;
    std::vector<double> wmatdr, wmatdp, wmatd_cri;
};

// src/rdati.f90 without the synthetic noise: dat = -log(R) - i phase/1000, the error
// model 1/wmatdr = (A/100 + B/R)^2 and the phase errors, or the individual errors
int read_data(const std::string &path, const crt_crtomo_cfg &c, int eanz, Data &D) {
    std::ifstream f(path);
    if (!f) return 1;
    std::vector<std::string> recs;
    std::string l;
    while (std::getline(f, l)) recs.push_back(l);
    size_t p = 0;
    auto toks = [&](size_t i) {
        std::vector<std::string> v;
        std::string r = recs[i];
        for (char &ch : r)
            if (ch == ',' || ch == '\t') ch = ' ';
        size_t q = 0;
        while ((q = r.find_first_not_of(' ', q)) != std::string::npos) {
            size_t e = r.find(' ', q);
            v.push_back(r.substr(q, e == std::string::npos ? std::string::npos : e - q));
            q = e == std::string::npos ? r.size() : e;
        }
        return v;
    };
    while (p < recs.size() && toks(p).empty()) ++p;
    if (p >= recs.size()) return 2;
    {
        auto v = toks(p++);
        long long n = 0;
        if (!to_int(v[0], n) || n <= 0) return 3;
        D.nanz = (int)n;
        bool li = false;
        if (v.size() > 1 && to_logical(v[1], li)) D.lindiv = li;
    }
    const int nanz = D.nanz;
    D.strnr.resize(nanz);
    D.vnr.resize(nanz);
    D.dat.resize(nanz);
    D.wmatdr.assign(nanz, 0.0);
    D.wmatdp.assign(nanz, 0.0);
    D.wmatd_cri.assign(nanz, 0.0);
    D.wdfak.assign(nanz, 1);
    bool crtf = true, first = true;
    std::vector<double> vals;
    for (int i = 0; i < nanz; ++i) {
        while (p < recs.size() && toks(p).empty()) ++p;
        if (p >= recs.size()) return 2;
        auto v = toks(p++);
        long long a = 0;
        if (!to_int(v[0], a)) return 3;
        if (first) {
            crtf = a - 10000 >= 0;  // 10002 40003 codes, else four electrode numbers
            first = false;
        }
        long long e1 = 0, e2 = 0, e3 = 0, e4 = 0;
        size_t q = 0;
        if (crtf) {
            long long b = 0;
            if (v.size() < 2 || !to_int(v[1], b)) return 3;
            D.strnr[i] = (int)a;
            D.vnr[i] = (int)b;
            q = 2;
        } else {
            if (v.size() < 4 || !to_int(v[1], e2) || !to_int(v[2], e3) || !to_int(v[3], e4)) return 3;
            e1 = a;
            D.strnr[i] = (int)(e1 * 10000 + e2);
            D.vnr[i] = (int)(e3 * 10000 + e4);
            q = 4;
        }
        const int need = c.ldc ? (D.lindiv ? 2 : 1) : (D.lindiv ? 4 : 2);
        vals.assign(need, 0.0);
        for (int k = 0; k < need; ++k)
            if (q + k >= v.size() || !to_real(v[q + k], vals[k])) return 3;
        double bet = vals[0], pha = c.ldc ? 0.0 : vals[1], stabw = 0.0, stabwp = 0.0;
        if (D.lindiv) {
            stabw = c.ldc ? vals[1] : vals[2];
            if (!c.ldc) {
                stabwp = vals[3];
                if (stabwp <= 0.0) return 88;
                stabwp *= 1e-3;
            }
            stabw /= bet;
            const int el1 = D.strnr[i] % 10000, el2 = (D.strnr[i] - el1) / 10000;
            const int el3 = D.vnr[i] % 10000, el4 = (D.vnr[i] - el3) / 10000;
            if (el1 == el2 || el3 == el4 || (((el1 == el3) || (el1 == el4)) && el1 != 0) ||
                (((el2 == el3) || (el2 == el4)) && el2 != 0))
                return 73;
            if (el1 < 0 || el1 > eanz || el2 < 0 || el2 > eanz || el3 < 0 || el3 > eanz || el4 < 0 || el4 > eanz)
                return 46;
            if (stabw <= 0.0) return 88;
        } else {
            if (!c.ldc) stabwp = (c.stabpA1 * std::pow(bet, c.stabpB) + 1e-2 * c.stabpA2 * std::fabs(pha) + c.stabp0) * 1e-3;
            if (bet <= 0.0) return 94;
            stabw = 1e-2 * c.stabw0 + c.stabm0 / bet;
        }
        if (bet <= 0.0) return 94;
        if (!c.ldc && std::fabs(pha) > 1e3 * PI) return 95;
        D.dat[i] = cplx(-std::log(bet), -pha / 1e3);
        D.wmatdr[i] = 1.0 / std::pow(stabw, 2.0);
        if (!c.ldc) {
            D.wmatd_cri[i] = 1.0 / (std::pow(stabw, 2.0) + std::pow(stabwp, 2.0));
            D.wmatdp[i] = 1.0 / std::pow(stabwp, 2.0);
        }
        const int el1 = D.strnr[i] % 10000, el2 = (D.strnr[i] - el1) / 10000;
        const int el3 = D.vnr[i] % 10000, el4 = (D.vnr[i] - el3) / 10000;
        if (el1 < 0 || el1 > eanz || el2 < 0 || el2 > eanz || el3 < 0 || el3 > eanz || el4 < 0 || el4 > eanz)
            return 46;
    }
    if (D.lindiv) {
        while (p < recs.size() && toks(p).empty()) ++p;
        if (p >= recs.size()) return 2;
        auto v = toks(p++);
        double s1 = 0.0, s2 = 0.0;
        if (!to_real(v[0], s1)) return 3;
        if (c.ldc) {
            if (s1 <= 0.0) return 88;
            for (auto &w : D.wmatdr) w *= s1 * s1;
        } else {
            if (v.size() < 2 || !to_real(v[1], s2)) return 3;
            if (s1 <= 0.0 || s2 <= 0.0) return 88;
            for (auto &w : D.wmatdr) w *= s1 * s1;
            for (auto &w : D.wmatdp) w *= s2 * s2;
        }
    }
    return 0;
}

// src/rsigma.f90: count, then magnitude and phase (mrad) per element
int read_model(const std::string &path, int elanz, std::vector<cplx> &sigma) {
    Tok k;
    if (!k.load(path)) return 1;
    try {
        const int n = (int)k.i();
        if (n != elanz) return 3;
        sigma.resize(n);
        for (int i = 0; i < n; ++i) {
            const double bet = k.d(), pha = k.d() / 1e3;
            if (bet < 1e-12) return 11;
            sigma[i] = cplx(std::cos(pha) / bet, -std::sin(pha) / bet);
        }
    } catch (...) {
        return 3;
    }
    return 0;
}

// ------------------------------------------------------------ the inversion
struct Inv {
    crt_ctx *ctx = nullptr;
    crt_crtomo_cfg C{};
    std::string dir, ramd;
    Grid G;
    Data D;
    int manz = 0, ncgmax = 0, kwnanz = 0, lbeta = 0;
    double amin = 0.0, amax = 0.0, lamnull_cri = 0.0, lamnull_fpi = 0.0;
    // konvmod / invmod state
    int it = 0, itr = 0, npol = 0, ncg = 0;
    double nrmsd = 0, betrms = 0, pharms = 0, rmsalt = 0, lamalt = 1, bdpar = 1, bdalt = 1, lam = 0, lammax = 0;
    double dlam = 1, dlalt = 1, alam = 0, rmsreg = 0, step = 1, stpalt = 1, rough = 0, lam_cri = 0;
    bool llam = false, ldlami = true, lstep = false, lfstep = false, ldlamf = false, lsetup = true, lsetip = false;
    bool lfpi = false, lfphai = false, l_bsmat = true;
    const double nrmsdm = 1.0, mqrms = 2e-2, eps = 1e-4, stpmin = 1e-3, bdmin = 1e-3, fstart = 0.5, fstop = 0.9;
    const int nlam = 30;
    std::vector<cplx> sigma, sigma2, sigmaa, sgmaa2, par, dpar, dpar2;
    std::vector<double> wmatd, cgres, cgres2, smatm;
    std::vector<int> nachbar;
    cplx sigma0;
    Out run;
    int kind = 0;

    std::string path(const std::string &f) const { return ramd + "/" + f; }
    static const double *dp(const std::vector<cplx> &v) { return reinterpret_cast<const double *>(v.data()); }
    static double *dp(std::vector<cplx> &v) { return reinterpret_cast<double *>(v.data()); }

    // ---- src/dmisft.f90 (with chkpo2), appending its table to eps.ctr where the
    // reference does
    int dmisft(bool setup) {
        Out ep;
        if (!ep.open(path("eps.ctr"), "a")) return 1;
        const bool table = (llam && !lstep) || setup;
        if (table) {
            ep.line("");
            ep.line(Line().at(1, lfpi ? "PIT#" : "IT#").at(7, fI(it, 4)).s);
            ep.line("Datum\teps\tpsi\tpol\tRe(d)\tRe(f(m))\tIm(d)\tIm(f(m))");
        }
        nrmsd = betrms = pharms = 0.0;
        int idum = 0;
        for (int i = 0; i < D.nanz; ++i) {
            int wdlok = 1;
            double redat = D.dat[i].real(), imdat = D.dat[i].imag(), resig = sigmaa[i].real(), imsig = sigmaa[i].imag();
            const int idat = imdat > PI / 2 ? -1 : (imdat <= -PI / 2 ? 1 : 0);
            const int isig = imsig > PI / 2 ? -1 : (imsig <= -PI / 2 ? 1 : 0);
            if (idat == 0 && isig != 0) {
                imsig += isig * PI;
                imdat -= std::copysign(PI, imdat);  // lpol is off
                wdlok = 0;
            } else if (idat != 0 && isig == 0) {
                wdlok = 0;
            } else if (idat != 0 && isig != 0) {
                imsig += isig * PI;
                imdat += idat * PI;
            }
            const cplx cdat(redat, imdat), csig(resig, imsig), cdum = cdat - csig;
            const double psi = lfpi ? std::sqrt(wmatd[i]) * std::fabs(cdum.imag()) : std::sqrt(wmatd[i]) * std::abs(cdum);
            if (table)
                ep.line(fI(i + 1, 6) + fG((float)(1.0 / std::sqrt(wmatd[i])), 12, 3) + " " + fF((float)psi, 10, 6) +
                        "  " + fI(wdlok, 3) + "  " + " " + fG(csig.real(), 15, 7) + " " + fG(cdat.real(), 15, 7) + " " +
                        fG(cdat.imag(), 15, 7) + " " + fG(csig.imag(), 15, 7));
            idum += wdlok;
            nrmsd += psi * psi * wdlok;
            betrms += D.wmatdr[i] * wdlok * cdum.real() * cdum.real();
            pharms += D.wmatdp[i] * wdlok * cdum.imag() * cdum.imag();
        }
        if (idum == 0) return 99;
        npol = D.nanz - idum;
        nrmsd = std::sqrt(nrmsd / idum);
        betrms = std::sqrt(betrms / idum);
        pharms = std::sqrt(pharms / idum);
        return 0;
    }

    // ---- src/chkpol.f90 (lpol off): the potential electrodes of a datum whose modelled
    // phase crosses +-pi/2 are swapped, data with one side crossing are excluded
    int chkpol(bool setup) {
        (void)setup;
        bool swapped = false;
        for (int i = 0; i < D.nanz; ++i) {
            D.wdfak[i] = 1;
            const int e3 = D.vnr[i] % 10000, e4 = (D.vnr[i] - e3) / 10000;
            double redat = D.dat[i].real(), imdat = D.dat[i].imag(), resig = sigmaa[i].real(), imsig = sigmaa[i].imag();
            const int idat = imdat > PI / 2 ? -1 : (imdat <= -PI / 2 ? 1 : 0);
            const int isig = imsig > PI / 2 ? -1 : (imsig <= -PI / 2 ? 1 : 0);
            if (idat == 0 && isig != 0) {
                D.vnr[i] = e3 * 10000 + e4;
                imsig += isig * PI;
                sigmaa[i] = cplx(resig, imsig);
                imdat -= std::copysign(PI, imdat);
                D.dat[i] = cplx(redat, imdat);
                fprintf(run.f, "%4d : change polarity\n", i + 1);
                D.wdfak[i] = 0;
                swapped = true;
            } else if (idat != 0 && isig == 0) {
                D.wdfak[i] = 0;
            } else if (idat != 0 && isig != 0) {
                D.vnr[i] = e3 * 10000 + e4;
                imsig += isig * PI;
                sigmaa[i] = cplx(resig, imsig);
                imdat += idat * PI;
                D.dat[i] = cplx(redat, imdat);
                fprintf(run.f, "%4d : change polarity\n", i + 1);
                swapped = true;
            }
        }
        for (int i = 0; i < D.nanz; ++i)
            if (D.wdfak[i] == 0) fprintf(run.f, "%4d : excluded\n", i + 1);
        if (swapped) {
            // the swapped potential electrodes enter the next voltages and the Jacobian
            int e = crt_set_survey(ctx, D.nanz, D.strnr.data(), D.vnr.data());
            if (e) return e;
            std::vector<cplx> v(D.nanz), s(D.nanz);
            CRT_TRY(crt_voltages(ctx, 1, nullptr, dp(v), dp(s)));
        }
        return 0;
    }

    // ---- src/kont1.f90
    int kont1() {
        Out f;
        if (!f.open(path("inv.ctr"), "a")) return 1;
        auto l10 = [&](bool v, const char *t) { f.line(Line().at(1, fL(v)).at(20, t).s); };
        auto l11 = [&](double v, const std::string &t) { f.line(Line().at(1, fG(v, 11, 5)).at(20, t).s); };
        auto l12 = [&](long long v, const char *t) { f.line(Line().at(1, fI(v, 8)).at(20, t).s); };
        f.line("##");
        f.line("## Complex Resistivity Tomography (CRTomo)");
        f.line("##");
        f.line(std::string("## Library    ") + crt_version());
        f.line("##");
        f.line("");
        if (C.mswitch != 0) l12(C.mswitch, "#  mswitch");
        else f.line("***FILES***");
        f.line(std::string(rtrim(C.elem)) + " # FEM grid");
        f.line(std::string(rtrim(C.elec)) + " # Electrodes");
        f.line(std::string(rtrim(C.volt)) + " # Measurements");
        f.line(std::string(rtrim(C.ramd)) + " # Inversion results");
        l10(C.ldiff || C.lprior, "! difference inversion or (m - m_{prior})");
        f.line(rtrim(C.dd0));
        f.line(rtrim(C.dm0));
        f.line(rtrim(C.dfm0));
        if (C.lnsepri) f.line(ldI(C.iseedpri) + ldR8(C.modl_stdn));
        else f.line("***PARAMETERS***");
        l12(C.nx, "! nx-switch or # cells in x-direction");
        l12(C.nz, "! nz-switch or # cells in z-direction");
        l11(C.alfx, "! smoothing parameter in x-direction");
        l11(C.alfz, "! smoothing parameter in z-direction");
        l12(C.itmax, "! max. # inversion iterations");
        l10(C.ldc, "! DC inversion ?");
        l10(C.lrobust, "! robust inversion ?");
        l10(C.lfphai, "! final phase improvement ?");
        l11(C.stabw0, "! rel. resistance error level (%)  (parameter A1 in err(R) = A1*abs(R) + A2)");
        l11(C.stabm0, "! min. abs. resistance error (ohm) (parameter A2 in err(R) = A1*abs(R) + A2)");
        l11(C.stabpA1, "! phase error model parameter A1 (mrad/ohm^B) (in err(pha) = A1*abs(R)**B + A2*abs(pha) + A3)");
        l11(C.stabpB, "! phase error model parameter B  (-)          (in err(pha) = A1*abs(R)**B + A2*abs(pha) + A3)");
        l11(C.stabpA2, "! phase error model parameter A2 (%)          (in err(pha) = A1*abs(R)**B + A2*abs(pha) + A3)");
        l11(C.stabp0, "! phase error model parameter A3 (mrad)       (in err(pha) = A1*abs(R)**B + A2*abs(pha) + A3)");
        l10(C.lrho0, "! homogeneous background resistivity ?");
        l11(C.bet0, "! background magnitude (ohm*m)");
        l11(C.pha0, "! background phase (mrad)");
        l10(C.lagain, "! Another dataset?");
        l12(C.swrtr, "! 2D (=0) or 2.5D (=1)");
        l10(C.lsink, "! fictitious sink ?");
        l12(C.nsink, "! fictitious sink node number");
        l10(C.lrandb2, "! boundary values ?");
        f.line(rtrim(C.drandb));
        if (C.llamf & 1) f.line(fI(C.ltri + 32 + ((C.llamf & 2) ? 64 : 0), 2));
        else f.line(fI(C.ltri, 2));
        f.line(fI(C.ltri, 2));
        if (C.ltri & 32) f.line(ldR8(C.lamfix));
        if (C.ltri > 4 && C.ltri < 15) f.line(ldR8(C.betamgs));
        auto l100 = [&](const char *a, bool v) { f.line(Line().at(1, a).at(39, fL(v)).s); };
        auto l101 = [&](const char *a, double v) { f.line(Line().at(1, a).at(30, fG(v, 10, 5)).s); };
        auto l102 = [&](const char *a, long long v) { f.line(Line().at(1, a).at(30, fI(v, 8)).s); };
        auto l101i = [&](const char *a, long long v) { f.line(Line().at(1, a).at(30, fI(v, 10)).s); };
        f.line("");
        f.line("***Model stats***");
        l102("# Model parameters", manz);
        l102("# Data points", D.nanz);
        l100("Add data noise ?", C.lnse);
        l100("Couple to Err. Modl?", !C.lnse2);
        l102("    seed", C.iseed);
        l101("    Variance", 0.0);
        l100("Add model noise ?", C.lnsepri);
        l102("    seed", C.iseedpri);
        l101("    Variance", C.modl_stdn);
        f.line("");
        f.line("******** Regularization Part *********");
        l100("Prior regualrization", C.lprior);
        l100("Reference regualrization", false);
        l101i("Regularization-switch", C.ltri);
        l100("Regular grid smooth", C.ltri == 0);
        l100("Triangular regu", C.ltri == 1);
        l100("Triangular regu2", C.ltri == 2);
        l100("Levenberg damping", C.ltri == 3);
        l100("Marquardt damping", C.ltri == 4);
        l100("Minimum grad supp", C.ltri == 5);
        l100("MGS beta/sns1 (RM)", C.ltri == 6);
        l100("MGS beta/sns2 (RM)", C.ltri == 7);
        l100("MGS beta/sns1 (RB)", C.ltri == 8);
        l100("MGS beta/sns2 (RB)", C.ltri == 9);
        l100("TV (Huber)", C.ltri == 10);
        l100("Fixed lambda?", C.llamf & 1);
        if (C.llamf & 1) l101("Lambda=", C.lamfix);
        if (C.llamf & 2) l100("Cooling lambda?", true);
        if (C.nz < 0) {
            const double v = C.nz < -1 ? -(double)(float)C.nz : (double)std::max((float)manz, (float)D.nanz);
            f.line(std::string("Taking easy lam_0 : ") + ldR4(v));
        }
        f.line("");
        f.line("******** Additional output *********");
        l101i("mswitch", C.mswitch);
        l100("Read start model?", C.lstart);
        l100("Write coverage?", C.lsens);
        l100("Write MCM 1?", C.lcov1);
        l100("Write resolution?", C.lres);
        l100("Write MCM 2?", C.lcov2);
        l100("Using Gauss ols?", C.lgauss);
        l100("Forcing negative phase?", C.lphi0);
        l100("Calculate sytop?", C.lsytop);
        l100("Verbose?", C.lverb);
        l100("Error Ellipses?", C.lelerr);
        l100("Restart FPI with homogenous phase?", C.lffhom);
        f.line(Line().at(1, fL(D.lindiv)).at(18, rjust("! individual error ?", 20)).s);
        f.line("");
        f.line("***FIXED***");
        if (C.swrtr == 1) {
            f.line(Line().at(1, " # wavenumbers :").at(50, fI(kwnanz, 2)).s);
            f.line(Line().at(1, " Inverse Fourier transform range :").at(50, fG(amin, 11, 5)).at(62, "m").s);
            f.line(Line().at(50, fG(amax, 11, 5)).at(62, "m").s);
        }
        if (!C.lrho0 && !C.lstart) {
            const cplx r = 1.0 / sigma0;
            f.line(Line().at(1, " Background resistivity :").at(50, fG(std::abs(r), 11, 5)).at(62, "ohm*m").s);
            f.line(Line().at(50, fG(1e3 * std::atan2(r.imag(), r.real()), 11, 5)).at(62, "mrad").s);
        }
        f.line(Line().at(1, " Force negative phase ?").at(50, fL(C.lphi0)).s);
        f.line(Line().at(1, " Ratio dataset ?").at(50, "F").s);
        if (C.lrobust || C.lfphai) f.line(Line().at(1, " Min. L1 norm").at(50, fG(1.0, 11, 5)).s);
        f.line(Line().at(1, " Min. rel. decrease of data RMS :").at(50, fG(mqrms, 11, 5)).s);
        f.line(Line().at(1, " Min. steplength              :").at(50, fG(stpmin, 11, 5)).s);
        f.line(Line().at(1, " Min. stepsize (||\\delta m||) :").at(50, fG(bdmin, 11, 5)).s);
        f.line(Line().at(1, " Min. error in relaxation :").at(50, fG(eps, 11, 5)).s);
        f.line(Line().at(1, " Max. # relaxation iterations :").at(50, fI(ncgmax, 5)).s);
        f.line(Line().at(1, " Max. # regularization steps :").at(50, fI(nlam, 3)).s);
        f.line(Line().at(1, " Initial step factor :").at(50, fG(fstart, 11, 5)).s);
        f.line(Line().at(1, " Final   step factor :").at(50, fG(fstop, 11, 5)).s);
        f.line("");
        f.line(std::string(48, '-') + std::string(48, '-') + std::string(13, '-'));
        f.line("");
        f.line(Line().at(1, rjust("ID", 3)).at(5, "it.").at(11, "data RMS").at(23, "stepsize").at(34, " lambda ")
                   .at(46, " roughn.").at(58, "CG-steps").at(69, " mag RMS").at(81, " pha RMS").at(93, "- # data")
                   .at(105, "steplength").s);
        f.line("");
        return 0;
    }

    // ---- src/kont2.f90
    int kont2(bool setup) {
        Out f, g;
        if (!f.open(path("inv.ctr"), "a") || !g.open(path("cjg.ctr"), "a")) return 1;
        const std::string cdump = std::string(109, '*') + " ";
        ncg = cgres.empty() ? 0 : (int)cgres[0];
        const std::string id = lfpi ? "PIT" : "IT";
        if (setup) {
            f.line(cdump);
            f.line(Line().at(1, rjust(id, 3)).at(5, fI(it, 3)).at(11, fG(nrmsd, 10, 4)).at(69, fG(betrms, 10, 4))
                       .at(81, fG(pharms, 10, 4)).at(93, fI(npol, 4)).s);
            f.line(cdump);
        } else if (llam && !lstep) {
            f.line(cdump);
            f.line(Line().at(1, rjust(id, 3)).at(5, fI(it, 3)).at(11, fG(nrmsd, 10, 4)).at(23, fG(bdpar, 10, 4))
                       .at(34, fG(lam, 10, 4)).at(46, fG(rough, 10, 4)).at(58, fI(ncg, 6)).at(69, fG(betrms, 10, 4))
                       .at(81, fG(pharms, 10, 4)).at(93, fI(npol, 4)).at(105, fF(step, 5, 3)).s);
            f.line(cdump);
            g.line("");
            g.line(ldI(it));
            for (int k = 1; k <= ncg && k < (int)cgres.size(); ++k) g.line(ldR8(cgres[k]));
        } else {
            f.line(Line().at(1, rjust(lfpi ? "PUP" : "UP", 3)).at(5, fI(itr, 3)).at(11, fG(nrmsd, 10, 4))
                       .at(23, fG(bdpar, 9, 3)).at(34, fG(lam, 10, 4)).at(46, fG(rough, 10, 4)).at(58, fI(ncg, 6))
                       .at(105, fF(step, 5, 3)).s);
        }
        return 0;
    }

    // ---- src/wout.f90
    int wout() {
        char ci[8];
        snprintf(ci, sizeof ci, "%02d", it % 100);
        const std::string base = path(std::string("rho") + ci + ".");
        {
            FILE *lf = fopen((dir + "/inv.lastmod").c_str(), "w");
            if (lf) {
                fprintf(lf, " %s\n", (std::string(rtrim(C.ramd)) + "/rho" + ci + ".mag").c_str());
                fclose(lf);
            }
        }
        Out f;
        if (!f.open(base + "mag", "w")) return 1;
        f.line(ldI(G.elanz) + ldA("\t") + ldR8(betrms));
        for (int i = 0; i < G.elanz; ++i) {
            const cplx d = 1.0 / sigma[i];
            f.line(ldR4(G.espx[i]) + ldR4(G.espy[i]) + ldR4(std::log10(std::abs(d))));
        }
        f.close();
        if (!f.open(base + "modl", "w")) return 1;
        f.line(fI(G.elanz, 7));
        for (int i = 0; i < G.elanz; ++i) {
            const cplx d = 1.0 / sigma[i];
            const double d2 = (float)(1e3 * std::atan2(d.imag(), d.real()));
            f.line(" " + fG(1.0 / sigma[i].real(), 12, 4) + " " + fG(d2, 12, 4));
        }
        f.close();
        if (!C.ldc) {
            if (!f.open(base + "pha", "w")) return 1;
            f.line(ldI(G.elanz) + ldA("\t") + ldR8(pharms));
            for (int i = 0; i < G.elanz; ++i) {
                const cplx d = 1.0 / sigma[i];
                f.line(ldR4(G.espx[i]) + ldR4(G.espy[i]) + ldR4(1e3 * std::atan2(d.imag(), d.real())));
            }
            f.close();
            if (!f.open(base + "sig", "w")) return 1;
            f.line(ldI(G.elanz) + ldR8(betrms) + ldR8(pharms));
            for (int i = 0; i < G.elanz; ++i) f.line(ldR8(sigma[i].real()) + ldR8(sigma[i].imag()));
            f.close();
        }
        if (!f.open(path(std::string("volt") + ci + ".dat"), "w")) return 1;
        f.line(ldI(D.nanz));
        for (int i = 0; i < D.nanz; ++i) {
            std::string l = ldI(D.strnr[i]) + ldI(D.vnr[i]) + ldR4(1.0 / std::exp(sigmaa[i].real()));
            if (!C.ldc) l += ldR4(-1e3 * sigmaa[i].imag());
            f.line(l + ldI(D.wdfak[i]));
        }
        return 0;
    }

    // ---- forward: potentials, voltages (bvolti) and their logarithms
    int forward() {
        int e = crt_set_sigma(ctx, dp(sigma));
        if (!e) e = crt_forward(ctx);
        std::vector<cplx> volt(D.nanz);
        if (!e) e = crt_voltages(ctx, 1, nullptr, dp(volt), dp(sigmaa));
        return e;
    }

    // potentials and voltages of the current model on the device for the Jacobian, the
    // host's (polarity-checked) sigmaa kept: after a change of the inversion kind, where
    // the reference reuses its potentials
    int device_state() {
        int e = crt_set_sigma(ctx, dp(sigma));
        if (!e) e = crt_forward(ctx);
        std::vector<cplx> v(D.nanz), s(D.nanz);
        if (!e) e = crt_voltages(ctx, 1, nullptr, dp(v), dp(s));
        return e;
    }

    // ---- src/brough_mod.f90 (ltri 0, 1, 2; no prior)
    void brough() {
        rough = 0.0;
        const int ncol = C.ltri == 0 ? 3 : G.maxn + 1;
        for (int i = 0; i < manz; ++i) {
            cplx cd(0.0, 0.0);
            if (C.ltri == 0) {
                const int nx = C.nx;
                if (i > 0) cd = smatm[(size_t)1 * manz + (i - 1)] * par[i - 1];
                if (i < manz - 1) cd += smatm[(size_t)1 * manz + i] * par[i + 1];
                if (i + 1 > nx) cd += smatm[(size_t)2 * manz + (i - nx)] * par[i - nx];
                if (i + 1 < manz - nx + 1) cd += smatm[(size_t)2 * manz + i] * par[i + nx];
                cd += smatm[i] * par[i];
            } else {
                for (int j = 0; j < G.maxn; ++j) {
                    const int n = nachbar[(size_t)j * manz + i];
                    if (n != 0) cd += smatm[(size_t)j * manz + i] * par[n - 1];
                }
                cd += smatm[(size_t)(ncol - 1) * manz + i] * par[i];
            }
            if (lfpi) rough += cd.imag() * par[i].imag();
            else rough += (cd * std::conj(par[i])).real();
        }
    }

    // ---- src/blam0.f90: the presets, else the device computation
    int blam0() {
        if (((C.llamf & 1) || lamnull_cri > 2.220446049250313e-16) && !lfpi) {
            lammax = C.nz == -1 ? (double)std::max((float)manz, (float)D.nanz) : lamnull_cri;
            return 0;
        }
        if ((C.llamf & 1) || lamnull_fpi > 2.220446049250313e-16) {
            lammax = lamnull_fpi;
            return 0;
        }
        return crt_blam0(ctx, lfpi ? 1 : 0, wmatd.data(), D.wdfak.data(), C.alfx, C.alfz, &lammax);
    }

    // ---- src/update.f90 around the library's update: a new solve, or (llam) the saved
    // update of the previous lambda at the current step length
    int update() {
        const std::vector<cplx> p0 = par;
        if (!llam) {
            dpar2 = dpar;
            cgres2 = cgres;
            std::vector<cplx> dpn(manz);
            cgres.assign(ncgmax + 1, 0.0);
            int nc = 0;
            int e = crt_update(ctx, lam, eps, ncgmax, step, wmatd.data(), D.wdfak.data(), dp(D.dat), dp(sigmaa), dp(par),
                               dp(dpn), cgres.data(), &nc, nullptr, nullptr);
            if (e) return e;
            dpar = dpn;
        } else {
            dpar = dpar2;
            cgres = cgres2;
            for (int j = 0; j < manz; ++j) par[j] = p0[j] + step * dpar[j];
        }
        bdpar = 0.0;
        for (int j = 0; j < manz; ++j) {
            if (C.lphi0 && par[j].imag() < 0.0) par[j] = cplx(par[j].real(), 0.0);
            bdpar += std::norm(dpar[j]);
        }
        bdpar *= step;
        return 0;
    }

    int set_kind(int k) {
        kind = k;
        return crt_set_inversion_kind(ctx, k);
    }

    // ---- src/bbsens.f90
    int bbsens(const std::vector<double> &J) {
        std::vector<double> cs(manz, 0.0), csf(manz, 0.0);
        for (int j = 0; j < manz; ++j)
            for (int i = 0; i < D.nanz; ++i) {
                double w = std::sqrt(wmatd[i]) * D.wdfak[i];
                if (C.ldc) {
                    cs[j] += std::fabs(J[(size_t)j * D.nanz + i]) * w;
                    continue;
                }
                const cplx s(J[2 * ((size_t)j * D.nanz + i)], J[2 * ((size_t)j * D.nanz + i) + 1]);
                if (lfpi) {
                    csf[j] += std::fabs(s.real()) * w;
                    w = std::sqrt(C.lelerr ? D.wmatd_cri[i] : D.wmatdr[i]) * D.wdfak[i];
                }
                cs[j] += std::abs(s) * w;
            }
        const double mx = *std::max_element(cs.begin(), cs.end());
        Out f;
        if (!f.open(path("coverage.mag"), "w")) return 1;
        f.line(ldI(manz) + ldR8(mx));
        for (int i = 0; i < manz; ++i) f.line(ldR8(G.espx[i]) + ldR8(G.espy[i]) + ldR8(std::log10(cs[i] / mx)));
        f.close();
        if (lfpi) {
            const double mf = *std::max_element(csf.begin(), csf.end());
            if (!f.open(path("coverage.mag_fpi"), "w")) return 1;
            f.line(ldI(manz));
            for (int i = 0; i < manz; ++i) f.line(ldR8(G.espx[i]) + ldR8(G.espy[i]) + ldR8(std::log10(csf[i] / mf)));
            f.line(ldA("Max:") + ldR8(mf));
        }
        return 0;
    }

    // ---- src/bmcm_mod.f90 buncert: the library's dense analysis, the .diag files here
    int buncert(const std::vector<double> &J) {
        fprintf(run.f, "Calculating model uncertainty..\n");
        const double lu = lfpi ? lam_cri : lamalt;
        std::vector<double> w = wmatd;
        if (C.lelerr) w = D.wmatd_cri;
        else if (lfpi) w = D.wmatdr;
        std::vector<double> diag(5 * (size_t)manz);
        int e = crt_buncert(ctx, lu, w.data(), D.wdfak.data(), C.lres, C.lcov2, C.lgauss, nullptr, nullptr, nullptr,
                            nullptr, nullptr, diag.data());
        if (e && e != 108) return e;
        auto mm = [&](const double *d, double &lo, double &hi) {
            lo = *std::min_element(d, d + manz);
            hi = *std::max_element(d, d + manz);
        };
        Out f;
        double lo, hi;
        const double *ata = diag.data(), *reg = ata + manz, *cov = reg + manz, *res = cov + manz, *cov2 = res + manz;
        if (!f.open(path("ata.diag"), "w")) return 1;
        mm(ata, lo, hi);
        f.line(ldI(manz) + ldR8(lo) + ldR8(hi));
        for (int i = 0; i < manz; ++i) f.line(ldR8(ata[i]) + ldR8(std::log10(ata[i]) - std::log10(hi)));
        f.close();
        if (lfpi && !C.ldc) {
            std::vector<double> df(manz, 0.0);
            for (int k = 0; k < manz; ++k)
                for (int i = 0; i < D.nanz; ++i) {
                    const double re = J[2 * ((size_t)k * D.nanz + i)];
                    df[k] += re * re * D.wmatdp[i] * D.wdfak[i];
                }
            mm(df.data(), lo, hi);
            if (!f.open(path("ata.diag_fpi"), "w")) return 1;
            f.line(ldI(manz));
            for (int i = 0; i < manz; ++i) f.line(ldR8(df[i]) + ldR8(std::log10(df[i]) - std::log10(hi)));
            f.line(ldA("Max/Min:") + ldR8(hi) + ldA("/") + ldR8(lo));
            f.close();
        }
        if (!f.open(path("ata_reg.diag"), "w")) return 1;
        mm(reg, lo, hi);
        f.line(ldI(manz) + ldR8(lu));
        for (int i = 0; i < manz; ++i) f.line(ldR8(reg[i]) + ldR8(std::log10(reg[i])));
        f.line(ldA("Max/Min:") + ldR8(hi) + ldA("/") + ldR8(lo));
        f.close();
        if (e == 108) return 108;
        if (!f.open(path("cov1_m.diag"), "w")) return 1;
        mm(cov, lo, hi);
        f.line(ldI(manz) + ldR8(lu));
        for (int i = 0; i < manz; ++i) {
            const double s = std::sqrt(cov[i]);
            f.line(ldR8(s * (double)1e2f) + ldR8(sigma[i].real() * s) + ldR8(sigma[i].imag() * s));
        }
        f.line(ldA("Max/Min:") + ldR8(hi) + ldA("/") + ldR8(lo));
        f.close();
        if (C.lres) {
            if (!f.open(path("res_m.diag"), "w")) return 1;
            mm(res, lo, hi);
            f.line(ldI(manz) + ldR8(lu) + ldR8(lo) + ldR8(hi));
            for (int i = 0; i < manz; ++i)
                f.line(ldR8(std::fabs(res[i])) + ldR8(std::log10(std::fabs(res[i])) - std::log10(hi)));
            f.close();
            if (C.lcov2) {
                if (!f.open(path("cov2_m.diag"), "w")) return 1;
                mm(cov2, lo, hi);
                f.line(ldI(manz) + ldR8(lu));
                for (int i = 0; i < manz; ++i) {
                    const cplx ds = sigma[i] * std::sqrt(cov2[i]);
                    f.line(ldR8(std::sqrt(cov2[i]) * 1e2) + ldR8(ds.real()) + ldR8(ds.imag()));
                }
                f.line(ldA("Max/Min:") + ldR8(hi) + ldA("/") + ldR8(lo));
            }
        }
        return 0;
    }

    int run_inversion(int *iters, double *final_rms);
};

int Inv::run_inversion(int *iters, double *final_rms) {
    int e = 0;
    // ---- set-up (src/rall.f90:593-880, src/inv.f90:136-330)
    if (C.lagain) {
        fprintf(stderr, "crt_tomodir_crtomo: another dataset (lagain) is not supported\n");
        return CRT_ERR_UNSUPPORTED;
    }
    if (C.lnse || C.ldiff || C.lprior || C.lrobust || C.lvario || !(C.ltri == 0 || C.ltri == 1 || C.ltri == 2)) {
        fprintf(stderr,
                "crt_tomodir_crtomo: data noise, difference / prior inversion, robust inversion, variograms and "
                "regularisations other than ltri 0, 1, 2 are not supported\n");
        return CRT_ERR_UNSUPPORTED;
    }
    if (file_exists(join(dir, "decouplings.dat")) || file_exists(join(dir, "electrode_capacitances.dat"))) {
        fprintf(stderr, "crt_tomodir_crtomo: decouplings.dat / electrode_capacitances.dat are not supported\n");
        return CRT_ERR_UNSUPPORTED;
    }
    ramd = join(dir, rtrim(C.ramd));
    mkdir(ramd.c_str(), 0755);
    if ((e = read_grid(join(dir, C.elem), G))) return e;
    manz = C.ltri != 0 ? G.elanz : C.nx * C.nz;
    if (manz != G.elanz) return 50;
    ncgmax = manz;
    if (C.ltri < 5) ncgmax /= 10;
    std::vector<int> enr;
    {
        Tok k;
        if (!k.load(join(dir, C.elec))) return 1;
        try {
            const int n = (int)k.i();
            for (int i = 0; i < n; ++i) enr.push_back((int)k.i());
        } catch (...) {
            return 3;
        }
    }
    if (C.lsink && C.nsink > G.sanz) return 3;
    if ((e = read_data(join(dir, C.volt), C, (int)enr.size(), D))) return e;
    // crt.lamnull (src/rall.f90:735-790)
    {
        const std::string ln = join(dir, "crt.lamnull");
        std::ifstream lf(ln);
        if (lf) {
            if (!(lf >> lamnull_cri)) return 2;
            if (!(lf >> lamnull_fpi)) lamnull_fpi = (C.llamf & 1) ? C.lamfix : 0.0;
        } else if (C.llamf & 1) {
            lamnull_cri = lamnull_fpi = C.lamfix;
        } else if (C.nz == -1) {
            lamnull_cri = (double)std::max(manz, D.nanz);
            lamnull_fpi = 0.0;
        } else if (C.nz < 0) {
            lamnull_cri = std::fabs((double)C.nz);
            lamnull_fpi = 0.0;
        } else {
            lamnull_cri = lamnull_fpi = 0.0;
        }
        lf.close();
        FILE *w = fopen(ln.c_str(), "w");
        if (w) {
            fprintf(w, "%s\n%s\n", ldR8(lamnull_cri).c_str(), ldR8(lamnull_fpi).c_str());
            fclose(w);
        }
    }
    wmatd = C.ldc ? D.wmatdr : (C.lelerr ? D.wmatd_cri : D.wmatdr);
    if ((e = crt_set_grid(ctx, G.sanz, G.mb, G.typanz, G.typ.data(), G.nelanz.data(), G.selanz.data(), G.maxn,
                          G.nrel.data(), G.x.data(), G.y.data(), G.rnr.data(), (int)enr.size(), enr.data())))
        return e;
    CRT_TRY(crt_set_survey(ctx, D.nanz, D.strnr.data(), D.vnr.data()));
    if (C.swrtr == 0) {
        for (int t = 0; t < G.typanz; ++t)
            if (G.typ[t] == 11) return 110;
        // one wavenumber (1e-6 without Dirichlet nodes); the potentials are kpot itself
        double k0 = C.lrandb2 ? 0.0 : 1e-6, w0 = PI;
        kwnanz = 1;
        CRT_TRY(crt_set_wavenumbers(ctx, 1, &k0, &w0));
    } else {
        std::vector<double> kwn(64), kwnwi(64);
        CRT_TRY(crt_rwaven(ctx, &kwnanz, kwn.data(), kwnwi.data(), &amin, &amax));
    }
    if (C.lrandb2) {
        std::ifstream bf(join(dir, clear_string(C.drandb)));
        if (!bf) return 1;
        int n = 0;
        bf >> n;
        std::vector<int> bn(n);
        std::vector<double> bv(2 * (size_t)n, 0.0);
        for (int i = 0; i < n; ++i) {
            std::string v;
            if (!(bf >> bn[i] >> v)) return 3;
            if (v[0] == '(') {
                while (v.find(')') == std::string::npos) {
                    std::string m;
                    if (!(bf >> m)) return 3;
                    v += m;
                }
                for (char &ch : v)
                    if (ch == '(' || ch == ')' || ch == ',') ch = ' ';
                if (sscanf(v.c_str(), "%lf %lf", &bv[2 * i], &bv[2 * i + 1]) != 2) return 3;
            } else if (!to_real(v, bv[2 * i])) {
                return 3;
            }
        }
        CRT_TRY(crt_set_boundary_values(ctx, n, bn.data(), bv.data()));
    }
    CRT_TRY(crt_precal(ctx, C.lsytop, &lbeta));
    if (!lbeta) {
        if (!C.lsink) return 102;
    } else if (C.swrtr == 0) {
        return 106;
    }
    CRT_TRY(crt_set_fem_switches(ctx, 0, 0, C.lsink ? C.nsink : 0));
    if ((e = set_kind(C.ldc ? 1 : 0))) return e;
    // control files (src/inv.f90:262-289)
    {
        Out f;
        if (!f.open(path("inv.ctr"), "w") || (f.close(), !f.open(path("cjg.ctr"), "w"))) return 1;
        f.close();
        if (!run.open(path("run.ctr"), "w")) return 1;
        if (!f.open(path("eps.ctr"), "w")) return 1;
        if (C.ldc) {
            f.line("1/eps_r eps_r datum");
            for (int i = 0; i < D.nanz; ++i)
                f.line(fG(std::sqrt(D.wmatdr[i]), 12, 3) + "  " + fG(1.0 / std::sqrt(D.wmatdr[i]), 12, 3) + "  " +
                       fG(D.dat[i].real(), 14, 5));
        } else {
            f.line(Line().at(5, "1/eps_r").at(14, "1/eps_p").at(27, "1/eps").at(38, "eps_r").at(50, "eps_p")
                       .at(62, "eps").at(71, "-log(|R|)").at(87, "-Phase (rad)").s);
            for (int i = 0; i < D.nanz; ++i)
                f.line(fF(std::sqrt(D.wmatdr[i]), 10, 1) + fF(std::sqrt(D.wmatdp[i]), 10, 1) + fF(std::sqrt(wmatd[i]), 10, 1) +
                       "  " + fG(1.0 / std::sqrt(D.wmatdr[i]), 12, 3) + fG(1.0 / std::sqrt(D.wmatdp[i]), 12, 3) +
                       fG(1.0 / std::sqrt(wmatd[i]), 12, 3) + fG(D.dat[i].real(), 15, 7) + fG(D.dat[i].imag(), 15, 7));
        }
    }
    // starting model (src/bsigm0.f90)
    sigma.assign(G.elanz, cplx(0.0, 0.0));
    if (C.lstart) {
        if ((e = read_model(join(dir, C.dstart), G.elanz, sigma))) return e;
    } else if (C.lrho0) {
        for (auto &s : sigma) s = cplx(std::cos(C.pha0 / 1e3) / C.bet0, -std::sin(C.pha0 / 1e3) / C.bet0);
    } else if (lbeta) {
        std::vector<double> kfak(D.nanz);
        CRT_TRY(crt_bkfak(ctx, kfak.data()));
        cplx s0(0.0, 0.0);
        double dum = 0.0;
        for (int i = 0; i < D.nanz; ++i) {
            double imdat = D.dat[i].imag();
            const int idat = imdat > PI / 2 ? -1 : (imdat <= -PI / 2 ? 1 : 0);
            if (idat != 0) imdat += idat * PI;
            const double redat = D.dat[i].real() - std::log(std::fabs(kfak[i]));
            const double d2 = std::sqrt(wmatd[i]) * D.wdfak[i];
            s0 += cplx(redat, imdat) * d2;
            dum += d2;
        }
        if (dum == 0.0) return 99;
        for (auto &s : sigma) s = std::exp(s0 / dum);
        sigma0 = sigma[0];
    } else {
        return 100;
    }
    if (C.lrho0 || C.lstart) sigma0 = sigma[0];
    if ((e = kont1())) return e;
    sigma2 = sigma;
    sigmaa.assign(D.nanz, cplx(0.0, 0.0));
    sgmaa2 = sigmaa;
    par.assign(manz, cplx(0.0, 0.0));
    dpar = dpar2 = par;
    cgres.assign(ncgmax + 1, 0.0);
    cgres2 = cgres;
    lfphai = C.lfphai;
    if (C.llamf & 1) lamalt = C.lamfix;

    // ---- the loop (src/inv.f90:332-956)
    int errnr2 = 0;
    bool converged = false;
    while (!converged) {
        fprintf(run.f, " Iteration %3d, %3d : Calculating Potentials\n", it, itr);
        e = forward();
        if (e == CRT_ERR_ZEROVOLT) {
            sigma = sigma2;
            sigmaa = sgmaa2;
            break;
        }
        if (e) return e;
        if (lsetup || lsetip)
            if ((e = chkpol(lsetup || lsetip))) return e;
        if ((e = dmisft(lsetup || lsetip))) return e;
        if (nrmsd < 1e-12) nrmsd = nrmsdm * (1.0 - mqrms);
        if (it == 0 && (e = wout())) return e;
        if ((e = kont2(lsetup || lsetip))) return e;
        if (llam && !lstep) {
            if ((e = chkpol(lsetup || lsetip))) return e;
            if (stpalt == 0.0) errnr2 = 92;
            if (std::fabs(1.0 - rmsalt / nrmsd) <= mqrms) errnr2 = 81;
            if (std::fabs(1.0 - nrmsd / nrmsdm) <= mqrms && ldlamf) errnr2 = 80;
            if (it >= C.itmax) errnr2 = 79;
            if (std::fabs(1.0 - bdpar / bdalt) < bdmin) errnr2 = 109;
            if (errnr2 != 0) {
                if (lfphai && errnr2 != 79) {
                    fprintf(run.f, "CRI termination %d\n", errnr2);
                    if ((e = wout())) return e;
                    fprintf(run.f, " Final phase improvement\n");
                    {
                        Out f;
                        if (!f.open(path("inv.ctr"), "a")) return 1;
                        f.line("");
                        f.line(std::string(48, '-') + std::string(48, '-') + std::string(17, '-'));
                        f.line("");
                    }
                    wmatd = D.wmatdp;
                    lam_cri = lamalt;
                    lfpi = true;
                    lsetip = true;
                    lfphai = false;
                    llam = false;
                    ldlami = true;
                    lfstep = true;
                    step = 1.0;
                    errnr2 = 0;
                    if ((e = set_kind(2))) return e;
                    if (C.lffhom) {
                        {
                            Out f;
                            if (!f.open(path("inv.ctr"), "a")) return 1;
                            f.line("  ******* Resetting phase model ********");
                        }
                        for (auto &s : sigma)
                            s = cplx(std::cos(C.pha0 / 1e3) * std::abs(s), -std::sin(C.pha0 / 1e3) * std::abs(s));
                        lsetup = true;
                        continue;
                    } else {
                        lsetup = false;
                    }
                    if ((e = device_state())) return e;
                    if ((e = dmisft(lsetip))) return e;
                    if ((e = kont2(lsetip))) return e;
                } else {
                    break;
                }
            } else {
                lamalt = lam;
                if ((e = wout())) return e;
            }
        }
        if ((llam && !lstep) || lsetup || lsetip) {
            ++it;
            itr = 0;
            rmsreg = 0.0;
            dlam = 1.0;
            dlalt = 1.0;
            ldlamf = false;
            bdpar = 1.0;
            bdalt = 1.0;
            rmsalt = nrmsd;
            if (it > 1) lamalt = lam;
            sigma2 = sigma;
            sgmaa2 = sigmaa;
            fprintf(run.f, " Iteration %3d, %3d : Calculating Sensitivities\n", it, itr);
            CRT_TRY(crt_sens(ctx, 1, nullptr));
            if (l_bsmat) {
                smatm.assign((size_t)manz * (C.ltri == 0 ? 3 : G.maxn + 1), 0.0);
                CRT_TRY(crt_bsmatm(ctx, C.ltri, C.nx, C.nz, C.alfx, C.alfz, smatm.data()));
                if (C.ltri != 0) {
                    nachbar.assign((size_t)manz * (G.maxn + 1), 0);
                    std::vector<int> nr((size_t)G.elanz * G.maxn);
                    for (int i = 0; i < G.elanz; ++i)
                        for (int q = 0; q < G.maxn; ++q) nr[(size_t)i * G.maxn + q] = G.nrel[(size_t)i * G.maxn + q];
                    CRT_TRY(crt_bnachbar(G.elanz, G.maxn, nr.data(), nachbar.data()));
                }
                l_bsmat = false;
            }
        } else {
            sigma = sigma2;
            sigmaa = sgmaa2;
        }
        if (C.itmax == 0) {
            errnr2 = 109;
            break;
        }
        if (C.llamf & 1) {
            lam = C.lamfix;
            if ((C.llamf & 2) && it > 1) lam = C.lamfix / (2.0 * (double)(it - 1));
            llam = false;
            if (lsetup || lsetip) lsetup = lsetip = false;
        } else if (!lstep) {
            if (llam) {
                llam = false;
                step = 1.0;
            } else {
                ++itr;
                if ((((nrmsd < rmsreg && itr <= nlam) || (dlam > 1.0 && itr <= nlam)) && (!ldlamf || dlalt <= 1.0) &&
                     (std::fabs(1.0 - bdpar / bdalt) > bdmin) && (std::fabs(1.0 - rmsreg / nrmsdm) > mqrms)) ||
                    rmsreg == 0.0) {
                    if (lsetup || lsetip) {
                        fprintf(run.f, " Iteration %3d, %3d : Calculating 1st regularization parameter", it, itr);
                        if ((e = blam0())) return e;
                        fprintf(run.f, "lam_0 %s\n", fG(lammax, 10, 2).c_str());
                        lam = lammax;
                        lamalt = lammax;
                    } else {
                        dlalt = dlam;
                        if (ldlami) {
                            ldlami = false;
                            alam = std::max(std::fabs(std::log(nrmsd / nrmsdm)), std::log(1.0 + mqrms));
                            dlam = fstart;
                        } else {
                            alam = std::max(alam, std::fabs(std::log(nrmsd / nrmsdm)));
                            dlam = std::log(fstop) * std::copysign(1.0, std::log(nrmsd / nrmsdm)) +
                                   std::log(fstart / fstop) * std::log(nrmsd / nrmsdm) / alam;
                            dlam = std::exp(dlam);
                        }
                        lam = lam * dlam;
                        if (dlalt > 1.0 && dlam < 1.0) ldlamf = true;
                    }
                } else {
                    llam = true;
                    lam = lam / dlam;
                    if (lfstep) lfstep = false;
                    else {
                        lstep = true;
                        step = 0.5;
                    }
                }
                if (lsetup || lsetip) lsetup = lsetip = false;
                else if (!lstep) rmsreg = nrmsd;
            }
        } else {
            lstep = false;
            // src/parfit.f90: parabola through (0, rmsalt), (0.5, nrmsd), (1, rmsreg)
            const double fa = rmsalt, fb = nrmsd, fc = rmsreg, fmin = nrmsdm, smin = stpmin;
            const double a = 0.0, b = 0.5, c = 1.0;
            auto vertex = [&]() {
                const double bma = b - a, fbmfc = fb - fc, bmc = b - c, fbmfa = fb - fa;
                const double z = bma * bma * fbmfc - bmc * bmc * fbmfa, n = bma * fbmfc - bmc * fbmfa;
                return b - z / (2.0 * n);
            };
            if (fa > fb) {
                if (fb <= fc) {
                    if (fb <= fmin) step = (fa < fmin && fc > fa) ? b + (c - b) * (fmin - fb) / (fc - fb) : a + (b - a) * (fmin - fa) / (fb - fa);
                    else step = vertex();
                } else {
                    if (fb <= fmin) step = a + (b - a) * (fmin - fa) / (fb - fa);
                    else if (fc <= fmin) step = b + (c - b) * (fmin - fb) / (fc - fb);
                    else step = c;
                }
            } else if (fa < fb) {
                if (fb >= fc) {
                    if (fb >= fmin) step = (fa > fmin && fc < fa) ? b + (c - b) * (fmin - fb) / (fc - fb) : a + (b - a) * (fmin - fa) / (fb - fa);
                    else step = vertex();
                } else {
                    if (fb >= fmin) step = a + (b - a) * (fmin - fa) / (fb - fa);
                    else if (fc >= fmin) step = b + (c - b) * (fmin - fb) / (fc - fb);
                    else step = c;
                }
            } else if (fa == fb) {
                if (fb > fc) {
                    if (fb < fmin) step = vertex();
                    else if (fc <= fmin) step = b + (c - b) * (fmin - fb) / (fc - fb);
                    else step = c;
                } else if (fb < fc) {
                    if (fb > fmin) step = vertex();
                    else if (fc >= fmin) step = b + (c - b) * (fmin - fb) / (fc - fb);
                    else step = c;
                } else {
                    step = smin;
                }
            }
            step = std::min(c, std::max(smin, step));
            if (step == stpmin && stpalt == stpmin) stpalt = 0.0;
            else stpalt = step;
        }
        fprintf(run.f, "  Iteration %12d , %12d  : Updating\n", it, itr);
        // bpar, update, bsigma, brough
        for (int i = 0; i < G.elanz; ++i) par[i] = std::log(sigma[i]);
        bdalt = bdpar;
        if ((e = update())) return e;
        for (int i = 0; i < G.elanz; ++i) sigma[i] = std::exp(par[i]);
        brough();
        if (C.llamf & 1) llam = true;
    }
    // ---- termination (src/inv.f90:960-1040)
    const char *why = "";
    switch (errnr2) {
        case 92: why = " Min. step-length for 2nd time."; break;
        case 80: why = " Min. CHI."; break;
        case 81: why = " Min. rel. CHI decrease."; break;
        case 79: why = " Max. # iterations."; break;
        case 109: why = " Min. model changes reached"; break;
        default: break;
    }
    fprintf(run.f, " Iteration terminated:%s\n", why);
    if ((e = wout())) return e;
    std::vector<double> J;
    if (C.lsens || C.lcov1) {
        J.assign((C.ldc ? 1 : 2) * (size_t)D.nanz * manz, 0.0);
        CRT_TRY(crt_get_sens(ctx, J.data()));
    }
    if (C.lsens && (e = bbsens(J))) return e;
    if (C.lcov1 && (e = buncert(J))) return e;
    fprintf(run.f, "***finished***\n");
    {
        Out f;
        if (!f.open(path("inv.ctr"), "a")) return 1;
        f.line("***finished***");
    }
    if (iters) *iters = it;
    if (final_rms) *final_rms = nrmsd;
    return 0;
}

}  // namespace

extern "C" {

int crt_read_crtomo_cfg(const char *cfg_path, crt_crtomo_cfg *cfg) {
    if (!cfg_path || !cfg) return CRT_ERR_ARG;
    return read_crtomo_cfg(cfg_path, *cfg);
}

int crt_tomodir_crtomo(crt_ctx *ctx, const char *cfg_path, int *iterations, double *final_rms) {
    if (!ctx || !cfg_path) return CRT_ERR_ARG;
    Inv v;
    v.ctx = ctx;
    int e = read_crtomo_cfg(cfg_path, v.C);
    if (e) return e;
    const std::string p(cfg_path);
    v.dir = p.find('/') == std::string::npos ? "." : p.substr(0, p.rfind('/'));
    return v.run_inversion(iterations, final_rms);
}

}  // extern "C"
