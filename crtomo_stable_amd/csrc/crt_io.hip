// crt_io.hip -- the reference's on-disk formats on the host side of the library, so a
// CRMod tomodir runs without the Fortran host.
//
//   crmod.cfg  the ***FILES*** block: elem, elec, rho, config paths, then the
//              potentials / measurements / sensitivities switches with their output
//              paths (the layout of the reference's tests/*/exe/crmod.cfg)
//   elem.dat   src/relem.f90:60-151 (header, type table, node table, element node
//              lists, boundary-element neighbours) with the border-to-element check
//              of :197-229
//   elec.dat   src/relectr.f90 (count, then node numbers)
//   rho.dat    src/rsigma.f90:168 (count, then magnitude and phase in mrad per
//              element; sigma = (cos p - i sin p) / |rho|)
//   config.dat src/rdatm.f90 (count, then current and potential electrode codes)
//   volt.dat   src/wdatm.f90:58-76 (count, then strnr, vnr, REAL*4 magnitude and
//              1000 * phase; a phase near +-pi swaps the potential electrodes)
//   pot*.dat   src/bpot.f90 + src/wpot.f90 (one file per measurement: node x, y and the
//              complex potential of its current dipole)
//   sens*.dat  src/bsens.f90 + src/wsens.f90 (one file per measurement: |sum J|, its
//              phase in mrad, then element centroid x, y and dV/dsigma per element)
// Text files hold REAL*4 values; the reference writes them list-directed, whose layout
// is the compiler's, so the values (not the spacing) are what matches.
#include "crt_common.hpp"
#include "../../include/crtomo_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <sstream>
#include <string>
#include <vector>

namespace {

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
    bool more(size_t n = 1) const { return p + n <= t.size(); }
    long long i() { return std::stoll(t[p++]); }
    double d() {
        std::string s = t[p++];
        for (char &ch : s)
            if (ch == 'D' || ch == 'd') ch = 'E';  // Fortran exponent letter
        return std::stod(s);
    }
};

struct ElemFile {
    int sanz = 0, typanz = 0, mb = 0, maxn = 0, elanz = 0, relanz = 0;
    std::vector<int> typ, nelanz, selanz, nrel, rnr;
    std::vector<double> x, y;
};

// src/relem.f90:60-151 and the border check of :197-229
int read_elem(const std::string &path, ElemFile &g) {
    Tok k;
    if (!k.load(path)) return 1;
    try {
        g.sanz = (int)k.i();
        g.typanz = (int)k.i();
        g.mb = (int)k.i();
        if (g.sanz <= 0 || g.typanz <= 0) return CRT_ERR_ARG;
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
            if (snr[i] < 1 || snr[i] > g.sanz) return CRT_ERR_ARG;
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
        return CRT_ERR_ARG;
    }
    // every boundary element belongs to the area element (first type group) that holds
    // both of its nodes
    const int s0 = g.selanz[0];
    std::vector<std::pair<long long, int>> owner;
    owner.reserve((size_t)g.elanz * s0 * (s0 - 1) / 2);
    for (int e = 0; e < g.elanz; ++e)
        for (int a = 0; a < s0; ++a)
            for (int b = a + 1; b < s0; ++b) {
                long long n1 = g.nrel[(size_t)e * g.maxn + a], n2 = g.nrel[(size_t)e * g.maxn + b];
                if (n1 > n2) std::swap(n1, n2);
                owner.push_back({(n1 << 32) | n2, e + 1});
            }
    // the last element holding a pair wins, as in a forward scan that overwrites
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
    return 0;
}

int read_counted_ints(const std::string &path, int per, std::vector<long long> &v, int &n) {
    Tok k;
    if (!k.load(path)) return 1;
    try {
        n = (int)k.i();
        if (n < 0) return CRT_ERR_ARG;
        v.resize((size_t)n * per);
        for (auto &x : v) x = k.i();
    } catch (...) {
        return CRT_ERR_ARG;
    }
    return 0;
}

int read_rho(const std::string &path, std::vector<double> &sigma, int &n) {
    Tok k;
    if (!k.load(path)) return 1;
    try {
        n = (int)k.i();
        sigma.resize(2 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            const double bet = k.d(), pha = k.d() / 1e3;
            if (bet < 1e-12) return 11;  // src/rsigma.f90: errnr 11
            sigma[2 * i] = std::cos(pha) / bet;
            sigma[2 * i + 1] = -std::sin(pha) / bet;
        }
    } catch (...) {
        return CRT_ERR_ARG;
    }
    return 0;
}

struct CrmodCfg {
    std::string elem, elec, rho, config, pot, volt, sens, randb;
    bool wpot = false, wvolt = false, wsens = false;
    // src/fem.f90:177-193: another dataset, 2D (0) or 2.5D (1), fictitious sink and its
    // node, boundary values and their file, the mswitch bits (lana 0, wkfak 1, lsr 2)
    bool lagain = false, lsink = false, lrandb2 = false;
    int swrtr = 1, nsink = 0, mswitch = 0;
};

std::string join(const std::string &dir, const std::string &p) {
    if (p.empty() || p[0] == '/' || dir.empty()) return p;
    return dir + "/" + p;
}

int read_crmod_cfg(const std::string &path, CrmodCfg &c) {
    std::ifstream f(path);
    if (!f) return 1;
    std::vector<std::string> lines;
    std::string l;
    while (std::getline(f, l)) lines.push_back(l);
    auto first = [](const std::string &s) {
        std::istringstream is(s);
        std::string w;
        is >> w;
        return w;
    };
    if (lines.size() < 11) return CRT_ERR_ARG;
    const std::string dir = path.find('/') == std::string::npos ? "." : path.substr(0, path.rfind('/'));
    c.elem = join(dir, first(lines[1]));
    c.elec = join(dir, first(lines[2]));
    c.rho = join(dir, first(lines[3]));
    c.config = join(dir, first(lines[4]));
    auto flag = [&](const std::string &s) { const std::string w = first(s); return !w.empty() && (w[0] == 'T' || w[0] == 't'); };
    c.wpot = flag(lines[5]);
    c.pot = join(dir, first(lines[6]));
    c.wvolt = flag(lines[7]);
    c.volt = join(dir, first(lines[8]));
    c.wsens = flag(lines[9]);
    c.sens = join(dir, first(lines[10]));
    // the switches after the files (src/fem.f90:177-193); the reference stops with errnr 3
    // when one of them is missing, mswitch alone is optional
    if (lines.size() < 17) return 3;
    try {
        c.lagain = flag(lines[11]);
        c.swrtr = std::stoi(first(lines[12]));
        c.lsink = flag(lines[13]);
        c.nsink = std::stoi(first(lines[14]));
        c.lrandb2 = flag(lines[15]);
        c.randb = join(dir, first(lines[16]));
        c.mswitch = lines.size() > 17 && !first(lines[17]).empty() ? std::stoi(first(lines[17])) : 0;
    } catch (...) {
        return 3;
    }
    return 0;
}

// boundary.dat (src/rrandb.f90:85-120, complex case): rwdanz, then node and value per
// line; a value is "(re,im)" as Fortran writes a complex, or a real number
int read_boundary(const std::string &path, std::vector<int> &nodes, std::vector<double> &vals) {
    std::ifstream f(path);
    if (!f) return 1;
    int n = 0;
    if (!(f >> n) || n < 0) return 3;
    nodes.resize(n);
    vals.assign(2 * (size_t)n, 0.0);
    for (int i = 0; i < n; ++i) {
        std::string v;
        if (!(f >> nodes[i] >> v)) return 3;
        if (v[0] == '(') {
            std::string rest = v;
            while (rest.find(')') == std::string::npos) {
                std::string more;
                if (!(f >> more)) return 3;
                rest += more;
            }
            for (char &ch : rest)
                if (ch == '(' || ch == ')' || ch == ',') ch = ' ';
            std::istringstream is(rest);
            if (!(is >> vals[2 * i] >> vals[2 * i + 1])) return 3;
        } else {
            try { vals[2 * i] = std::stod(v); } catch (...) { return 3; }
        }
    }
    return 0;
}

// Name of file n of a numbered series (src/wpot.f90:34-44, src/wsens.f90:46-60 with
// filpat.f90 and intcha.f90): the directory length as filpat reports it (at least 1),
// the name up to its first '.', then n in lnanz = int(log10(real(nanz))) + 1 digits,
// zero-filled (asterisks when n has more digits, as Fortran's I format), then the four
// characters from the '.' on.  base_len gets the name's length before the number.
std::string numbered_name(const std::string &datei, int n, int nanz, int &lnanz, int &base_len) {
    lnanz = (int)std::log10((float)nanz) + 1;
    const size_t sl = datei.rfind('/');
    const int id2 = sl == std::string::npos ? 0 : (int)sl + 1;  // 1-based last slash
    const int idum2 = id2 >= 2 ? id2 - 1 : 1;
    const size_t dot = (size_t)idum2 <= datei.size() ? datei.find('.', idum2) : std::string::npos;
    const int idx = dot == std::string::npos ? 0 : (int)dot - idum2 + 1;
    const int idum = idum2 + idx - 1;
    base_len = idum - idum2 - 1;
    std::string digits = std::to_string(n);
    if ((int)digits.size() > lnanz) digits = std::string(lnanz, '*');
    else digits = std::string(lnanz - digits.size(), '0') + digits;
    std::string ext = idum >= 0 && (size_t)idum < datei.size() ? datei.substr(idum, 4) : std::string();
    while (!ext.empty() && ext.back() == ' ') ext.pop_back();
    return datei.substr(0, std::max(idum, 0)) + digits + ext;
}

// one REAL*4 value per column, enough digits to read back exactly
void put_row(FILE *f, std::initializer_list<double> v) {
    for (double x : v) fprintf(f, " %16.8E", (double)(float)x);
    fputc('\n', f);
}

}  // namespace

extern "C" {

int crt_tomodir_sizes(const char *cfg_path, int *sanz, int *mb, int *elanz, int *eanz, int *nanz) {
    if (!cfg_path) return CRT_ERR_ARG;
    CrmodCfg cfg;
    int e = read_crmod_cfg(cfg_path, cfg);
    if (e) return e;
    ElemFile g;
    if ((e = read_elem(cfg.elem, g))) return e;
    std::vector<long long> el, cf;
    int ne = 0, nc = 0;
    if ((e = read_counted_ints(cfg.elec, 1, el, ne))) return e;
    if ((e = read_counted_ints(cfg.config, 2, cf, nc))) return e;
    if (sanz) *sanz = g.sanz;
    if (mb) *mb = g.mb;
    if (elanz) *elanz = g.elanz;
    if (eanz) *eanz = ne;
    if (nanz) *nanz = nc;
    return 0;
}

int crt_write_volt(const char *path, int nanz, const int *strnr, const int *vnr, const double *sigmaa) {
    if (!path || nanz < 0 || (nanz && (!strnr || !vnr || !sigmaa))) return CRT_ERR_ARG;
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    const double pi = std::acos(-1.0);
    fprintf(f, "%12d\n", nanz);
    for (int i = 0; i < nanz; ++i) {
        const double re = sigmaa[2 * i], im = sigmaa[2 * i + 1];
        const double bet = std::hypot(re, im);
        double pha = std::atan2(im, re);
        int v = vnr[i];
        const double npi = std::round(pha / pi) * pi;
        if (std::fabs(npi) > 1e-12) {
            pha -= npi;
            const int e1 = v % 10000, e2 = (v - e1) / 10000;
            v = e1 * 10000 + e2;
        }
        // REAL*4 values, printed with enough digits to read back exactly
        fprintf(f, "%12d %12d %16.8E %16.8E\n", strnr[i], v, (double)(float)bet, (double)(float)(1e3 * pha));
    }
    return fclose(f) == 0 ? 0 : 4;
}

int crt_tomodir_crmod(crt_ctx *ctx, const char *cfg_path) {
    if (!ctx || !cfg_path) return CRT_ERR_ARG;
    CrmodCfg cfg;
    int e = read_crmod_cfg(cfg_path, cfg);
    if (e) return e;
    ElemFile g;
    if ((e = read_elem(cfg.elem, g))) return e;
    std::vector<long long> el, cf;
    int ne = 0, nc = 0, nr = 0;
    if ((e = read_counted_ints(cfg.elec, 1, el, ne))) return e;
    if ((e = read_counted_ints(cfg.config, 2, cf, nc))) return e;
    std::vector<double> sigma;
    if ((e = read_rho(cfg.rho, sigma, nr))) return e;
    if (nr != g.elanz) return CRT_ERR_ARG;
    std::vector<int> enr(el.begin(), el.end()), strnr(nc), vnr(nc);
    for (int i = 0; i < nc; ++i) {
        strnr[i] = (int)cf[2 * (size_t)i];
        vnr[i] = (int)cf[2 * (size_t)i + 1];
    }
    if ((e = crt_set_grid(ctx, g.sanz, g.mb, g.typanz, g.typ.data(), g.nelanz.data(), g.selanz.data(), g.maxn,
                          g.nrel.data(), g.x.data(), g.y.data(), g.rnr.data(), ne, enr.data())))
        return e;
    if ((e = crt_set_survey(ctx, nc, strnr.data(), vnr.data()))) return e;
    // another dataset in the same run (lagain) is a host loop this call does not make
    if (cfg.lagain) {
        fprintf(stderr, "crt_tomodir_crmod: another dataset (lagain) is not supported\n");
        return CRT_ERR_UNSUPPORTED;
    }
    if (cfg.swrtr != 0 && cfg.swrtr != 1) return 3;
    // mswitch bits 0 (lana) and 2 (lsr; lsr switches lana off), both off in 2D
    // (src/fem.f90:191-203, 241-243); the fictitious sink (src/fem.f90:179-180)
    {
        const bool lsr = (cfg.mswitch & 4) && cfg.swrtr != 0, lana = (cfg.mswitch & 1) && cfg.swrtr != 0 && !lsr;
        if ((e = crt_set_fem_switches(ctx, lsr ? 1 : 0, lana ? 1 : 0, cfg.lsink ? cfg.nsink : 0))) return e;
    }
    if (cfg.lrandb2) {
        std::vector<int> bn;
        std::vector<double> bv;
        if ((e = read_boundary(cfg.randb, bn, bv))) return e;
        if ((e = crt_set_boundary_values(ctx, (int)bn.size(), bn.data(), bv.data()))) return e;
    }
    int kwnanz = 0;
    double amin = 0.0, amax = 0.0;
    std::vector<double> kwn(64), kwnwi(64);
    if (cfg.swrtr == 0) {
        // 2D (src/fem.f90:241-250): one wavenumber k = 0; rtrafo and bsens take kpot
        // itself (src/fem.f90:393), which the back-transform gives with the weight pi
        kwnanz = 1;
        kwn[0] = 0.0;
        kwnwi[0] = std::acos(-1.0);
        if ((e = crt_set_wavenumbers(ctx, 1, kwn.data(), kwnwi.data()))) return e;
    } else if ((e = crt_rwaven(ctx, &kwnanz, kwn.data(), kwnwi.data(), &amin, &amax))) return e;
    int lbeta = 0;
    if ((e = crt_precal(ctx, 0, &lbeta))) return e;
    // wkfak with mixed boundaries: bvolt multiplies by bkfak's geometric factors
    // (src/fem.f90:417-423), otherwise kfak = 1
    std::vector<double> kfak;
    if ((cfg.mswitch & 2) && lbeta) {
        kfak.resize(nc);
        if ((e = crt_bkfak(ctx, kfak.data()))) return e;
    }
    if ((e = crt_set_sigma(ctx, sigma.data()))) return e;
    if ((e = crt_forward(ctx))) return e;
    std::vector<double> volt(2 * (size_t)nc), sigmaa(2 * (size_t)nc);
    if ((e = crt_voltages(ctx, 0, kfak.empty() ? nullptr : kfak.data(), volt.data(), sigmaa.data()))) return e;
    if (cfg.wvolt && (e = crt_write_volt(cfg.volt.c_str(), nc, strnr.data(), vnr.data(), sigmaa.data()))) return e;
    if (cfg.wpot) {
        // bpot (src/bpot.f90:25-40): per measurement the potential of its current
        // dipole, (hpot(:,B) - hpot(:,A)) * strom with strom = 1, written by wpot
        // (src/wpot.f90:45-50): node coordinates and the complex potential per node
        std::vector<double> hpot(2 * (size_t)g.sanz * ne);
        if ((e = crt_get_hpot(ctx, hpot.data()))) return e;
        for (int i = 0; i < nc; ++i) {
            const int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
            int lnanz = 0, blen = 0;
            const std::string name = numbered_name(cfg.pot, i + 1, nc, lnanz, blen);
            if (blen > 8 - lnanz) return 15;  // src/wpot.f90:36-40
            FILE *f = fopen(name.c_str(), "w");
            if (!f) return 1;
            for (int j = 0; j < g.sanz; ++j) {
                double re = 0.0, im = 0.0;
                if (e2 > 0) { re += hpot[2 * ((size_t)(e2 - 1) * g.sanz + j)]; im += hpot[2 * ((size_t)(e2 - 1) * g.sanz + j) + 1]; }
                if (e1 > 0) { re -= hpot[2 * ((size_t)(e1 - 1) * g.sanz + j)]; im -= hpot[2 * ((size_t)(e1 - 1) * g.sanz + j) + 1]; }
                put_row(f, {g.x[j], g.y[j], re, im});
            }
            if (fclose(f) != 0) return 4;
        }
    }
    if (cfg.wsens) {
        // bsens (dV/dsigma, strom = 1) and wsens (src/wsens.f90:62-80): per measurement
        // |sum_j J(i,j)| and 1000 atan2 of that sum, then element centroids (the mean of
        // the element's node coordinates, src/relem.f90:128-140) and J per element
        if ((e = crt_sens(ctx, 0, nullptr))) return e;
        std::vector<double> J(2 * (size_t)nc * g.elanz);
        if ((e = crt_get_sens(ctx, J.data()))) return e;
        std::vector<double> cx(g.elanz, 0.0), cy(g.elanz, 0.0);
        for (int t = 0, iel = 0, ia = 0; t < g.typanz; ++t)
            for (int q = 0; q < g.nelanz[t]; ++q, ++iel) {
                if (g.typ[t] > 10) continue;
                for (int k = 0; k < g.selanz[t]; ++k) {
                    const int nd = g.nrel[(size_t)iel * g.maxn + k] - 1;
                    cx[ia] += g.x[nd];
                    cy[ia] += g.y[nd];
                }
                cx[ia] /= g.selanz[t];
                cy[ia] /= g.selanz[t];
                ++ia;
            }
        for (int i = 0; i < nc; ++i) {
            int lnanz = 0, blen = 0;
            const std::string name = numbered_name(cfg.sens, i + 1, nc, lnanz, blen);
            FILE *f = fopen(name.c_str(), "w");
            if (!f) return 1;
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < g.elanz; ++j) {
                sr += J[2 * ((size_t)j * nc + i)];
                si += J[2 * ((size_t)j * nc + i) + 1];
            }
            if (std::hypot(sr, si) > 0.0) put_row(f, {std::hypot(sr, si), 1e3 * std::atan2(si, sr)});
            else put_row(f, {0.0, 0.0});
            for (int j = 0; j < g.elanz; ++j)
                put_row(f, {cx[j], cy[j], J[2 * ((size_t)j * nc + i)], J[2 * ((size_t)j * nc + i) + 1]});
            if (fclose(f) != 0) return 4;
        }
    }
    return 0;
}

}  // extern "C"
