// SGP4/SDP4 propagation of a two-line element set and its rotation to
// Earth-fixed coordinates: the host-side source of a TLE transmitter's
// position grid (TleTrajectory, src/geometry/tle.rs:15-85).
//
// The reference propagates with the third-party `sgp4` crate (2.3, improved
// mode, WGS72). That crate is not in the reference tree; this file restates
// the published algorithm it implements: Hoots and Roehrich, Spacetrack
// Report #3 (1980), as revised by Vallado, Crawford, Hujsak and Kelso,
// "Revisiting Spacetrack Report #3" (AIAA 2006-6753): the near-Earth model
// with drag, and for periods of 225 min or more the deep-space model with
// lunar-solar secular and periodic terms and the 12 h and 24 h geopotential
// resonances integrated from epoch (every call starts the integration from
// epoch, as the crate's immutable propagate does). Improved mode: sidereal
// time from the IAU 1982 formula at the element epoch, no AFSPC angle wraps.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/lpk.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kX2o3 = 2.0 / 3.0;

// WGS72 geopotential (Spacetrack Report #3, table of constants)
constexpr double kMu = 398600.8;          // km^3 / s^2
constexpr double kRe = 6378.135;          // km
constexpr double kJ2 = 0.001082616;
constexpr double kJ3 = -0.00000253881;
constexpr double kJ4 = -0.00000165597;
constexpr double kJ3oJ2 = kJ3 / kJ2;
inline double xke() { return 60.0 / std::sqrt(kRe * kRe * kRe / kMu); }  // sqrt(mu) in earth radii^1.5 / min

// Earth rotation for the TEME -> ECEF velocity term (src/geometry/tle.rs:9)
constexpr double kOmegaE = 7.2921151467e-5;

// The propagator's state after initialisation.
struct Sat {
    // elements (radians, radians per minute)
    double bstar, ecco, argpo, inclo, mo, no_kozai, nodeo, no_unkozai;
    // near-Earth
    int isimp;
    bool deep;
    double a, con41, cc1, cc4, cc5, d2, d3, d4, delmo, eta, argpdot, omgcof, sinmao, t2cof, t3cof, t4cof, t5cof,
        x1mth2, x7thm1, mdot, nodedot, xlcof, xmcof, nodecf, aycof;
    // deep space
    int irez;
    double d2201, d2211, d3210, d3222, d4410, d4422, d5220, d5232, d5421, d5433, dedt, del1, del2, del3, didt,
        dmdt, dnodt, domdt, e3, ee2, peo, pgho, pho, pinco, plo, se2, se3, sgh2, sgh3, sgh4, sh2, sh3, si2, si3,
        sl2, sl3, sl4, gsto, xfact, xgh2, xgh3, xgh4, xh2, xh3, xi2, xi3, xl2, xl3, xl4, xlamo, zmol, zmos;
};

// Greenwich sidereal time (IAU 1982) at Julian date jd_ut1, radians in [0, 2 pi).
double gstime(double jd_ut1) {
    const double tut1 = (jd_ut1 - 2451545.0) / 36525.0;
    double temp = -6.2e-6 * tut1 * tut1 * tut1 + 0.093104 * tut1 * tut1 +
                  (876600.0 * 3600.0 + 8640184.812866) * tut1 + 67310.54841;  // seconds
    temp = std::fmod(temp * (kPi / 180.0) / 240.0, kTwoPi);
    if (temp < 0.0) temp += kTwoPi;
    return temp;
}

// Lunar-solar terms common to the secular and periodic parts (dscom).
struct Dscom {
    double snodm, cnodm, sinim, cosim, sinomm, cosomm, day, em, emsq, gam, rtemsq, nm;
    double s1, s2, s3, s4, s5, s6, s7, ss1, ss2, ss3, ss4, ss5, ss6, ss7;
    double sz1, sz2, sz3, sz11, sz12, sz13, sz21, sz22, sz23, sz31, sz32, sz33;
    double z1, z2, z3, z11, z12, z13, z21, z22, z23, z31, z32, z33;
};

void dscom(double epoch, double ep, double argpp, double tc, double inclp, double nodep, double np, Sat &s,
           Dscom &o) {
    const double zes = 0.01675, zel = 0.05490, c1ss = 2.9864797e-6, c1l = 4.7968065e-7;
    const double zsinis = 0.39785416, zcosis = 0.91744867, zcosgs = 0.1945905, zsings = -0.98088458;
    o.nm = np;
    o.em = ep;
    o.snodm = std::sin(nodep);
    o.cnodm = std::cos(nodep);
    o.sinomm = std::sin(argpp);
    o.cosomm = std::cos(argpp);
    o.sinim = std::sin(inclp);
    o.cosim = std::cos(inclp);
    o.emsq = o.em * o.em;
    const double betasq = 1.0 - o.emsq;
    o.rtemsq = std::sqrt(betasq);

    s.peo = s.pinco = s.plo = s.pgho = s.pho = 0.0;
    o.day = epoch + 18261.5 + tc / 1440.0;
    const double xnodce = std::fmod(4.5236020 - 9.2422029e-4 * o.day, kTwoPi);
    const double stem = std::sin(xnodce), ctem = std::cos(xnodce);
    const double zcosil = 0.91375164 - 0.03568096 * ctem;
    const double zsinil = std::sqrt(1.0 - zcosil * zcosil);
    const double zsinhl = 0.089683511 * stem / zsinil;
    const double zcoshl = std::sqrt(1.0 - zsinhl * zsinhl);
    o.gam = 5.8351514 + 0.0019443680 * o.day;
    double zx = 0.39785416 * stem / zsinil;
    const double zy = zcoshl * ctem + 0.91744867 * zsinhl * stem;
    zx = std::atan2(zx, zy);
    zx = o.gam + zx - xnodce;
    const double zcosgl = std::cos(zx), zsingl = std::sin(zx);

    // the sun first (lsflg 1), then the moon
    double zcosg = zcosgs, zsing = zsings, zcosi = zcosis, zsini = zsinis;
    double zcosh = o.cnodm, zsinh = o.snodm, cc = c1ss;
    const double xnoi = 1.0 / o.nm;
    for (int lsflg = 1; lsflg <= 2; ++lsflg) {
        const double a1 = zcosg * zcosh + zsing * zcosi * zsinh;
        const double a3 = -zsing * zcosh + zcosg * zcosi * zsinh;
        const double a7 = -zcosg * zsinh + zsing * zcosi * zcosh;
        const double a8 = zsing * zsini;
        const double a9 = zsing * zsinh + zcosg * zcosi * zcosh;
        const double a10 = zcosg * zsini;
        const double a2 = o.cosim * a7 + o.sinim * a8;
        const double a4 = o.cosim * a9 + o.sinim * a10;
        const double a5 = -o.sinim * a7 + o.cosim * a8;
        const double a6 = -o.sinim * a9 + o.cosim * a10;

        const double x1 = a1 * o.cosomm + a2 * o.sinomm;
        const double x2 = a3 * o.cosomm + a4 * o.sinomm;
        const double x3 = -a1 * o.sinomm + a2 * o.cosomm;
        const double x4 = -a3 * o.sinomm + a4 * o.cosomm;
        const double x5 = a5 * o.sinomm;
        const double x6 = a6 * o.sinomm;
        const double x7 = a5 * o.cosomm;
        const double x8 = a6 * o.cosomm;

        o.z31 = 12.0 * x1 * x1 - 3.0 * x3 * x3;
        o.z32 = 24.0 * x1 * x2 - 6.0 * x3 * x4;
        o.z33 = 12.0 * x2 * x2 - 3.0 * x4 * x4;
        o.z1 = 3.0 * (a1 * a1 + a2 * a2) + o.z31 * o.emsq;
        o.z2 = 6.0 * (a1 * a3 + a2 * a4) + o.z32 * o.emsq;
        o.z3 = 3.0 * (a3 * a3 + a4 * a4) + o.z33 * o.emsq;
        o.z11 = -6.0 * a1 * a5 + o.emsq * (-24.0 * x1 * x7 - 6.0 * x3 * x5);
        o.z12 = -6.0 * (a1 * a6 + a3 * a5) + o.emsq * (-24.0 * (x2 * x7 + x1 * x8) - 6.0 * (x3 * x6 + x4 * x5));
        o.z13 = -6.0 * a3 * a6 + o.emsq * (-24.0 * x2 * x8 - 6.0 * x4 * x6);
        o.z21 = 6.0 * a2 * a5 + o.emsq * (24.0 * x1 * x5 - 6.0 * x3 * x7);
        o.z22 = 6.0 * (a4 * a5 + a2 * a6) + o.emsq * (24.0 * (x2 * x5 + x1 * x6) - 6.0 * (x4 * x7 + x3 * x8));
        o.z23 = 6.0 * a4 * a6 + o.emsq * (24.0 * x2 * x6 - 6.0 * x4 * x8);
        o.z1 = o.z1 + o.z1 + betasq * o.z31;
        o.z2 = o.z2 + o.z2 + betasq * o.z32;
        o.z3 = o.z3 + o.z3 + betasq * o.z33;
        o.s3 = cc * xnoi;
        o.s2 = -0.5 * o.s3 / o.rtemsq;
        o.s4 = o.s3 * o.rtemsq;
        o.s1 = -15.0 * o.em * o.s4;
        o.s5 = x1 * x3 + x2 * x4;
        o.s6 = x2 * x3 + x1 * x4;
        o.s7 = x2 * x4 - x1 * x3;

        if (lsflg == 1) {
            o.ss1 = o.s1; o.ss2 = o.s2; o.ss3 = o.s3; o.ss4 = o.s4; o.ss5 = o.s5; o.ss6 = o.s6; o.ss7 = o.s7;
            o.sz1 = o.z1; o.sz2 = o.z2; o.sz3 = o.z3;
            o.sz11 = o.z11; o.sz12 = o.z12; o.sz13 = o.z13;
            o.sz21 = o.z21; o.sz22 = o.z22; o.sz23 = o.z23;
            o.sz31 = o.z31; o.sz32 = o.z32; o.sz33 = o.z33;
            zcosg = zcosgl;
            zsing = zsingl;
            zcosi = zcosil;
            zsini = zsinil;
            zcosh = zcoshl * o.cnodm + zsinhl * o.snodm;
            zsinh = o.snodm * zcoshl - o.cnodm * zsinhl;
            cc = c1l;
        }
    }
    s.zmol = std::fmod(4.7199672 + 0.22997150 * o.day - o.gam, kTwoPi);
    s.zmos = std::fmod(6.2565837 + 0.017201977 * o.day, kTwoPi);

    // solar periodic coefficients
    s.se2 = 2.0 * o.ss1 * o.ss6;
    s.se3 = 2.0 * o.ss1 * o.ss7;
    s.si2 = 2.0 * o.ss2 * o.sz12;
    s.si3 = 2.0 * o.ss2 * (o.sz13 - o.sz11);
    s.sl2 = -2.0 * o.ss3 * o.sz2;
    s.sl3 = -2.0 * o.ss3 * (o.sz3 - o.sz1);
    s.sl4 = -2.0 * o.ss3 * (-21.0 - 9.0 * o.emsq) * zes;
    s.sgh2 = 2.0 * o.ss4 * o.sz32;
    s.sgh3 = 2.0 * o.ss4 * (o.sz33 - o.sz31);
    s.sgh4 = -18.0 * o.ss4 * zes;
    s.sh2 = -2.0 * o.ss2 * o.sz22;
    s.sh3 = -2.0 * o.ss2 * (o.sz23 - o.sz21);
    // lunar periodic coefficients
    s.ee2 = 2.0 * o.s1 * o.s6;
    s.e3 = 2.0 * o.s1 * o.s7;
    s.xi2 = 2.0 * o.s2 * o.z12;
    s.xi3 = 2.0 * o.s2 * (o.z13 - o.z11);
    s.xl2 = -2.0 * o.s3 * o.z2;
    s.xl3 = -2.0 * o.s3 * (o.z3 - o.z1);
    s.xl4 = -2.0 * o.s3 * (-21.0 - 9.0 * o.emsq) * zel;
    s.xgh2 = 2.0 * o.s4 * o.z32;
    s.xgh3 = 2.0 * o.s4 * (o.z33 - o.z31);
    s.xgh4 = -18.0 * o.s4 * zel;
    s.xh2 = -2.0 * o.s2 * o.z22;
    s.xh3 = -2.0 * o.s2 * (o.z23 - o.z21);
}

// Lunar-solar periodics (dpper). `init`: the call made while initialising,
// which evaluates the terms and applies nothing. Later calls subtract the
// epoch values peo..pho, which dscom leaves at zero, so the full periodics
// apply.
void dpper(const Sat &s, double t, bool init, double &ep, double &inclp, double &nodep, double &argpp, double &mp) {
    const double zns = 1.19459e-5, zes = 0.01675, znl = 1.5835218e-4, zel = 0.05490;
    double zm = init ? s.zmos : s.zmos + zns * t;
    double zf = zm + 2.0 * zes * std::sin(zm);
    double sinzf = std::sin(zf);
    double f2 = 0.5 * sinzf * sinzf - 0.25;
    double f3 = -0.5 * sinzf * std::cos(zf);
    const double ses = s.se2 * f2 + s.se3 * f3;
    const double sis = s.si2 * f2 + s.si3 * f3;
    const double sls = s.sl2 * f2 + s.sl3 * f3 + s.sl4 * sinzf;
    const double sghs = s.sgh2 * f2 + s.sgh3 * f3 + s.sgh4 * sinzf;
    const double shs = s.sh2 * f2 + s.sh3 * f3;
    zm = init ? s.zmol : s.zmol + znl * t;
    zf = zm + 2.0 * zel * std::sin(zm);
    sinzf = std::sin(zf);
    f2 = 0.5 * sinzf * sinzf - 0.25;
    f3 = -0.5 * sinzf * std::cos(zf);
    const double sel = s.ee2 * f2 + s.e3 * f3;
    const double sil = s.xi2 * f2 + s.xi3 * f3;
    const double sll = s.xl2 * f2 + s.xl3 * f3 + s.xl4 * sinzf;
    const double sghl = s.xgh2 * f2 + s.xgh3 * f3 + s.xgh4 * sinzf;
    const double shll = s.xh2 * f2 + s.xh3 * f3;
    double pe = ses + sel, pinc = sis + sil, pl = sls + sll, pgh = sghs + sghl, ph = shs + shll;
    if (init) return;

    pe -= s.peo;
    pinc -= s.pinco;
    pl -= s.plo;
    pgh -= s.pgho;
    ph -= s.pho;
    inclp += pinc;
    ep += pe;
    const double sinip = std::sin(inclp), cosip = std::cos(inclp);
    if (inclp >= 0.2) {
        // periodics applied directly
        ph /= sinip;
        pgh -= cosip * ph;
        argpp += pgh;
        nodep += ph;
        mp += pl;
    } else {
        // Lyddane's modification near zero inclination
        const double sinop = std::sin(nodep), cosop = std::cos(nodep);
        double alfdp = sinip * sinop, betdp = sinip * cosop;
        const double dalf = ph * cosop + pinc * cosip * sinop;
        const double dbet = -ph * sinop + pinc * cosip * cosop;
        alfdp += dalf;
        betdp += dbet;
        nodep = std::fmod(nodep, kTwoPi);
        double xls = mp + argpp + cosip * nodep;
        const double dls = pl + pgh - pinc * nodep * sinip;
        xls += dls;
        const double xnoh = nodep;
        nodep = std::atan2(alfdp, betdp);
        if (std::fabs(xnoh - nodep) > kPi) nodep = nodep < xnoh ? nodep + kTwoPi : nodep - kTwoPi;
        mp += pl;
        argpp = xls - mp - cosip * nodep;
    }
}

constexpr double kRptim = 4.37526908801129966e-3;  // Earth rotation, rad / min

// Deep-space secular rates and resonance set-up (dsinit).
void dsinit(Sat &s, const Dscom &c, double t, double tc, double xpidot, double eccsq, double &em, double &argpm,
            double &inclm, double &mm, double &nm, double &nodem) {
    const double q22 = 1.7891679e-6, q31 = 2.1460748e-6, q33 = 2.2123015e-7;
    const double root22 = 1.7891679e-6, root44 = 7.3636953e-9, root54 = 2.1765803e-9;
    const double root32 = 3.7393792e-7, root52 = 1.1428639e-7;
    const double znl = 1.5835218e-4, zns = 1.19459e-5;
    const double cosim = c.cosim, sinim = c.sinim;
    double emsq = c.emsq;

    s.irez = 0;
    if (nm < 0.0052359877 && nm > 0.0034906585) s.irez = 1;
    if (nm >= 8.26e-3 && nm <= 9.24e-3 && em >= 0.5) s.irez = 2;

    // solar terms
    const double ses = c.ss1 * zns * c.ss5;
    const double sis = c.ss2 * zns * (c.sz11 + c.sz13);
    const double sls = -zns * c.ss3 * (c.sz1 + c.sz3 - 14.0 - 6.0 * emsq);
    const double sghs = c.ss4 * zns * (c.sz31 + c.sz33 - 6.0);
    double shs = -zns * c.ss2 * (c.sz21 + c.sz23);
    if (inclm < 5.2359877e-2 || inclm > kPi - 5.2359877e-2) shs = 0.0;
    if (sinim != 0.0) shs = shs / sinim;
    const double sgs = sghs - cosim * shs;

    // lunar terms
    s.dedt = ses + c.s1 * znl * c.s5;
    s.didt = sis + c.s2 * znl * (c.z11 + c.z13);
    s.dmdt = sls - znl * c.s3 * (c.z1 + c.z3 - 14.0 - 6.0 * emsq);
    const double sghl = c.s4 * znl * (c.z31 + c.z33 - 6.0);
    double shll = -znl * c.s2 * (c.z21 + c.z23);
    if (inclm < 5.2359877e-2 || inclm > kPi - 5.2359877e-2) shll = 0.0;
    s.domdt = sgs + sghl;
    s.dnodt = shs;
    if (sinim != 0.0) {
        s.domdt = s.domdt - cosim / sinim * shll;
        s.dnodt = s.dnodt + shll / sinim;
    }

    // secular updates over t (zero at initialisation)
    const double theta = std::fmod(s.gsto + tc * kRptim, kTwoPi);
    em = em + s.dedt * t;
    inclm = inclm + s.didt * t;
    argpm = argpm + s.domdt * t;
    nodem = nodem + s.dnodt * t;
    mm = mm + s.dmdt * t;

    if (s.irez == 0) return;
    const double aonv = std::pow(nm / xke(), kX2o3);
    if (s.irez == 2) {
        // geopotential resonance of 12 h orbits
        const double cosisq = cosim * cosim;
        const double emo = em;
        em = s.ecco;
        const double emsqo = emsq;
        emsq = eccsq;
        const double eoc = em * emsq;
        const double g201 = -0.306 - (em - 0.64) * 0.440;
        double g211, g310, g322, g410, g422, g520, g521, g532, g533;
        if (em <= 0.65) {
            g211 = 3.616 - 13.2470 * em + 16.2900 * emsq;
            g310 = -19.302 + 117.3900 * em - 228.4190 * emsq + 156.5910 * eoc;
            g322 = -18.9068 + 109.7927 * em - 214.6334 * emsq + 146.5816 * eoc;
            g410 = -41.122 + 242.6940 * em - 471.0940 * emsq + 313.9530 * eoc;
            g422 = -146.407 + 841.8800 * em - 1629.014 * emsq + 1083.4350 * eoc;
            g520 = -532.114 + 3017.977 * em - 5740.032 * emsq + 3708.2760 * eoc;
        } else {
            g211 = -72.099 + 331.819 * em - 508.738 * emsq + 266.724 * eoc;
            g310 = -346.844 + 1582.851 * em - 2415.925 * emsq + 1246.113 * eoc;
            g322 = -342.585 + 1554.908 * em - 2366.899 * emsq + 1215.972 * eoc;
            g410 = -1052.797 + 4758.686 * em - 7193.992 * emsq + 3651.957 * eoc;
            g422 = -3581.690 + 16178.110 * em - 24462.770 * emsq + 12422.520 * eoc;
            g520 = em > 0.715 ? -5149.66 + 29936.92 * em - 54087.36 * emsq + 31324.56 * eoc
                              : 1464.74 - 4664.75 * em + 3763.64 * emsq;
        }
        if (em < 0.7) {
            g533 = -919.22770 + 4988.6100 * em - 9064.7700 * emsq + 5542.21 * eoc;
            g521 = -822.71072 + 4568.6173 * em - 8491.4146 * emsq + 5337.524 * eoc;
            g532 = -853.66600 + 4690.2500 * em - 8624.7700 * emsq + 5341.4 * eoc;
        } else {
            g533 = -37995.780 + 161616.52 * em - 229838.20 * emsq + 109377.94 * eoc;
            g521 = -51752.104 + 218913.95 * em - 309468.16 * emsq + 146349.42 * eoc;
            g532 = -40023.880 + 170470.89 * em - 242699.48 * emsq + 115605.82 * eoc;
        }
        const double sini2 = sinim * sinim;
        const double f220 = 0.75 * (1.0 + 2.0 * cosim + cosisq);
        const double f221 = 1.5 * sini2;
        const double f321 = 1.875 * sinim * (1.0 - 2.0 * cosim - 3.0 * cosisq);
        const double f322 = -1.875 * sinim * (1.0 + 2.0 * cosim - 3.0 * cosisq);
        const double f441 = 35.0 * sini2 * f220;
        const double f442 = 39.3750 * sini2 * sini2;
        const double f522 = 9.84375 * sinim *
                            (sini2 * (1.0 - 2.0 * cosim - 5.0 * cosisq) +
                             0.33333333 * (-2.0 + 4.0 * cosim + 6.0 * cosisq));
        const double f523 = sinim * (4.92187512 * sini2 * (-2.0 - 4.0 * cosim + 10.0 * cosisq) +
                                     6.56250012 * (1.0 + 2.0 * cosim - 3.0 * cosisq));
        const double f542 = 29.53125 * sinim * (2.0 - 8.0 * cosim + cosisq * (-12.0 + 8.0 * cosim + 10.0 * cosisq));
        const double f543 = 29.53125 * sinim * (-2.0 - 8.0 * cosim + cosisq * (12.0 + 8.0 * cosim - 10.0 * cosisq));
        const double xno2 = nm * nm;
        const double ainv2 = aonv * aonv;
        double temp1 = 3.0 * xno2 * ainv2;
        double temp = temp1 * root22;
        s.d2201 = temp * f220 * g201;
        s.d2211 = temp * f221 * g211;
        temp1 = temp1 * aonv;
        temp = temp1 * root32;
        s.d3210 = temp * f321 * g310;
        s.d3222 = temp * f322 * g322;
        temp1 = temp1 * aonv;
        temp = 2.0 * temp1 * root44;
        s.d4410 = temp * f441 * g410;
        s.d4422 = temp * f442 * g422;
        temp1 = temp1 * aonv;
        temp = temp1 * root52;
        s.d5220 = temp * f522 * g520;
        s.d5232 = temp * f523 * g532;
        temp = 2.0 * temp1 * root54;
        s.d5421 = temp * f542 * g521;
        s.d5433 = temp * f543 * g533;
        s.xlamo = std::fmod(s.mo + s.nodeo + s.nodeo - theta - theta, kTwoPi);
        s.xfact = s.mdot + s.dmdt + 2.0 * (s.nodedot + s.dnodt - kRptim) - s.no_unkozai;
        em = emo;
        emsq = emsqo;
    }
    if (s.irez == 1) {
        // synchronous (24 h) resonance
        const double g200 = 1.0 + emsq * (-2.5 + 0.8125 * emsq);
        const double g310 = 1.0 + 2.0 * emsq;
        const double g300 = 1.0 + emsq * (-6.0 + 6.60937 * emsq);
        const double f220 = 0.75 * (1.0 + cosim) * (1.0 + cosim);
        const double f311 = 0.9375 * sinim * sinim * (1.0 + 3.0 * cosim) - 0.75 * (1.0 + cosim);
        double f330 = 1.0 + cosim;
        f330 = 1.875 * f330 * f330 * f330;
        s.del1 = 3.0 * nm * nm * aonv * aonv;
        s.del2 = 2.0 * s.del1 * f220 * g200 * q22;
        s.del3 = 3.0 * s.del1 * f330 * g300 * q33 * aonv;
        s.del1 = s.del1 * f311 * g310 * q31 * aonv;
        s.xlamo = std::fmod(s.mo + s.nodeo + s.argpo - theta, kTwoPi);
        s.xfact = s.mdot + xpidot - kRptim + s.dmdt + s.domdt + s.dnodt - s.no_unkozai;
    }
    nm = s.no_unkozai;  // + dndt, which is zero here
}

// Deep-space secular terms and the resonance integration from epoch (dspace).
void dspace(const Sat &s, double t, double tc, double &em, double &argpm, double &inclm, double &mm, double &nodem,
            double &nm) {
    const double fasx2 = 0.13130908, fasx4 = 2.8843198, fasx6 = 0.37448087;
    const double g22 = 5.7686396, g32 = 0.95240898, g44 = 1.8014998, g52 = 1.0508330, g54 = 4.4108898;
    const double stepp = 720.0, stepn = -720.0, step2 = 259200.0;

    const double theta = std::fmod(s.gsto + tc * kRptim, kTwoPi);
    em += s.dedt * t;
    inclm += s.didt * t;
    argpm += s.domdt * t;
    nodem += s.dnodt * t;
    mm += s.dmdt * t;
    if (s.irez == 0) return;

    // Euler-Maclaurin integration of the resonance terms from epoch
    double atime = 0.0, xni = s.no_unkozai, xli = s.xlamo;
    const double delt = t > 0.0 ? stepp : stepn;
    double ft = 0.0, xndt = 0.0, xldot = 0.0, xnddt = 0.0;
    for (;;) {
        if (s.irez != 2) {
            xndt = s.del1 * std::sin(xli - fasx2) + s.del2 * std::sin(2.0 * (xli - fasx4)) +
                   s.del3 * std::sin(3.0 * (xli - fasx6));
            xldot = xni + s.xfact;
            xnddt = s.del1 * std::cos(xli - fasx2) + 2.0 * s.del2 * std::cos(2.0 * (xli - fasx4)) +
                    3.0 * s.del3 * std::cos(3.0 * (xli - fasx6));
            xnddt = xnddt * xldot;
        } else {
            const double xomi = s.argpo + s.argpdot * atime;
            const double x2omi = xomi + xomi, x2li = xli + xli;
            xndt = s.d2201 * std::sin(x2omi + xli - g22) + s.d2211 * std::sin(xli - g22) +
                   s.d3210 * std::sin(xomi + xli - g32) + s.d3222 * std::sin(-xomi + xli - g32) +
                   s.d4410 * std::sin(x2omi + x2li - g44) + s.d4422 * std::sin(x2li - g44) +
                   s.d5220 * std::sin(xomi + xli - g52) + s.d5232 * std::sin(-xomi + xli - g52) +
                   s.d5421 * std::sin(xomi + x2li - g54) + s.d5433 * std::sin(-xomi + x2li - g54);
            xldot = xni + s.xfact;
            xnddt = s.d2201 * std::cos(x2omi + xli - g22) + s.d2211 * std::cos(xli - g22) +
                    s.d3210 * std::cos(xomi + xli - g32) + s.d3222 * std::cos(-xomi + xli - g32) +
                    s.d5220 * std::cos(xomi + xli - g52) + s.d5232 * std::cos(-xomi + xli - g52) +
                    2.0 * (s.d4410 * std::cos(x2omi + x2li - g44) + s.d4422 * std::cos(x2li - g44) +
                           s.d5421 * std::cos(xomi + x2li - g54) + s.d5433 * std::cos(-xomi + x2li - g54));
            xnddt = xnddt * xldot;
        }
        if (std::fabs(t - atime) < stepp) {
            ft = t - atime;
            break;
        }
        xli = xli + xldot * delt + xndt * step2;
        xni = xni + xndt * delt + xnddt * step2;
        atime = atime + delt;
    }
    nm = xni + xndt * ft + xnddt * ft * ft * 0.5;
    const double xl = xli + xldot * ft + xndt * ft * ft * 0.5;
    if (s.irez != 1)
        mm = xl - 2.0 * nodem + 2.0 * theta;
    else
        mm = xl - nodem - argpm + theta;
}

// Initialisation (sgp4init with initl): epoch in days since 1949 December 31 0h UT.
bool init(Sat &s, double epoch, double jd_epoch, double bstar, double ecco, double argpo, double inclo, double mo,
          double no_kozai, double nodeo) {
    std::memset(&s, 0, sizeof(s));
    s.bstar = bstar;
    s.ecco = ecco;
    s.argpo = argpo;
    s.inclo = inclo;
    s.mo = mo;
    s.no_kozai = no_kozai;
    s.nodeo = nodeo;
    const double temp4 = 1.5e-12;
    const double ss = 78.0 / kRe + 1.0;
    const double qzms2ttemp = (120.0 - 78.0) / kRe;
    const double qzms2t = qzms2ttemp * qzms2ttemp * qzms2ttemp * qzms2ttemp;

    // initl: un-Kozai the mean motion and the epoch quantities
    const double eccsq = ecco * ecco;
    const double omeosq = 1.0 - eccsq;
    const double rteosq = std::sqrt(omeosq);
    const double cosio = std::cos(inclo);
    const double cosio2 = cosio * cosio;
    const double ak = std::pow(xke() / no_kozai, kX2o3);
    const double d1 = 0.75 * kJ2 * (3.0 * cosio2 - 1.0) / (rteosq * omeosq);
    double del = d1 / (ak * ak);
    const double adel = ak * (1.0 - del * del - del * (1.0 / 3.0 + 134.0 * del * del / 81.0));
    del = d1 / (adel * adel);
    s.no_unkozai = no_kozai / (1.0 + del);
    const double ao = std::pow(xke() / s.no_unkozai, kX2o3);
    const double sinio = std::sin(inclo);
    const double po = ao * omeosq;
    const double con42 = 1.0 - 5.0 * cosio2;
    s.con41 = -con42 - cosio2 - cosio2;
    const double posq = po * po;
    const double rp = ao * (1.0 - ecco);
    s.gsto = gstime(jd_epoch);

    s.a = std::pow(s.no_unkozai / xke(), -kX2o3);
    if (!(omeosq >= 0.0 || s.no_unkozai >= 0.0)) return false;
    s.isimp = rp < (220.0 / kRe + 1.0) ? 1 : 0;
    double sfour = ss, qzms24 = qzms2t;
    const double perige = (rp - 1.0) * kRe;
    if (perige < 156.0) {
        // perigees below 156 km alter s and qoms2t
        sfour = perige - 78.0;
        if (perige < 98.0) sfour = 20.0;
        const double qzms24temp = (120.0 - sfour) / kRe;
        qzms24 = qzms24temp * qzms24temp * qzms24temp * qzms24temp;
        sfour = sfour / kRe + 1.0;
    }
    const double pinvsq = 1.0 / posq;
    const double tsi = 1.0 / (ao - sfour);
    s.eta = ao * ecco * tsi;
    const double etasq = s.eta * s.eta;
    const double eeta = ecco * s.eta;
    const double psisq = std::fabs(1.0 - etasq);
    const double coef = qzms24 * std::pow(tsi, 4.0);
    const double coef1 = coef / std::pow(psisq, 3.5);
    const double cc2 = coef1 * s.no_unkozai *
                       (ao * (1.0 + 1.5 * etasq + eeta * (4.0 + etasq)) +
                        0.375 * kJ2 * tsi / psisq * s.con41 * (8.0 + 3.0 * etasq * (8.0 + etasq)));
    s.cc1 = bstar * cc2;
    double cc3 = 0.0;
    if (ecco > 1.0e-4) cc3 = -2.0 * coef * tsi * kJ3oJ2 * s.no_unkozai * sinio / ecco;
    s.x1mth2 = 1.0 - cosio2;
    s.cc4 = 2.0 * s.no_unkozai * coef1 * ao * omeosq *
            (s.eta * (2.0 + 0.5 * etasq) + ecco * (0.5 + 2.0 * etasq) -
             kJ2 * tsi / (ao * psisq) *
                 (-3.0 * s.con41 * (1.0 - 2.0 * eeta + etasq * (1.5 - 0.5 * eeta)) +
                  0.75 * s.x1mth2 * (2.0 * etasq - eeta * (1.0 + etasq)) * std::cos(2.0 * argpo)));
    s.cc5 = 2.0 * coef1 * ao * omeosq * (1.0 + 2.75 * (etasq + eeta) + eeta * etasq);
    const double cosio4 = cosio2 * cosio2;
    const double temp1 = 1.5 * kJ2 * pinvsq * s.no_unkozai;
    const double temp2 = 0.5 * temp1 * kJ2 * pinvsq;
    const double temp3 = -0.46875 * kJ4 * pinvsq * pinvsq * s.no_unkozai;
    s.mdot = s.no_unkozai + 0.5 * temp1 * rteosq * s.con41 +
             0.0625 * temp2 * rteosq * (13.0 - 78.0 * cosio2 + 137.0 * cosio4);
    s.argpdot = -0.5 * temp1 * con42 + 0.0625 * temp2 * (7.0 - 114.0 * cosio2 + 395.0 * cosio4) +
                temp3 * (3.0 - 36.0 * cosio2 + 49.0 * cosio4);
    const double xhdot1 = -temp1 * cosio;
    s.nodedot = xhdot1 + (0.5 * temp2 * (4.0 - 19.0 * cosio2) + 2.0 * temp3 * (3.0 - 7.0 * cosio2)) * cosio;
    const double xpidot = s.argpdot + s.nodedot;
    s.omgcof = bstar * cc3 * std::cos(argpo);
    s.xmcof = 0.0;
    if (ecco > 1.0e-4) s.xmcof = -kX2o3 * coef * bstar / eeta;
    s.nodecf = 3.5 * omeosq * xhdot1 * s.cc1;
    s.t2cof = 1.5 * s.cc1;
    if (std::fabs(cosio + 1.0) > 1.5e-12)
        s.xlcof = -0.25 * kJ3oJ2 * sinio * (3.0 + 5.0 * cosio) / (1.0 + cosio);
    else
        s.xlcof = -0.25 * kJ3oJ2 * sinio * (3.0 + 5.0 * cosio) / temp4;
    s.aycof = -0.5 * kJ3oJ2 * sinio;
    const double delmotemp = 1.0 + s.eta * std::cos(mo);
    s.delmo = delmotemp * delmotemp * delmotemp;
    s.sinmao = std::sin(mo);
    s.x7thm1 = 7.0 * cosio2 - 1.0;

    if (kTwoPi / s.no_unkozai >= 225.0) {
        // deep space
        s.deep = true;
        s.isimp = 1;
        const double tc = 0.0;
        double inclm = inclo;
        Dscom c{};
        dscom(epoch, ecco, argpo, tc, inclo, nodeo, s.no_unkozai, s, c);
        double ep = ecco, ip = inclo, np = nodeo, ap = argpo, mp = mo;
        // the initialising call evaluates and applies nothing; the epoch
        // values peo..pho stay zero (set so by dscom), so later calls apply
        // the full periodics
        dpper(s, 0.0, true, ep, ip, np, ap, mp);
        double em = c.em, argpm = 0.0, mm = 0.0, nm = c.nm, nodem = 0.0;
        dsinit(s, c, 0.0, tc, xpidot, eccsq, em, argpm, inclm, mm, nm, nodem);
    } else {
        s.deep = false;
    }
    if (s.isimp != 1) {
        const double cc1sq = s.cc1 * s.cc1;
        s.d2 = 4.0 * ao * tsi * cc1sq;
        const double temp = s.d2 * tsi * s.cc1 / 3.0;
        s.d3 = (17.0 * ao + sfour) * temp;
        s.d4 = 0.5 * temp * ao * tsi * (221.0 * ao + 31.0 * sfour) * s.cc1;
        s.t3cof = s.d2 + 2.0 * cc1sq;
        s.t4cof = 0.25 * (3.0 * s.d3 + s.cc1 * (12.0 * s.d2 + 10.0 * cc1sq));
        s.t5cof = 0.2 * (3.0 * s.d4 + 12.0 * s.cc1 * s.d3 + 6.0 * s.d2 * s.d2 + 15.0 * cc1sq * (2.0 * s.d2 + cc1sq));
    }
    return true;
}

// Position (km) and velocity (km/s) in TEME at t minutes from epoch; 0 or
// the SGP4 error number (1 eccentricity, 2 mean motion, 3 perturbed
// eccentricity, 4 semi-latus rectum, 6 decay).
int propagate(const Sat &s, double t, double r[3], double v[3]) {
    const double temp4 = 1.5e-12;
    const double vkmpersec = kRe * xke() / 60.0;
    const double xmdf = s.mo + s.mdot * t;
    const double argpdf = s.argpo + s.argpdot * t;
    const double nodedf = s.nodeo + s.nodedot * t;
    double argpm = argpdf, mm = xmdf;
    const double t2 = t * t;
    double nodem = nodedf + s.nodecf * t2;
    double tempa = 1.0 - s.cc1 * t;
    double tempe = s.bstar * s.cc4 * t;
    double templ = s.t2cof * t2;
    if (s.isimp != 1) {
        const double delomg = s.omgcof * t;
        const double delmtemp = 1.0 + s.eta * std::cos(xmdf);
        const double delm = s.xmcof * (delmtemp * delmtemp * delmtemp - s.delmo);
        const double temp = delomg + delm;
        mm = xmdf + temp;
        argpm = argpdf - temp;
        const double t3 = t2 * t, t4 = t3 * t;
        tempa = tempa - s.d2 * t2 - s.d3 * t3 - s.d4 * t4;
        tempe = tempe + s.bstar * s.cc5 * (std::sin(mm) - s.sinmao);
        templ = templ + s.t3cof * t3 + t4 * (s.t4cof + t * s.t5cof);
    }
    double nm = s.no_unkozai, em = s.ecco, inclm = s.inclo;
    if (s.deep) dspace(s, t, t, em, argpm, inclm, mm, nodem, nm);
    if (nm <= 0.0) return 2;
    const double am = std::pow(xke() / nm, kX2o3) * tempa * tempa;
    nm = xke() / std::pow(am, 1.5);
    em = em - tempe;
    if (em >= 1.0 || em < -0.001) return 1;
    if (em < 1.0e-6) em = 1.0e-6;
    mm = mm + s.no_unkozai * templ;
    double xlm = mm + argpm + nodem;
    nodem = std::fmod(nodem, kTwoPi);
    argpm = std::fmod(argpm, kTwoPi);
    xlm = std::fmod(xlm, kTwoPi);
    mm = std::fmod(xlm - argpm - nodem, kTwoPi);

    const double sinim = std::sin(inclm), cosim = std::cos(inclm);
    double ep = em, xincp = inclm, argpp = argpm, nodep = nodem, mp = mm, sinip = sinim, cosip = cosim;
    double aycof = s.aycof, xlcof = s.xlcof, con41 = s.con41, x1mth2 = s.x1mth2, x7thm1 = s.x7thm1;
    if (s.deep) {
        dpper(s, t, false, ep, xincp, nodep, argpp, mp);
        if (xincp < 0.0) {
            xincp = -xincp;
            nodep = nodep + kPi;
            argpp = argpp - kPi;
        }
        if (ep < 0.0 || ep > 1.0) return 3;
        sinip = std::sin(xincp);
        cosip = std::cos(xincp);
        aycof = -0.5 * kJ3oJ2 * sinip;
        if (std::fabs(cosip + 1.0) > 1.5e-12)
            xlcof = -0.25 * kJ3oJ2 * sinip * (3.0 + 5.0 * cosip) / (1.0 + cosip);
        else
            xlcof = -0.25 * kJ3oJ2 * sinip * (3.0 + 5.0 * cosip) / temp4;
    }
    // long-period periodics
    const double axnl = ep * std::cos(argpp);
    double temp = 1.0 / (am * (1.0 - ep * ep));
    const double aynl = ep * std::sin(argpp) + temp * aycof;
    const double xl = mp + argpp + nodep + temp * xlcof * axnl;

    // Kepler's equation
    const double u = std::fmod(xl - nodep, kTwoPi);
    double eo1 = u, tem5 = 9999.9, sineo1 = 0.0, coseo1 = 0.0;
    for (int ktr = 1; std::fabs(tem5) >= 1.0e-12 && ktr <= 10; ++ktr) {
        sineo1 = std::sin(eo1);
        coseo1 = std::cos(eo1);
        tem5 = 1.0 - coseo1 * axnl - sineo1 * aynl;
        tem5 = (u - aynl * coseo1 + axnl * sineo1 - eo1) / tem5;
        if (std::fabs(tem5) >= 0.95) tem5 = tem5 > 0.0 ? 0.95 : -0.95;
        eo1 = eo1 + tem5;
    }

    // short-period preliminaries
    const double ecose = axnl * coseo1 + aynl * sineo1;
    const double esine = axnl * sineo1 - aynl * coseo1;
    const double el2 = axnl * axnl + aynl * aynl;
    const double pl = am * (1.0 - el2);
    if (pl < 0.0) return 4;
    const double rl = am * (1.0 - ecose);
    const double rdotl = std::sqrt(am) * esine / rl;
    const double rvdotl = std::sqrt(pl) / rl;
    const double betal = std::sqrt(1.0 - el2);
    temp = esine / (1.0 + betal);
    const double sinu = am / rl * (sineo1 - aynl - axnl * temp);
    const double cosu = am / rl * (coseo1 - axnl + aynl * temp);
    double su = std::atan2(sinu, cosu);
    const double sin2u = (cosu + cosu) * sinu;
    const double cos2u = 1.0 - 2.0 * sinu * sinu;
    temp = 1.0 / pl;
    const double temp1 = 0.5 * kJ2 * temp;
    const double temp2 = temp1 * temp;
    if (s.deep) {
        const double cosisq = cosip * cosip;
        con41 = 3.0 * cosisq - 1.0;
        x1mth2 = 1.0 - cosisq;
        x7thm1 = 7.0 * cosisq - 1.0;
    }
    const double mrt = rl * (1.0 - 1.5 * temp2 * betal * con41) + 0.5 * temp1 * x1mth2 * cos2u;
    su = su - 0.25 * temp2 * x7thm1 * sin2u;
    const double xnode = nodep + 1.5 * temp2 * cosip * sin2u;
    const double xinc = xincp + 1.5 * temp2 * cosip * sinip * cos2u;
    const double mvt = rdotl - nm * temp1 * x1mth2 * sin2u / xke();
    const double rvdot = rvdotl + nm * temp1 * (x1mth2 * cos2u + 1.5 * con41) / xke();

    // orientation vectors
    const double sinsu = std::sin(su), cossu = std::cos(su);
    const double snod = std::sin(xnode), cnod = std::cos(xnode);
    const double sini = std::sin(xinc), cosi = std::cos(xinc);
    const double xmx = -snod * cosi, xmy = cnod * cosi;
    const double ux = xmx * sinsu + cnod * cossu;
    const double uy = xmy * sinsu + snod * cossu;
    const double uz = sini * sinsu;
    const double vx = xmx * cossu - cnod * sinsu;
    const double vy = xmy * cossu - snod * sinsu;
    const double vz = sini * cossu;
    r[0] = (mrt * ux) * kRe;
    r[1] = (mrt * uy) * kRe;
    r[2] = (mrt * uz) * kRe;
    v[0] = (mvt * ux + rvdot * vx) * vkmpersec;
    v[1] = (mvt * uy + rvdot * vy) * vkmpersec;
    v[2] = (mvt * uz + rvdot * vz) * vkmpersec;
    if (mrt < 1.0) return 6;
    return 0;
}

// ------------------------------------------------------------ TLE parsing

bool field(const std::string &l, size_t c0, size_t c1, std::string &out) {
    // columns c0..c1, 1-based inclusive
    if (l.size() < c1) return false;
    out = l.substr(c0 - 1, c1 - c0 + 1);
    return true;
}
bool to_double(std::string s, double &v) {
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    if (a == std::string::npos) return false;
    s = s.substr(a, b - a + 1);
    char *end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0' && std::isfinite(v);
}
// "+12345-6" style field: sign, five mantissa digits with an implied leading
// decimal point, exponent sign and digit.
bool implied_exp(std::string s, double &v) {
    size_t a = s.find_first_not_of(' ');
    if (a == std::string::npos) return false;
    s = s.substr(a);
    double sign = 1.0;
    if (s[0] == '-' || s[0] == '+') {
        sign = s[0] == '-' ? -1.0 : 1.0;
        s = s.substr(1);
    }
    const size_t e = s.find_last_of("+-");
    if (e == std::string::npos || e == 0) return false;
    double mant, ex;
    if (!to_double("0." + s.substr(0, e), mant) || !to_double(s.substr(e), ex)) return false;
    v = sign * mant * std::pow(10.0, ex);
    return true;
}
int checksum(const std::string &l) {
    int sum = 0;
    for (size_t i = 0; i < 68 && i < l.size(); ++i) {
        const char ch = l[i];
        if (ch >= '0' && ch <= '9') sum += ch - '0';
        else if (ch == '-') sum += 1;
    }
    return sum % 10;
}
// days since 1970-01-01 of a civil date (proleptic Gregorian)
int64_t days_from_civil(int64_t y, unsigned m, unsigned d) {
    y -= m <= 2;
    const int64_t era = (y >= 0 ? y : y - 399) / 400;
    const unsigned yoe = static_cast<unsigned>(y - era * 400);
    const unsigned doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
    const unsigned doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
    return era * 146097 + static_cast<int64_t>(doe) - 719468;
}

bool sat_of(const lpk_tle *e, Sat &s) {
    const double d2r = kPi / 180.0, xpdotp = 1440.0 / kTwoPi;
    return init(s, e->jd_epoch - 2433281.5, e->jd_epoch, e->bstar, e->ecco, e->argpo_deg * d2r, e->inclo_deg * d2r,
                e->mo_deg * d2r, e->no_revday / xpdotp, e->nodeo_deg * d2r);
}

// GMST as the reference's TEME -> ECEF rotation takes it (src/geometry/tle.rs:15-23)
double gmst_rad(double t_unix) {
    const double jd = t_unix / 86400.0 + 2440587.5;
    const double tc = (jd - 2451545.0) / 36525.0;
    const double gmst_sec = 67310.54841 + (876600.0 * 3600.0 + 8640184.812866) * tc + 0.093104 * tc * tc -
                            6.2e-6 * tc * tc * tc;
    double f = std::fmod(gmst_sec / 86400.0, 1.0);  // rem_euclid(1.0)
    if (f < 0.0) f += 1.0;
    return f * kTwoPi;
}

int state_ecef(const Sat &s, const lpk_tle *e, double t_unix, double pos[3], double vel[3]) {
    const double dt_min = (t_unix - e->epoch_unix) / 60.0;
    double r[3], v[3];
    if (propagate(s, dt_min, r, v) != 0) return LPK_E_INVALID;
    const double th = gmst_rad(t_unix);
    const double sn = std::sin(th), cs = std::cos(th);
    const double xt = r[0] * 1000.0, yt = r[1] * 1000.0, zt = r[2] * 1000.0;
    const double xe = xt * cs + yt * sn, ye = -xt * sn + yt * cs;
    pos[0] = xe;
    pos[1] = ye;
    pos[2] = zt;
    if (vel) {
        const double vxt = v[0] * 1000.0, vyt = v[1] * 1000.0, vzt = v[2] * 1000.0;
        vel[0] = vxt * cs + vyt * sn + kOmegaE * ye;
        vel[1] = -vxt * sn + vyt * cs - kOmegaE * xe;
        vel[2] = vzt;
    }
    return LPK_OK;
}

}  // namespace

extern "C" {

int lpk_tle_parse(const char *line1, const char *line2, lpk_tle *out) {
    if (!line1 || !line2 || !out) return LPK_E_INVALID;
    std::string l1(line1), l2(line2);
    while (!l1.empty() && (l1.back() == ' ' || l1.back() == '\r' || l1.back() == '\n')) l1.pop_back();
    while (!l2.empty() && (l2.back() == ' ' || l2.back() == '\r' || l2.back() == '\n')) l2.pop_back();
    if (l1.size() != 69 || l2.size() != 69 || l1[0] != '1' || l2[0] != '2') return LPK_E_INVALID;
    if (checksum(l1) != l1[68] - '0' || checksum(l2) != l2[68] - '0') return LPK_E_INVALID;
    lpk_tle e{};
    std::string f;
    double v;
    field(l1, 3, 7, f);
    if (!to_double(f, v)) return LPK_E_INVALID;
    e.satnum = static_cast<uint32_t>(v);
    field(l1, 19, 20, f);
    if (!to_double(f, v)) return LPK_E_INVALID;
    const int yy = static_cast<int>(v);
    e.epoch_year = yy < 57 ? 2000 + yy : 1900 + yy;  // two-digit years 1957-2056
    field(l1, 21, 32, f);
    if (!to_double(f, e.epoch_day) || !(e.epoch_day >= 1.0 && e.epoch_day < 367.0)) return LPK_E_INVALID;
    field(l1, 34, 43, f);
    if (!to_double(f, e.ndot)) return LPK_E_INVALID;
    field(l1, 45, 52, f);
    if (!implied_exp(f, e.nddot)) return LPK_E_INVALID;
    field(l1, 54, 61, f);
    if (!implied_exp(f, e.bstar)) return LPK_E_INVALID;
    field(l2, 9, 16, f);
    if (!to_double(f, e.inclo_deg)) return LPK_E_INVALID;
    field(l2, 18, 25, f);
    if (!to_double(f, e.nodeo_deg)) return LPK_E_INVALID;
    field(l2, 27, 33, f);
    if (!to_double("0." + f, e.ecco)) return LPK_E_INVALID;
    field(l2, 35, 42, f);
    if (!to_double(f, e.argpo_deg)) return LPK_E_INVALID;
    field(l2, 44, 51, f);
    if (!to_double(f, e.mo_deg)) return LPK_E_INVALID;
    field(l2, 53, 63, f);
    if (!to_double(f, e.no_revday) || !(e.no_revday > 0.0)) return LPK_E_INVALID;

    // The epoch as a calendar time: day of year and the second of that day
    // to the nanosecond, rounded; the UTC unix time then keeps whole
    // microseconds (src/geometry/tle.rs:37-39).
    const double iday = std::floor(e.epoch_day);
    const double sec = (e.epoch_day - iday) * 86400.0;
    const double whole = std::floor(sec);
    double nanos = std::round((sec - whole) * 1e9);
    double whole_s = whole;
    if (nanos >= 1e9) {
        nanos -= 1e9;
        whole_s += 1.0;
    }
    const int64_t d1970 = days_from_civil(e.epoch_year, 1, 1) + static_cast<int64_t>(iday) - 1;
    e.epoch_unix = static_cast<double>(d1970) * 86400.0 + whole_s + std::floor(nanos / 1000.0) * 1e-6;
    e.jd_epoch = 2440587.5 + static_cast<double>(d1970) + (whole_s + nanos * 1e-9) / 86400.0;
    Sat s;
    if (!sat_of(&e, s)) return LPK_E_INVALID;
    *out = e;
    return LPK_OK;
}

int lpk_sgp4_teme(const lpk_tle *tle, double minutes, double r_km[3], double v_kms[3]) {
    if (!tle || !r_km || !v_kms || !std::isfinite(minutes)) return LPK_E_INVALID;
    Sat s;
    if (!sat_of(tle, s)) return LPK_E_INVALID;
    return propagate(s, minutes, r_km, v_kms) == 0 ? LPK_OK : LPK_E_INVALID;
}

int lpk_tle_state_ecef(const lpk_tle *tle, double t_unix, double pos_m[3], double vel_ms[3]) {
    if (!tle || !pos_m || !std::isfinite(t_unix)) return LPK_E_INVALID;
    Sat s;
    if (!sat_of(tle, s)) return LPK_E_INVALID;
    return state_ecef(s, tle, t_unix, pos_m, vel_ms);
}

int lpk_grid_fill_tle(const lpk_tle *tle, double scene_epoch_unix, double duration_s, double grid_dt_s,
                      double *t0_out, double *xyz, uint64_t n) {
    if (!tle || !(grid_dt_s > 0.0) || !t0_out || !xyz) return LPK_E_INVALID;
    if (n != lpk_grid_size(duration_s, grid_dt_s)) return LPK_E_INVALID;
    Sat s;
    if (!sat_of(tle, s)) return LPK_E_INVALID;
    // PrecomputedTrajectory::sample (src/geometry/precomputed.rs:47-70)
    const double start = scene_epoch_unix - 0.5;
    for (uint64_t i = 0; i < n; ++i) {
        double p[3];
        if (state_ecef(s, tle, start + double(i) * grid_dt_s, p, nullptr) != LPK_OK) return LPK_E_INVALID;
        xyz[i] = p[0];
        xyz[n + i] = p[1];
        xyz[2 * n + i] = p[2];
    }
    *t0_out = start;
    return LPK_OK;
}

}  // extern "C"
