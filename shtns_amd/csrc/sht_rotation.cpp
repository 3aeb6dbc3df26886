// sht_rotation.cpp -- rotations of spherical harmonic expansions on the host (SHTns's
// shtns_rotation_* objects and the SH_Yrotate / SH_Yrotate90 / SH_Xrotate90 shorthands,
// shtns.h:193-218, sht_func.c:342-398 and 1394-2229).
//
// A rotation R acts on a field by f'(x) = f(R^-1 x) (an active rotation, right-handed: a
// rotation by +a about z maps phi to phi + a, as SH_Zrotate does).  In orthonormal harmonics
// with the Condon-Shortley phase the coefficients of each degree l transform as
//     a'_{m'} = sum_m D^l_{m'm} a_m,
// and for R = Rz(g) Ry(b) Rz(a) (a about z first, then b about y, then g about z; the order
// in which shtns_rotation_apply_real applies e^{-i m alpha}, the Wigner matrix and
// e^{-i m gamma}, sht_func.c:1721-1940) D^l_{m'm} = e^{-i m' g} d^l_{m'm}(b) e^{-i m a}, with
// d^l_{m'm}(b) = <l m'| exp(-i b J_y) |l m>.
//
// The Wigner d functions are generated per pair (m', m) by the three-term recurrence in the
// degree (a Jacobi-polynomial recurrence, stable upward like the Legendre one):
//     j sqrt(((j+1)^2-m^2)((j+1)^2-m'^2)) d^{j+1}
//         = (2j+1) (j(j+1) cos b - m m') d^j - (j+1) sqrt((j^2-m^2)(j^2-m'^2)) d^{j-1},
// started at j0 = max(|m|,|m'|) from the closed form (for m' >= |m|)
//     d^{m'}_{m'm} = sqrt(C(2m', m'+m)) cos(b/2)^{m'+m} (-sin(b/2))^{m'-m},
// the other pairs by the symmetries d_{m'm} = (-1)^{m-m'} d_{mm'} = d_{-m,-m'}.  The seed is
// carried as a mantissa with a separate power-of-two exponent, so pairs whose seed lies far
// below the double range (large m, small b) still grow correctly.  This replaces SHTns's
// recursion over m' (Gumerov-Duraiswami); both give the same matrices.  Work is O(lmax^3),
// spread over host threads by output order m' (each thread owns its outputs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "sht_internal.h"

using shtns_amd::set_error;
using std::complex;
typedef complex<double> zd;

struct shtns_rot_ {
    int lmax, mmax;
    bool cs_phase;        // orthonormal harmonics with (true) or without the Condon-Shortley phase
    bool real_norm;       // SHT_REAL_NORM: m = 0 coefficients of real fields scaled by 1/sqrt(2)
    double alpha = 0.0, beta = 0.0, gamma = 0.0;
    double cb = 1.0, sb = 0.0;           // cos(beta), sin(beta)
    long double ch = 1.0L, sh = 0.0L;    // cos(beta/2), sin(beta/2)
    std::vector<double> rt, irt;         // rt[|m|][j] = sqrt(j^2 - m^2), irt = 1/rt (j > |m|)
};

namespace {

// start of row |m| of the triangular tables (row m holds degrees j = m .. lmax+1)
inline long tri(int m, int lmax) { return (long)m * (lmax + 2) - (long)m * (m - 1) / 2; }

// row |m| of the sqrt tables holds j = |m| .. lmax+1
inline const double *rt_row(const shtns_rot_ *r, const std::vector<double> &t, int m) {
    const int am = m < 0 ? -m : m;
    return t.data() + tri(am, r->lmax) - am;   // index by j
}

// d^{j}_{mp,m}(beta) for j = 0..jmax into out[j] (zero below j0 = max(|m|,|mp|))
void wigner_column(const shtns_rot_ *r, int mp, int m, double *out, int jmax) {
    const int lmax = jmax;
    const int am = std::abs(m), amp = std::abs(mp);
    const int j0 = std::max(am, amp);
    // map to the fundamental domain a >= |b| for the seed: d_{mp,m} = (-1)^(m-mp) d_{m,mp} = d_{-m,-mp}
    int a = mp, b = m;
    double sgn = 1.0;
    if (amp < am) {
        if (m > 0) { a = m; b = mp; sgn = ((m - mp) & 1) ? -1.0 : 1.0; }
        else { a = -m; b = -mp; }
    } else if (mp < 0) {   // |mp| >= |m| with mp < 0: d_{mp,m} = d_{-m,-mp}, then swap
        // d_{-m,-mp} with -mp >= |m|: = (-1)^(-mp - (-m)) d_{-mp,-m}
        a = -mp;
        b = -m;
        sgn = ((m - mp) & 1) ? -1.0 : 1.0;
    }
    // seed sqrt(C(2a, a+b)) ch^(a+b) (-sh)^(a-b), as mantissa * 2^E
    const long double ch = r->ch, sh = r->sh;
    double val = 0.0;
    int E = 0;
    const bool zc = (ch == 0.0L && a + b > 0), zs = (sh == 0.0L && a - b > 0);
    if (!zc && !zs) {
        long double lg = 0.5L * (lgammal(2.0L * a + 1) - lgammal((long double)a + b + 1) - lgammal((long double)a - b + 1)) /
                         logl(2.0L);
        if (a + b > 0) lg += (a + b) * log2l(fabsl(ch));
        if (a - b > 0) lg += (a - b) * log2l(fabsl(sh));
        long double s = 1.0L;
        if (ch < 0 && ((a + b) & 1)) s = -s;
        if (sh > 0 && ((a - b) & 1)) s = -s;   // (-sh)^(a-b)
        const long double fl = floorl(lg);
        E = (int)fl;
        val = (double)(s * exp2l(lg - fl)) * sgn;
    }
    // recurrence on scaled values: true d = v * 2^E (E <= 0: |d| <= 1)
    const double *rm = rt_row(r, r->rt, m), *rmp = rt_row(r, r->rt, mp);
    const double *im_ = rt_row(r, r->irt, m), *imp = rt_row(r, r->irt, mp);
    const double cb = r->cb, mm = (double)m * mp;
    double v0 = 0.0, v1 = val;
    for (int j = 0; j < j0 && j <= lmax; ++j) out[j] = 0.0;
    if (j0 > lmax) return;
    out[j0] = (E < -1070) ? 0.0 : ldexp(v1, E);
    for (int j = j0; j < lmax; ++j) {
        double v2;
        if (j == 0) {
            v2 = cb * v1;   // only m = mp = 0 starts at j = 0
        } else {
            const double c1 = (2.0 * j + 1.0) * ((double)j * (j + 1) * cb - mm);
            const double c0 = (j + 1.0) * ((j > am && j > amp) ? rm[j] * rmp[j] : 0.0);
            v2 = (c1 * v1 - c0 * v0) * (imp[j + 1] * im_[j + 1] / j);
        }
        v0 = v1;
        v1 = v2;
        if (E < 0 && fabs(v1) > 0x1p300) {
            v0 = ldexp(v0, -256);
            v1 = ldexp(v1, -256);
            E += 256;
        }
        out[j + 1] = (E < -1070) ? 0.0 : ldexp(v1, E);
    }
}

// index of (l, m), -l <= m <= l, |m| <= mmax, in a complex field's spectrum (shtns.h:110-114)
inline long lm_cplx(int l, int m, int mmax) {
    return (l <= mmax) ? (long)l * (l + 1) + m : (long)mmax * (2 * l - mmax) + l + m;
}

int n_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h ? h : 1u, 64u));
}

// Rotate full complex spectra Z (indexed by lm_cplx, degrees 0..lmax, |m| <= mmax) into R
// (R != Z).  Output orders m' >= 0 are written, and m' < 0 too when both_signs.
void rotate_spectra(const shtns_rot_ *r, const zd *Z, zd *R, bool both_signs) {
    const int L = r->lmax, M = r->mmax;
    // Z multiplied by e^{-i m alpha} once
    std::vector<zd> Za((size_t)lm_cplx(L, M, M) + 1);
    {
        std::vector<zd> ph(2 * M + 1);
        for (int m = -M; m <= M; ++m) ph[m + M] = std::polar(1.0, -r->alpha * m);
        for (int l = 0; l <= L; ++l) {
            const int ml = std::min(l, M);
            for (int m = -ml; m <= ml; ++m) Za[lm_cplx(l, m, M)] = Z[lm_cplx(l, m, M)] * ph[m + M];
        }
    }
    const int nt = n_threads();
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) {
        pool.emplace_back([&, t]() {
            std::vector<double> d(L + 1);
            std::vector<zd> acc(L + 1), accn(L + 1);
            for (int mp = t; mp <= M; mp += nt) {
                std::fill(acc.begin(), acc.end(), zd(0.0));
                std::fill(accn.begin(), accn.end(), zd(0.0));
                for (int m = -M; m <= M; ++m) {
                    wigner_column(r, mp, m, d.data(), L);
                    const int j0 = std::max(mp, std::abs(m));
                    // d_{-mp,-m} = (-1)^(mp-m) d_{mp,m}
                    const double par = ((mp - m) & 1) ? -1.0 : 1.0;
                    for (int j = j0; j <= L; ++j) {
                        acc[j] += d[j] * Za[lm_cplx(j, m, M)];
                        if (both_signs && mp > 0) accn[j] += (par * d[j]) * Za[lm_cplx(j, -m, M)];
                    }
                }
                const zd gp = std::polar(1.0, -r->gamma * mp), gn = std::polar(1.0, r->gamma * mp);
                for (int j = mp; j <= L; ++j) {
                    R[lm_cplx(j, mp, M)] = acc[j] * gp;
                    if (both_signs && mp > 0) R[lm_cplx(j, -mp, M)] = accn[j] * gn;
                }
            }
        });
    }
    for (auto &th : pool) th.join();
}

void build_tables(shtns_rot_ *r) {
    const int L = r->lmax;
    const long n = tri(L + 1, L);
    r->rt.assign(n, 0.0);
    r->irt.assign(n, 0.0);
    for (int am = 0; am <= L; ++am) {
        double *row = r->rt.data() + tri(am, L) - am;
        double *irow = r->irt.data() + tri(am, L) - am;
        for (int j = am; j <= L + 1; ++j) {
            const double v = sqrt((double)j * j - (double)am * am);
            row[j] = v;
            irow[j] = (v > 0) ? 1.0 / v : 0.0;
        }
    }
}

}  // namespace

extern "C" shtns_rot shtns_rotation_create(const int lmax, const int mmax, int norm) {
    shtns_amd::clear_error();
    if (lmax < 1 || mmax < 0 || mmax > lmax) {
        set_error(1, "shtns_rotation_create: need 0 <= mmax <= lmax, lmax >= 1 (lmax=%d, mmax=%d)", lmax, mmax);
        return nullptr;
    }
    shtns_rot_ *r = new shtns_rot_();
    r->lmax = lmax;
    r->mmax = mmax;
    r->cs_phase = !(norm & SHT_NO_CS_PHASE);
    r->real_norm = (norm & SHT_REAL_NORM) != 0;
    build_tables(r);
    return r;
}

extern "C" void shtns_rotation_destroy(shtns_rot r) { delete r; }

extern "C" void shtns_rotation_set_angles_ZYZ(shtns_rot r, double alpha, double beta, double gamma) {
    shtns_amd::clear_error();
    if (!(fabs(beta) <= M_PI)) {   // SHTns stops here (sht_func.c:1446-1449)
        fprintf(stderr, "ERROR: angle 'beta' must be between -pi and pi\n");
        exit(1);
    }
    if (beta == 0.0) {   // a rotation about z only
        alpha += gamma;
        gamma = 0.0;
    }
    r->alpha = alpha;
    r->beta = beta;
    r->gamma = gamma;
    const long double hb = 0.5L * (long double)beta;
    r->ch = cosl(hb);
    r->sh = sinl(hb);
    r->cb = cos(beta);
    r->sb = sin(beta);
}

extern "C" void shtns_rotation_set_angles_ZXZ(shtns_rot r, double alpha, double beta, double gamma) {
    // Rz(g) Rx(b) Rz(a) = Rz(g - pi/2) Ry(b) Rz(a + pi/2)   (sht_func.c:1480-1484)
    shtns_rotation_set_angles_ZYZ(r, alpha + M_PI / 2, beta, gamma - M_PI / 2);
}

extern "C" void shtns_rotation_set_angle_axis(shtns_rot r, double theta, double Vx, double Vy, double Vz) {
    const double n = sqrt(Vx * Vx + Vy * Vy + Vz * Vz);
    if (!(n > 0.0)) {
        set_error(1, "shtns_rotation_set_angle_axis: null axis");
        return;
    }
    const double x = Vx / n, y = Vy / n, z = Vz / n, c = cos(theta), s = sin(theta), t = 1.0 - c;
    // Rodrigues' rotation matrix, then R = Rz(g) Ry(b) Rz(a):
    //   R[2][2] = cos b, (R[0][2], R[1][2]) = sin b (cos g, sin g), (R[2][0], R[2][1]) = sin b (-cos a, sin a)
    const double R00 = c + x * x * t, R01 = x * y * t - z * s, R02 = x * z * t + y * s;
    const double R10 = y * x * t + z * s, R11 = c + y * y * t, R12 = y * z * t - x * s;
    const double R20 = z * x * t - y * s, R21 = z * y * t + x * s, R22 = c + z * z * t;
    const double sbeta = sqrt(R02 * R02 + R12 * R12);
    double a, b, g;
    if (sbeta > 1e-14) {
        b = atan2(sbeta, R22);
        g = atan2(R12, R02);
        a = atan2(R21, -R20);
    } else if (R22 > 0) {   // b = 0: Rz(a + g)
        b = 0.0;
        g = 0.0;
        a = atan2(R10, R00);
    } else {                // b = pi: Ry(pi) Rz(a) = [[-ca, sa, 0], [sa, ca, 0], [0, 0, -1]]
        b = M_PI;
        g = 0.0;
        a = atan2(R10, R11);
        (void)R01;
    }
    shtns_rotation_set_angles_ZYZ(r, a, b, g);
}

extern "C" int shtns_rotation_wigner_d_matrix(shtns_rot r, const int l, double *mx) {
    shtns_amd::clear_error();
    if (l < 0 || l > r->lmax) {
        set_error(1, "shtns_rotation_wigner_d_matrix: 0 <= l <= lmax not satisfied (l=%d)", l);
        return 0;
    }
    const int w = 2 * l + 1;
    std::vector<double> col(r->lmax + 1);
    for (int mp = -l; mp <= l; ++mp)
        for (int m = -l; m <= l; ++m) {
            wigner_column(r, mp, m, col.data(), l);
            mx[(size_t)(mp + l) * w + (m + l)] = col[l];
        }
    return w;
}

extern "C" void shtns_rotation_apply_cplx(shtns_rot r, cplx *Zlm, cplx *Rlm) {
    const int L = r->lmax, M = r->mmax;
    const long n = lm_cplx(L, M, M) + 1;
    const zd *Z = reinterpret_cast<const zd *>(Zlm);
    std::vector<zd> in(Z, Z + n), out(n, zd(0.0));
    if (!r->cs_phase)   // coefficients of Y without (-1)^m: a(CS) = (-1)^m a
        for (int l = 0; l <= L; ++l)
            for (int m = -std::min(l, M); m <= std::min(l, M); ++m)
                if (m & 1) in[lm_cplx(l, m, M)] = -in[lm_cplx(l, m, M)];
    rotate_spectra(r, in.data(), out.data(), true);
    if (!r->cs_phase)
        for (int l = 0; l <= L; ++l)
            for (int m = -std::min(l, M); m <= std::min(l, M); ++m)
                if (m & 1) out[lm_cplx(l, m, M)] = -out[lm_cplx(l, m, M)];
    std::copy(out.begin(), out.end(), reinterpret_cast<zd *>(Rlm));
}

extern "C" void shtns_rotation_apply_real(shtns_rot r, cplx *Qlm, cplx *Rlm) {
    const int L = r->lmax, M = r->mmax;
    const long n = lm_cplx(L, M, M) + 1;
    const zd *Q = reinterpret_cast<const zd *>(Qlm);
    zd *Ro = reinterpret_cast<zd *>(Rlm);
    // real field, m >= 0 stored in LM order: a_{l,-m} = (-1)^m conj(a_{lm}) (CS phase)
    auto lm_real = [&](int l, int m) { return (long)m * (2 * L + 1 - m) / 2 + l; };
    const double m0 = r->real_norm ? sqrt(2.0) : 1.0;
    std::vector<zd> in(n), out(n);
    for (int l = 0; l <= L; ++l) {
        in[lm_cplx(l, 0, M)] = zd(Q[lm_real(l, 0)].real() * m0, 0.0);
        for (int m = 1; m <= std::min(l, M); ++m) {
            zd a = Q[lm_real(l, m)];
            if (!r->cs_phase && (m & 1)) a = -a;
            in[lm_cplx(l, m, M)] = a;
            in[lm_cplx(l, -m, M)] = (m & 1) ? -std::conj(a) : std::conj(a);
        }
    }
    rotate_spectra(r, in.data(), out.data(), false);
    for (int l = 0; l <= L; ++l) {
        Ro[lm_real(l, 0)] = zd(out[lm_cplx(l, 0, M)].real() / m0, 0.0);
        for (int m = 1; m <= std::min(l, M); ++m) {
            zd a = out[lm_cplx(l, m, M)];
            if (!r->cs_phase && (m & 1)) a = -a;
            Ro[lm_real(l, m)] = a;
        }
    }
}

// ---- SHTns's rotations of a configuration's spectra (sht_func.c:342-398): mres = 1 and
//      mmax = lmax are required, as in SHTns.
static void cfg_rotate(shtns_cfg c, cplx *Q, cplx *R, double a, double b, double g, const char *who) {
    if (c->mres != 1 || c->mmax < c->lmax) {
        fprintf(stderr, "*** [shtns_amd] %s: truncature makes rotation not closed.\n", who);
        exit(1);
    }
    // factors depending on l only commute with rotations: only the phase and real-norm flags matter
    shtns_rot r = shtns_rotation_create(c->lmax, c->mmax, c->norm & (SHT_NO_CS_PHASE | SHT_REAL_NORM));
    shtns_rotation_set_angles_ZYZ(r, a, b, g);
    shtns_rotation_apply_real(r, Q, R);
    shtns_rotation_destroy(r);
}
extern "C" void SH_Yrotate(shtns_cfg c, cplx *Qlm, double alpha, cplx *Rlm) {
    cfg_rotate(c, Qlm, Rlm, 0.0, alpha, 0.0, "SH_Yrotate");
}
extern "C" void SH_Yrotate90(shtns_cfg c, cplx *Qlm, cplx *Rlm) {
    cfg_rotate(c, Qlm, Rlm, 0.0, M_PI / 2, 0.0, "SH_Yrotate90");
}
extern "C" void SH_Xrotate90(shtns_cfg c, cplx *Qlm, cplx *Rlm) {
    // Rx(pi/2) = Rz(-pi/2) Ry(pi/2) Rz(pi/2)
    cfg_rotate(c, Qlm, Rlm, M_PI / 2, M_PI / 2, -M_PI / 2, "SH_Xrotate90");
}
