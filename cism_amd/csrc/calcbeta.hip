// calcbeta.hip -- basal traction coefficient beta at the vertices from the
// basal velocity: glissade_calcbeta (glissade_basal_traction.F90:69-719)
// for every which_ho_babc option but no-slip (6), the
// f_ground weighting with which_ho_beta_limit (:650-685) and the beta >= 0
// check (:687-699).  One thread per vertex; the fields are read once and
// beta written once (about 40 us at 1027 x 1027 for the Schoof law, bound
// by its FP64 pow calls), so the Picard loop recomputes beta without a host
// round trip.
//
// Roundings: contraction is off and every expression is evaluated in the
// reference's left-to-right order; pow and tan are the device library's
// (within an ulp or two of the host's correctly rounded ones).
#include "cismgpu_internal.h"

namespace cg {

// parameter slots, in the order of cismgpu_calcbeta_in's doubles
enum {
    CB_BETA_CONST, CB_BETA_SMALL, CB_BETA_LARGE, CB_PP_Q, CB_PP_U0, CB_PHIMIN, CB_PHIMAX, CB_BEDMIN, CB_BEDMAX,
    CB_EUS, CB_M, CB_ZI_UT, CB_ZI_NMAX, CB_BUMP_SLOPE, CB_BUMP_WAVELENGTH, CB_POWERLAW_K, CB_BETA_GMIN, CB_UMAX,
    CB_DEW, CB_DNS, CB_NPARAM
};

struct CBDev {
    int babc = 0, limit = 0;
    double p[CB_NPARAM] = {};
    const double *fg = nullptr, *N = nullptr, *cc = nullptr, *pc = nullptr, *mintauf = nullptr, *csf = nullptr,
                 *bext = nullptr, *topg = nullptr, *flwa = nullptr;
    const int32_t *bpmp = nullptr, *imask = nullptr;   // imask: thck > 0 (cells)
};

// glimmer_physcon.F90:57
constexpr double CB_PI = 3.14159265358979;

// ISMIP-HOM C (:343-397) on one task (two halo cells): the global field
// 1000 + 1000 sin(w dew ew) sin(w dns ns), w = 2 pi / (global_ewn dew), on
// the owned cells (ew = i - 2, ns = j - 2), parallel_halo's cell rule
// outside them (periodic: wrap; outflow, no-ice: zero); vertex (i, j) takes
// cell (i, j).
__device__ __forceinline__ double ishomc_at(const Dims& d, double dew, double dns, int i, int j) {
#pragma clang fp contract(off)
    const int n = d.nx - 4, m = d.ny - 4;
    if (d.bc == 0 || d.bc == 2) {
        if (i <= 2) i += n; else if (i >= d.nx - 1) i -= n;
        if (j <= 2) j += m; else if (j >= d.ny - 1) j -= m;
    } else {
        const int w = d.bc == 1 ? 2 : 3;
        if (i <= w || i >= d.nx - w + 1 || j <= w || j >= d.ny - w + 1) return 0.0;
    }
    const double ldomain = n * dew;
    const double omega = 2.0 * CB_PI / ldomain;
    const double x = dew * (i - 2), y = dns * (j - 2);
    return 1000.0 + 1000.0 * sin(omega * x) * sin(omega * y);
}

// beta at every vertex t = (i-1) + nvx (j-1).  Velocity: level lev of a
// system vector (vidx), or a vertex array when lev < 0.  bad: smallest
// vertex index whose beta is negative or NaN (initialised to all ones).
__global__ void k_calcbeta(Dims d, CBDev c, const double* __restrict__ u, const double* __restrict__ v, int lev,
                           double* __restrict__ beta, unsigned long long* __restrict__ bad) {
#pragma clang fp contract(off)
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)d.nvx * d.nvy) return;
    const int i = 1 + (int)(t % d.nvx), j = 1 + (int)(t / d.nvx);
    const long long xi = lev >= 0 ? vidx(d, lev, i, j) : t;
    const double uu = u[xi], vv = v[xi];
    const double smallnum = 1.0e-2;
    // speed with the floor smallnum (:180) and the cap beta_powerlaw_umax (:187-189)
    double speed = sqrt(uu * uu + vv * vv + smallnum * smallnum);
    const double umax = c.p[CB_UMAX];
    if (umax > 0.0 && speed > umax) speed = umax;
    const double m = c.p[CB_M];
    double b = 0.0;
    switch (c.babc) {
    case 0:   // constant (:202-204)
        b = c.p[CB_BETA_CONST];
        break;
    case 1:   // bpmp (:206-212)
        b = c.bpmp[t] == 1 ? c.p[CB_BETA_SMALL] : c.p[CB_BETA_LARGE];
        break;
    case 2: { // pseudo-plastic, tau_c = N coulomb_c (:214-245)
        const double q = c.p[CB_PP_Q], u0 = c.p[CB_PP_U0];
        const double tau_c = c.N[t] * c.cc[t];
        b = tau_c / (pow(u0, q) * pow(speed, 1.0 - q));
        break;
    }
    case 3: { // old pseudo-plastic: phi linear in the bed elevation of the
              // cell with the vertex's indices (:247-289)
        const double q = c.p[CB_PP_Q], u0 = c.p[CB_PP_U0];
        const double phimin = c.p[CB_PHIMIN], phimax = c.p[CB_PHIMAX];
        const double bedmin = c.p[CB_BEDMIN], bedmax = c.p[CB_BEDMAX];
        const double bed = c.topg[c2(d, i, j)] - c.p[CB_EUS];
        double phi;
        if (bed <= bedmin) phi = phimin;
        else if (bed >= bedmax) phi = phimax;
        else phi = phimin + ((bed - bedmin) / (bedmax - bedmin)) * (phimax - phimin);
        const double tau_c = tan(phi * CB_PI / 180.0) * c.N[t];
        b = tau_c / (pow(u0, q) * pow(speed, 1.0 - q));
        break;
    }
    case 4:   // large (:304-309)
        b = c.p[CB_BETA_LARGE];
        break;
    case 5:   // external (:399-413)
        b = c.bext[t];
        break;
    case 7: { // Zoet-Iverson, N capped at zoet_iverson_nmax (:311-341)
        const double nc = c.N[t] < c.p[CB_ZI_NMAX] ? c.N[t] : c.p[CB_ZI_NMAX];
        b = c.cc[t] * nc * pow(speed, 1.0 / m - 1.0) / pow(speed + c.p[CB_ZI_UT], 1.0 / m);
        break;
    }
    case 8:   // ISMIP-HOM C (:343-397)
        b = ishomc_at(d, c.p[CB_DEW], c.p[CB_DNS], i, j);
        break;
    case 9:   // power law (:415-433)
        b = c.pc[t] * pow(speed, 1.0 / m - 1.0);
        break;
    case 10: {// Coulomb friction with the bump parameters (:456-507); the basal
              // rate factor staggered over the cells with thck > 0 (:476-488)
        const double fs = stagger_masked_at(d, c.imask, c.flwa, i, j);
        const double big_lambda = (c.p[CB_BUMP_WAVELENGTH] / c.p[CB_BUMP_SLOPE]) * fs;
        const double N = c.N[t];
        b = c.cc[t] * N * pow(speed, 1.0 / m - 1.0) * pow(speed + pow(N, m) * big_lambda, -1.0 / m);
        if (c.csf) b = b * c.csf[t];
        if (b > 1.0e8) b = 1.0e8;
        break;
    }
    case 11: {// Schoof in the form of Asay-Davis et al. (2016) (:509-579)
        const double num = c.pc[t] * c.cc[t] * c.N[t];
        const double den = pow(pow(c.pc[t], m) * speed + pow(c.cc[t] * c.N[t], m), 1.0 / m);
        b = (num / den) * pow(speed, 1.0 / m - 1.0);
        if (c.csf) b = b * c.csf[t];
        if (b > 1.0e8) b = 1.0e8;
        break;
    }
    case 12: {// modified Schoof, 1/tau_b = 1/tau_p + 1/tau_c (:581-614); with
              // neither stress positive the reference leaves tau_b unset: 0 here
        const double tp = c.pc[t] * pow(speed, 1.0 / m);
        const double tc = c.cc[t] * c.N[t];
        double taub = 0.0;
        if (tc > 0.0 && tp > 0.0) taub = 1.0 / (1.0 / tp + 1.0 / tc);
        else if (tp > 0.0) taub = tp;
        else if (tc > 0.0) taub = tc;
        b = taub / speed;
        break;
    }
    case 13: {// Tsai: the smaller stress (:616-642)
        const double tp = c.pc[t] * pow(speed, 1.0 / m);
        const double tc = c.cc[t] * c.N[t];
        b = (tc <= tp ? tc : tp) / speed;
        if (c.csf) b = b * c.csf[t];
        break;
    }
    case 14: {// power law with effective pressure, p = 3, q = 1 (:435-454)
        const double pp = 3.0, pq = 1.0;
        b = pow(c.p[CB_POWERLAW_K], -1.0 / pp) * pow(c.N[t], pq / pp) * pow(speed, 1.0 / pp - 1.0);
        break;
    }
    case 15:  // yield stress, no speed cap (:291-302)
        b = c.mintauf[t] / sqrt(uu * uu + vv * vv + smallnum * smallnum);
        break;
    default:
        break;
    }
    // f_ground and beta_grounded_min (:667-685)
    const double fg = c.fg[t], gmin = c.p[CB_BETA_GMIN];
    if (c.limit == 0) {
        b = b * fg;
        if (fg > 0.0 && b < gmin) b = gmin;
    } else {
        if (fg > 0.0 && b < gmin) b = gmin;
        b = b * fg;
    }
    beta[t] = b;
    if (!(b >= 0.0)) atomicMin(bad, (unsigned long long)t);
}

// compute_basal_friction_heatflx (glissade_velo_higher.F90:6831-7042) before
// its cell halo, and the basal traction beta u_b (:3817-3818).  One thread
// per cell c = (i-1) + nx (j-1); thread c < nvx nvy also writes the traction
// of vertex c.  which 0 (standard): owned active cells, beta u.u at the four
// Gauss points from the bilinear interpolants (weights 1), over 4 scyr;
// which 1 (local): glissade_unstagger's mean, over cells 2..nx-1 and
// 2..ny-1, of beta (u^2 + v^2) / scyr at active vertices (0 elsewhere).
// Velocity: level lev of a system vector, or vertex arrays when lev < 0.
__global__ void k_basal_friction(Dims d, int which, int nhalo, const int32_t* __restrict__ acell,
                                 const int32_t* __restrict__ avert, const double* __restrict__ u,
                                 const double* __restrict__ v, int lev, const double* __restrict__ beta,
                                 double* __restrict__ bfric, double* __restrict__ btrac) {
#pragma clang fp contract(off)
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n2 = (long long)d.nvx * d.nvy;
    if (c >= (long long)d.nx * d.ny) return;
    auto vel = [&](const double* a, int i, int j) { return a[lev >= 0 ? vidx(d, lev, i, j) : v2(d, i, j)]; };
    if (btrac && c < n2) {
        const int i = 1 + (int)(c % d.nvx), j = 1 + (int)(c / d.nvx);
        btrac[c] = beta[c] * vel(u, i, j);
        btrac[n2 + c] = beta[c] * vel(v, i, j);
    }
    if (!bfric) return;
    const double scyr = 31536000.0;   // glimmer_physcon.F90:90
    const int i = 1 + (int)(c % d.nx), j = 1 + (int)(c / d.nx);
    double f = 0.0;
    if (which == 0) {
        if (i >= nhalo + 1 && i <= d.nx - nhalo && j >= nhalo + 1 && j <= d.ny - nhalo && acell[c]) {
            // nodes SW, SE, NE, NW: vertices (i-1, j-1), (i, j-1), (i, j), (i-1, j)
            const int vi[4] = {i - 1, i, i, i - 1}, vj[4] = {j - 1, j - 1, j, j};
            double un[4], vn[4], bn[4];
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                un[n] = vel(u, vi[n], vj[n]);
                vn[n] = vel(v, vi[n], vj[n]);
                bn[n] = beta[v2(d, vi[n], vj[n])];
            }
            double sum_w = 0.0;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                double uq = 0.0, vq = 0.0, bq = 0.0;
#pragma unroll
                for (int n = 0; n < 4; ++n) {
                    uq = uq + c_phi2[n][p] * un[n];
                    vq = vq + c_phi2[n][p] * vn[n];
                    bq = bq + c_phi2[n][p] * bn[n];
                }
                f = f + 1.0 * bq * (uq * uq + vq * vq);
                sum_w = sum_w + 1.0;
            }
            f = f / (sum_w * scyr);
        }
    } else if (i >= 2 && i <= d.nx - 1 && j >= 2 && j <= d.ny - 1) {
        // stagbfricflx at the four vertices, summed in glissade_unstagger's order
        auto st = [&](int a, int b) {
            if (!avert[v2(d, a, b)]) return 0.0;
            const double ua = vel(u, a, b), va = vel(v, a, b);
            const double s = beta[v2(d, a, b)] * (ua * ua + va * va);
            return s / scyr;
        };
        f = (st(i, j) + st(i - 1, j) + st(i, j - 1) + st(i - 1, j - 1)) / 4.0;
    }
    bfric[c] = f;
}

}  // namespace cg
