/*
 * gwh_basis.h -- the parameter bases of the Fisher and sampler paths (src/fisher.cpp:1841-2511), usable on the
 * host and on the device: unpack maps a parameter set to the basis vector, repack maps a basis vector back.
 */
#ifndef GWH_BASIS_H
#define GWH_BASIS_H

#include <hip/hip_runtime.h>

#include <cmath>

#include "gwat_hip.h"
#include "gwh_model.h"

#define GWH_HD __host__ __device__ inline

/* ------------------------------------------------------------------------- */
/* parameter bases (fisher.cpp:1841-2511)                                    */
/* ------------------------------------------------------------------------- */

GWH_HD static int basis_dim(int basis)
{
    switch (basis) {
    case GWH_BASIS_PHENOMD_SKYAVG: return 7;
    case GWH_BASIS_PHENOMD_FULL: return 11;
    case GWH_BASIS_PHENOMD_MCMC: return 11;
    case GWH_BASIS_PV2_FULL: return 13;
    case GWH_BASIS_PV2_MCMC: return 15;
    }
    return -1;
}

static constexpr double kMsolSec = GWH_MSOL_SEC;
GWH_HD static double mpc_sec() { return GWH_MPC_SI / GWH_C_SI; }
GWH_HD static double chirpmass(double m1, double m2) { return pow(m1 * m2, 3. / 5) / pow(m1 + m2, 1. / 5); }
GWH_HD static double sym_eta(double m1, double m2) { return (m1 * m2) / pow(m1 + m2, 2); }
GWH_HD static double mass1_of(double mc, double eta)
{
    double e = pow(eta, 3. / 5);
    return 1. / 2 * (mc / e + sqrt(1. - 4 * eta) * mc / e);
}
GWH_HD static double mass2_of(double mc, double eta)
{
    double e = pow(eta, 3. / 5);
    return 1. / 2 * (mc / e - sqrt(1. - 4 * eta) * mc / e);
}
GWH_HD static double amp_k(int sky) { return sky ? sqrt(M_PI / 30) : sqrt(M_PI * 40. / 192.); }

GWH_HD static void cart_to_sph(const double *c, double *s)
{
    s[0] = sqrt(c[0] * c[0] + c[2] * c[2] + c[1] * c[1]);
    s[1] = acos(c[2] / s[0]);
    s[2] = atan2(c[1], c[0]);
    if (s[2] < 0) s[2] += 2 * M_PI;
}
GWH_HD static void sph_to_cart(const double *s, double *c)
{
    c[0] = s[0] * sin(s[1]) * cos(s[2]);
    c[1] = s[0] * sin(s[1]) * sin(s[2]);
    c[2] = s[0] * cos(s[1]);
}

GWH_HD static void unpack(const gwh_params &p, int basis, double *v, int *lf)
{
    int dim = basis_dim(basis);
    for (int i = 0; i < dim; i++) lf[i] = 0;
    double mc = chirpmass(p.mass1, p.mass2), eta = sym_eta(p.mass1, p.mass2);
    if (basis == GWH_BASIS_PHENOMD_SKYAVG) {
        lf[0] = lf[3] = lf[4] = 1;
        v[3] = mc;
        v[0] = amp_k(p.sky_average) * (mc * kMsolSec) * (mc * kMsolSec) / (p.DL * mpc_sec()) *
               pow(M_PI * mc * kMsolSec, -7. / 6);
        v[1] = p.phiRef;
        v[2] = p.tc;
        v[4] = eta;
        v[5] = (p.spin1[2] + p.spin2[2]) / 2.;
        v[6] = (p.spin1[2] - p.spin2[2]) / 2.;
        return;
    }
    bool mcmc = (basis == GWH_BASIS_PHENOMD_MCMC || basis == GWH_BASIS_PV2_MCMC);
    v[0] = p.RA;
    v[1] = mcmc ? sin(p.DEC) : p.DEC;
    v[2] = p.psi;
    v[3] = mcmc ? cos(p.incl) : p.incl;
    v[4] = p.phiRef;
    v[5] = p.tc;
    if (mcmc) {
        v[6] = log(p.DL);
        v[7] = log(mc);
    } else {
        lf[6] = lf[7] = 1;
        v[6] = p.DL;
        v[7] = mc;
    }
    v[8] = eta;
    if (basis == GWH_BASIS_PV2_MCMC) {
        double s1[3], s2[3];
        cart_to_sph(p.spin1, s1);
        cart_to_sph(p.spin2, s2);
        v[9] = s1[0];
        v[10] = s2[0];
        v[11] = cos(s1[1]);
        v[12] = cos(s2[1]);
        v[13] = s1[2];
        v[14] = s2[2];
    } else {
        v[9] = p.spin1[2];
        v[10] = p.spin2[2];
        if (basis == GWH_BASIS_PV2_FULL) {
            v[11] = p.chip;
            v[12] = p.phip;
        }
    }
}

/* a: a gen_params carrying only the non-parameter options (fisher.cpp:2513-2572) */
GWH_HD static void repack(const double *v, int basis, gwh_params &a)
{
    if (basis == GWH_BASIS_PHENOMD_SKYAVG) {
        a.mass1 = mass1_of(v[3], v[4]);
        a.mass2 = mass2_of(v[3], v[4]);
        double mcs = v[3] * kMsolSec;
        a.DL = amp_k(a.sky_average) * mcs * mcs / v[0] * pow(M_PI * mcs, -7. / 6) / mpc_sec();
        a.tc = v[2];
        a.phiRef = v[1];
        a.spin1[2] = v[5] + v[6];
        a.spin2[2] = v[5] - v[6];
        return;
    }
    bool mcmc = (basis == GWH_BASIS_PHENOMD_MCMC || basis == GWH_BASIS_PV2_MCMC);
    a.mass1 = mass1_of(mcmc ? exp(v[7]) : v[7], v[8]);
    a.mass2 = mass2_of(mcmc ? exp(v[7]) : v[7], v[8]);
    a.DL = mcmc ? exp(v[6]) : v[6];
    a.RA = v[0];
    a.DEC = mcmc ? asin(v[1]) : v[1];
    a.psi = v[2];
    a.incl = mcmc ? acos(v[3]) : v[3];
    a.phiRef = v[4];
    a.tc = v[5];
    if (basis == GWH_BASIS_PV2_MCMC) {
        /* bump cos(theta) back into [-1, 1] (fisher.cpp:2188-2210) */
        double t1 = v[11] > 1 ? 0 : (v[11] < -1 ? M_PI : acos(v[11]));
        double t2 = v[12] > 1 ? 0 : (v[12] < -1 ? M_PI : acos(v[12]));
        double s1[3] = {v[9], t1, v[13]}, s2[3] = {v[10], t2, v[14]};
        sph_to_cart(s1, a.spin1);
        sph_to_cart(s2, a.spin2);
    } else {
        a.spin1[2] = v[9];
        a.spin2[2] = v[10];
        if (basis == GWH_BASIS_PV2_FULL) {
            a.chip = v[11];
            a.phip = v[12];
        }
    }
}

#endif
