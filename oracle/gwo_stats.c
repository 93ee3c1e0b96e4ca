/*
 * Detector projection, likelihood and Fisher oracle (TEST INFRASTRUCTURE):
 * FP64 restatement of src/detector_util.cpp, src/waveform_util.cpp,
 * src/mcmc_gw.cpp and src/fisher.cpp for the terrestrial-detector path.
 */
#include <stdlib.h>
#include <string.h>
#include "gwo_internal.h"
#include "../gw_analysis_tools_amd/csrc/gwh_tables.h"

static const double geometric_factor[GWH_NDET_TABLE] = {1., 1., 1., 1., 1., 1., 0.86602540378, 0.86602540378,
                                                        0.86602540378};

/* detector_util.cpp:1083-1158 name dispatch */
int gwo_detector_id(const char *d)
{
    if (!strcmp(d, "Hanford") || !strcmp(d, "hanford")) return GWH_DET_HANFORD;
    if (!strcmp(d, "Livingston") || !strcmp(d, "livingston")) return GWH_DET_LIVINGSTON;
    if (!strcmp(d, "Virgo") || !strcmp(d, "virgo")) return GWH_DET_VIRGO;
    if (!strcmp(d, "Kagra") || !strcmp(d, "kagra")) return GWH_DET_KAGRA;
    if (!strcmp(d, "Indigo") || !strcmp(d, "indigo")) return GWH_DET_INDIGO;
    if (!strcmp(d, "Cosmic Explorer") || !strcmp(d, "cosmic explorer") || !strcmp(d, "CE")) return GWH_DET_CE;
    if (!strcmp(d, "Einstein Telescope 1") || !strcmp(d, "einstein telescope 1") || !strcmp(d, "ET1")) return GWH_DET_ET1;
    if (!strcmp(d, "Einstein Telescope 2") || !strcmp(d, "einstein telescope 2") || !strcmp(d, "ET2")) return GWH_DET_ET2;
    if (!strcmp(d, "Einstein Telescope 3") || !strcmp(d, "einstein telescope 3") || !strcmp(d, "ET3")) return GWH_DET_ET3;
    return -1;
}

/* detector_util.cpp:900-1016: F+ and Fx from the response tensor (Anderson et al. 2001, B4-B7) */
int gwo_antenna(int det, double ra, double dec, double psi, double gmst, double *fplus, double *fcross)
{
    if (det < 0 || det >= GWH_NDET_TABLE) return 0;
    const double *D = gwh_det_tensor[det];
    double gha = gmst - ra;
    double cosgha = cos(gha), singha = sin(gha), cosdec = cos(dec), sindec = sin(dec);
    double cospsi = cos(psi), sinpsi = sin(psi);
    double X[3], Y[3];
    X[0] = -cospsi * singha - sinpsi * cosgha * sindec;
    X[1] = -cospsi * cosgha + sinpsi * singha * sindec;
    X[2] = sinpsi * cosdec;
    Y[0] = sinpsi * singha - cospsi * cosgha * sindec;
    Y[1] = sinpsi * cosgha + cospsi * singha * sindec;
    Y[2] = cospsi * cosdec;
    double fp = 0.0, fx = 0.0;
    for (int i = 0; i < 3; i++) {
        double DX = D[3 * i] * X[0] + D[3 * i + 1] * X[1] + D[3 * i + 2] * X[2];
        double DY = D[3 * i] * Y[0] + D[3 * i + 1] * Y[1] + D[3 * i + 2] * Y[2];
        fp += X[i] * DX - Y[i] * DY;
        fx += X[i] * DY + Y[i] * DX;
    }
    *fplus = fp * geometric_factor[det];
    *fcross = fx * geometric_factor[det];
    return 1;
}

/* detector_util.cpp:676-790: arrival-time difference det1 - det2 (seconds) */
double gwo_dtoa(double ra, double dec, double gmst, int det1, int det2)
{
    const double *l1 = gwh_det_location[det1], *l2 = gwh_det_location[det2];
    double dx[3] = {l1[0] - l2[0], l1[1] - l2[1], l1[2] - l2[2]};
    double ha = gmst - ra;
    double e[3] = {cos(dec) * cos(ha), cos(dec) * -sin(ha), sin(dec)};
    return (dx[0] * e[0] + dx[1] * e[1] + dx[2] * e[2]) / GWO_C_SI;
}

/* util.cpp:1797-1860 */
double gwo_gps_to_gmst_radian(double gps)
{
    double J2000 = 2451545, J2000_GPST = 630763213.;
    double JD = J2000 + (gps - J2000_GPST) / (86400.);
    double JD0 = floor(JD) + .5;
    double H = (JD - JD0) * 24;
    double D0 = JD0 - J2000, D = JD - J2000;
    double tau = D / 36525.;
    double g = 6.697374558 + 0.06570982441908 * D0 + 1.00273790935 * H + 0.000026 * tau * tau;
    double hours = (double)((int)floor(g) % 24);
    double fraction = g - floor(g);
    return (hours + fraction) * M_PI / 12.;
}

/* waveform_util.cpp:864-873 and :935-985 (terrestrial, equatorial coordinates) */
int gwo_detector_response(const double *f, int n, double *response, int det, int model, const gwo_params *p)
{
    double *hp = (double *)malloc(sizeof(double) * 2 * n);
    double *hc = (double *)malloc(sizeof(double) * 2 * n);
    int status = gwo_fourier_waveform(f, n, hp, hc, model, p);
    double fp, fx;
    if (!gwo_antenna(det, p->RA, p->DEC, p->psi, p->gmst, &fp, &fx)) status = 0;
    for (int i = 0; i < n; i++) {
        response[2 * i] = fp * hp[2 * i] + fx * hc[2 * i];
        response[2 * i + 1] = fp * hp[2 * i + 1] + fx * hc[2 * i + 1];
    }
    free(hp);
    free(hc);
    return status;
}

/* util.h:841-856: weights 1,4,2,4,...,(2|4),1 by index parity whatever n is */
double gwo_simpsons_sum(double dx, int n, const double *integrand)
{
    double sum = 0;
    for (int i = 1; i < n - 1; i++) {
        if (i % 2 == 0)
            sum += 2 * integrand[i];
        else
            sum += 4 * integrand[i];
    }
    sum += integrand[0] + integrand[n - 1];
    return sum * dx / 3;
}

/* mcmc_gw.cpp:801-868: -1/2[(h|h) - 2(d|h)] */
double gwo_loglike_internal(const double *data, const double *psd, const double *f, const double *weights,
                            const double *r, int n, int log10f, int gaussleg)
{
    double delta_f = f[n / 2] - f[n / 2 - 1];
    double *integrand = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) integrand[i] = (r[2 * i] * r[2 * i] + r[2 * i + 1] * r[2 * i + 1]) / psd[i];
    double HH = 0;
    if (!gaussleg) {
        HH = 4. * gwo_simpsons_sum(delta_f, n, integrand);
    } else {
        for (int i = 0; i < n; i++) HH += weights[i] * integrand[i] * (log10f ? f[i] * GWO_LOG10 : 1.0);
        HH *= 4;
    }
    for (int i = 0; i < n; i++)
        integrand[i] = (data[2 * i] * r[2 * i] + data[2 * i + 1] * r[2 * i + 1]) / psd[i];
    double DH = 0;
    if (!gaussleg) {
        DH = 4. * gwo_simpsons_sum(delta_f, n, integrand);
    } else {
        for (int i = 0; i < n; i++) DH += weights[i] * integrand[i] * (log10f ? f[i] * GWO_LOG10 : 1.0);
        DH *= 4;
    }
    free(integrand);
    return -0.5 * (HH - 2 * DH);
}

/* mcmc_gw.cpp:570-591 */
double gwo_log_likelihood(const double *data, const double *psd, const double *f, int n, const gwo_params *p,
                          int det, int model)
{
    double *r = (double *)malloc(sizeof(double) * 2 * n);
    gwo_detector_response(f, n, r, det, model, p);
    double ll = gwo_loglike_internal(data, psd, f, NULL, r, n, 0, 0);
    free(r);
    return ll;
}

/* mcmc_gw.cpp:2374-2489 + waveform_util.cpp:152-184 (one waveform, per-detector
 * projection and time-of-arrival phase e^{-2 pi i f dt}) */
double gwo_network_loglike(const gwo_params *pp, int model, const double *f, int n, const double *const *data,
                           const double *const *psd, const int *dets, int ndet)
{
    return gwo_network_loglike_ex(pp, model, f, n, data, psd, dets, ndet, NULL, 0, 0);
}

/* the same with the integration rule of Log_Likelihood_internal (mcmc_gw.cpp:801-868): Simpson, or the caller's
 * Gauss-Legendre weights (times f ln 10 on a log10 grid), as MCMC_likelihood_extrinsic passes them through */
double gwo_network_loglike_ex(const gwo_params *pp, int model, const double *f, int n, const double *const *data,
                              const double *const *psd, const int *dets, int ndet, const double *weights, int log10f,
                              int gaussleg)
{
    double *hp = (double *)malloc(sizeof(double) * 2 * n);
    double *hc = (double *)malloc(sizeof(double) * 2 * n);
    double *r = (double *)malloc(sizeof(double) * 2 * n);
    gwo_params p = *pp;
    gwo_fourier_waveform(f, n, hp, hc, model, &p);
    double ll = 0;
    for (int d = 0; d < ndet; d++) {
        double dt = gwo_dtoa(p.RA, p.DEC, p.gmst, dets[0], dets[d]);
        double tcd = -dt;
        tcd *= 2 * M_PI;
        double fp, fx;
        gwo_antenna(dets[d], p.RA, p.DEC, p.psi, p.gmst, &fp, &fx);
        for (int j = 0; j < n; j++) {
            double rr = fp * hp[2 * j] + fx * hc[2 * j];
            double ri = fp * hp[2 * j + 1] + fx * hc[2 * j + 1];
            double ph = tcd * f[j];
            gwo_cmul(rr, ri, cos(ph), sin(ph), &r[2 * j], &r[2 * j + 1]);
        }
        ll += gwo_loglike_internal(data[d], psd[d], f, weights, r, n, log10f, gaussleg);
    }
    free(hp);
    free(hc);
    free(r);
    return ll;
}

double gwo_mcmc_likelihood_extrinsic(const gwo_params *p_in, int model, const double *f, int n,
                                     const double *const *data, const double *const *psd, const int *dets,
                                     int ndet, double T_eff)
{
    gwo_params p = *p_in;
    p.tc = T_eff - p_in->tc;
    return gwo_network_loglike(&p, model, f, n, data, psd, dets, ndet);
}

/* waveform_util.cpp:479-510 (SIMPSONS) */
double gwo_snr(const double *psd, const double *w, const double *f, int n)
{
    double *integrand = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) integrand[i] = 4. * (w[2 * i] * w[2 * i] + w[2 * i + 1] * w[2 * i + 1]) / psd[i];
    double v = gwo_simpsons_sum(f[1] - f[0], n, integrand);
    free(integrand);
    return sqrt(v);
}

/* ------------------------------- Fisher ------------------------------- */

static int has(const char *s, const char *sub) { return strstr(s, sub) != NULL; }

/* fisher.cpp:1754-1770 */
static void local_method(const char *m, char *out)
{
    size_t L = strlen(m);
    if (has(m, "MCMC") && has(m, "Full")) {
        memcpy(out, m + 5, L - 10);
        out[L - 10] = 0;
    } else if (has(m, "MCMC")) {
        strcpy(out, m + 5);
    } else {
        strcpy(out, m);
    }
}

/* the waveform model of a (local) generation method; the ppE and gIMR methods keep their variant */
static int model_of(const char *m)
{
    const int pv2 = has(m, "IMRPhenomPv2");
    if (has(m, "gIMR")) return pv2 ? GWO_GIMRPHENOMPV2 : GWO_GIMRPHENOMD;
    if (has(m, "ppE")) {
        const int ins = has(m, "Inspiral");
        return pv2 ? (ins ? GWO_PPE_IMRPHENOMPV2_INSPIRAL : GWO_PPE_IMRPHENOMPV2_IMR)
                   : (ins ? GWO_PPE_IMRPHENOMD_INSPIRAL : GWO_PPE_IMRPHENOMD_IMR);
    }
    return pv2 ? GWO_IMRPHENOMPV2 : GWO_IMRPHENOMD;
}

static void cart_sph(const double *c, double *s)
{
    s[0] = sqrt(c[0] * c[0] + c[2] * c[2] + c[1] * c[1]);
    s[1] = acos(c[2] / s[0]);
    s[2] = atan2(c[1], c[0]);
    if (s[2] < 0) s[2] += 2 * M_PI;
}
static void sph_cart(const double *s, double *c)
{
    c[0] = s[0] * sin(s[1]) * cos(s[2]);
    c[1] = s[0] * sin(s[1]) * sin(s[2]);
    c[2] = s[0] * cos(s[1]);
}

/* fisher.cpp:1841-2161 (terrestrial, non-equatorial-orientation branches) */
static int unpack_gr(double *v, int *lf, const gwo_params *p, const char *m, int dim);
static int repack_gr(const double *v, gwo_params *a, const char *m);

/* the table entry of a gIMR list code: 0..13 delta_phi of PN order code - 4, 14..16 delta_sigma 2..4,
 * 17..18 delta_beta 2..3, 19..22 delta_alpha 2..5; NULL (-1): an index the waveform does not read */
static double *gimr_slot(gwo_params *p, int code)
{
    if (code < 0 || code > 22) return NULL;
    if (code < 14) return &p->gimr_dphi[code];
    if (code < 17) return &p->gimr_dsigma[code - 14];
    if (code < 19) return &p->gimr_dbeta[code - 17];
    return &p->gimr_dalpha[code - 19];
}

/* the ppE methods append betappe[0..Nmod) to the GR basis (fisher.cpp:2083-2088, :2424-2429); the gIMR methods
 * append delta_phi, delta_sigma, delta_beta, delta_alpha in list order (:2128-2158, :2473-2505). The deltas are
 * held here as tables indexed by PN order / coefficient; the waveform reads the first list entry of an index
 * (check_list_id), so a repeated entry writes the table only through its first occurrence. */
int gwo_unpack_parameters(double *v, int *lf, const gwo_params *p, const char *m, int dim)
{
    if (!unpack_gr(v, lf, p, m, dim)) return 0;
    if (has(m, "gIMR")) {
        gwo_params q = *p;
        const int base = dim - p->gimr_nlist;
        for (int i = 0; i < p->gimr_nlist; i++) {
            const double *s = gimr_slot(&q, p->gimr_list[i]);
            v[base + i] = s ? *s : 0.0;
        }
    } else if (has(m, "ppE")) {
        const int base = dim - p->ppe_nmod;
        for (int i = 0; i < p->ppe_nmod; i++) v[base + i] = p->ppe_beta[i];
    }
    return 1;
}

int gwo_repack_parameters(const double *v, gwo_params *a, const char *m, int dim, const gwo_params *orig)
{
    (void)orig;
    if (!repack_gr(v, a, m)) return 0;
    if (has(m, "gIMR")) {
        const int base = dim - a->gimr_nlist;
        for (int i = 0; i < a->gimr_nlist; i++) {
            int first = 1;
            for (int j = 0; j < i; j++)
                if (a->gimr_list[j] == a->gimr_list[i]) first = 0;
            double *s = gimr_slot(a, a->gimr_list[i]);
            if (first && s) *s = v[base + i];
        }
    } else if (has(m, "ppE")) {
        const int base = dim - a->ppe_nmod;
        for (int i = 0; i < a->ppe_nmod; i++) a->ppe_beta[i] = v[base + i];
    }
    return 1;
}

static int unpack_gr(double *v, int *lf, const gwo_params *p, const char *m, int dim)
{
    for (int i = 0; i < dim; i++) lf[i] = 0;
    int mcmc = has(m, "MCMC");
    if (!p->sky_average) {
        if (has(m, "IMRPhenomPv2")) {
            v[0] = p->RA;
            v[1] = mcmc ? sin(p->DEC) : p->DEC;
            v[2] = p->psi;
            v[3] = mcmc ? cos(p->incl) : p->incl;
            v[4] = p->phiRef;
            v[5] = p->tc;
            if (mcmc) {
                v[6] = log(p->DL);
                v[7] = log(gwo_calculate_chirpmass(p->mass1, p->mass2));
                v[8] = gwo_calculate_eta(p->mass1, p->mass2);
                double s1[3], s2[3];
                cart_sph(p->spin1, s1);
                cart_sph(p->spin2, s2);
                v[9] = s1[0];
                v[10] = s2[0];
                v[11] = cos(s1[1]);
                v[12] = cos(s2[1]);
                v[13] = s1[2];
                v[14] = s2[2];
            } else {
                lf[6] = lf[7] = 1;
                v[6] = p->DL;
                v[7] = gwo_calculate_chirpmass(p->mass1, p->mass2);
                v[8] = gwo_calculate_eta(p->mass1, p->mass2);
                v[9] = p->spin1[2];
                v[10] = p->spin2[2];
                v[11] = p->chip;
                v[12] = p->phip;
            }
            return 1;
        }
        if (has(m, "IMRPhenomD")) {
            v[0] = p->RA;
            v[1] = mcmc ? sin(p->DEC) : p->DEC;
            v[2] = p->psi;
            v[3] = mcmc ? cos(p->incl) : p->incl;
            v[4] = p->phiRef;
            v[5] = p->tc;
            if (mcmc) {
                v[6] = log(p->DL);
                v[7] = log(gwo_calculate_chirpmass(p->mass1, p->mass2));
            } else {
                lf[6] = lf[7] = 1;
                v[6] = p->DL;
                v[7] = gwo_calculate_chirpmass(p->mass1, p->mass2);
            }
            v[8] = gwo_calculate_eta(p->mass1, p->mass2);
            v[9] = p->spin1[2];
            v[10] = p->spin2[2];
            return 1;
        }
        return 0;
    }
    if (has(m, "IMRPhenomPv2")) return 0; /* "Sky averaged IMRPhenomP is not supported" */
    if (has(m, "IMRPhenomD")) {
        if (mcmc) {
            v[0] = log(gwo_calculate_chirpmass(p->mass1, p->mass2));
            v[1] = gwo_calculate_eta(p->mass1, p->mass2);
            v[2] = p->spin1[2];
            v[3] = p->spin2[2];
        } else {
            lf[0] = lf[3] = lf[4] = 1;
            v[3] = gwo_calculate_chirpmass(p->mass1, p->mass2);
            v[0] = gwo_A0_from_DL(v[3] * GWO_MSOL_SEC, p->DL * gwo_mpc_sec(), p->sky_average);
            v[1] = p->phiRef;
            v[2] = p->tc;
            v[4] = gwo_calculate_eta(p->mass1, p->mass2);
            v[5] = (p->spin1[2] + p->spin2[2]) / 2.;
            v[6] = (p->spin1[2] - p->spin2[2]) / 2.;
        }
        return 1;
    }
    return 0;
}

/* fisher.cpp:2167-2511 */
static int repack_gr(const double *v, gwo_params *a, const char *m)
{
    int mcmc = has(m, "MCMC");
    if (!a->sky_average) {
        if (has(m, "IMRPhenomPv2")) {
            if (mcmc) {
                a->mass1 = gwo_calculate_mass1(exp(v[7]), v[8]);
                a->mass2 = gwo_calculate_mass2(exp(v[7]), v[8]);
                a->DL = exp(v[6]);
                a->RA = v[0];
                a->DEC = asin(v[1]);
                a->psi = v[2];
                a->incl = acos(v[3]);
                a->phiRef = v[4];
                double t1 = v[11] > 1 ? 0 : (v[11] < -1 ? M_PI : acos(v[11]));
                double t2 = v[12] > 1 ? 0 : (v[12] < -1 ? M_PI : acos(v[12]));
                a->tc = v[5];
                double s1[3] = {v[9], t1, v[13]}, s2[3] = {v[10], t2, v[14]};
                sph_cart(s1, a->spin1);
                sph_cart(s2, a->spin2);
            } else {
                a->mass1 = gwo_calculate_mass1(v[7], v[8]);
                a->mass2 = gwo_calculate_mass2(v[7], v[8]);
                a->DL = v[6];
                a->RA = v[0];
                a->DEC = v[1];
                a->psi = v[2];
                a->incl = v[3];
                a->phiRef = v[4];
                a->tc = v[5];
                a->spin1[2] = v[9];
                a->spin2[2] = v[10];
                a->chip = v[11];
                a->phip = v[12];
            }
            return 1;
        }
        if (has(m, "IMRPhenomD")) {
            if (mcmc) {
                a->mass1 = gwo_calculate_mass1(exp(v[7]), v[8]);
                a->mass2 = gwo_calculate_mass2(exp(v[7]), v[8]);
                a->DL = exp(v[6]);
                a->RA = v[0];
                a->DEC = asin(v[1]);
                a->psi = v[2];
                a->incl = acos(v[3]);
            } else {
                a->mass1 = gwo_calculate_mass1(v[7], v[8]);
                a->mass2 = gwo_calculate_mass2(v[7], v[8]);
                a->DL = v[6];
                a->RA = v[0];
                a->DEC = v[1];
                a->psi = v[2];
                a->incl = v[3];
            }
            a->spin1[2] = v[9];
            a->spin2[2] = v[10];
            a->phiRef = v[4];
            a->tc = v[5];
            return 1;
        }
        return 0;
    }
    if (has(m, "IMRPhenomD")) {
        if (mcmc) {
            a->mass1 = gwo_calculate_mass1(exp(v[0]), v[1]);
            a->mass2 = gwo_calculate_mass2(exp(v[0]), v[1]);
            a->DL = 1000;
            a->spin1[2] = v[2];
            a->spin2[2] = v[3];
            a->phiRef = 0;
            a->tc = 0;
            a->incl = 0;
        } else {
            a->mass1 = gwo_calculate_mass1(v[3], v[4]);
            a->mass2 = gwo_calculate_mass2(v[3], v[4]);
            a->DL = gwo_DL_from_A0(v[3] * GWO_MSOL_SEC, v[0], a->sky_average) / gwo_mpc_sec();
            a->tc = v[2];
            a->phiRef = v[1];
            a->spin1[2] = v[5] + v[6];
            a->spin2[2] = v[5] - v[6];
        }
        return 1;
    }
    return 0;
}

/* fisher.cpp:2513-2572: a fresh gen_params carrying only the non-parameter options */
static void repack_non_parameter_options(gwo_params *w, const gwo_params *in)
{
    gwo_default_params(w);
    w->sky_average = in->sky_average;
    w->f_ref = in->f_ref;
    w->gmst = in->gmst;
    w->shift_time = 0;
    w->shift_phase = in->shift_phase;
    w->dep_postmerger = in->dep_postmerger;
    /* fisher.cpp:2541-2570: the ppE terms and the gIMR index lists travel with the parameters */
    w->ppe_nmod = in->ppe_nmod;
    w->ppe_theory = in->ppe_theory; /* fisher.cpp:2540-2545: check_theory_support keeps the theory's terms */
    w->cosmology = in->cosmology;
    for (int i = 0; i < 4; i++) {
        w->ppe_b[i] = in->ppe_b[i];
        w->ppe_beta[i] = in->ppe_beta[i];
    }
    memcpy(w->gimr_dphi, in->gimr_dphi, sizeof(w->gimr_dphi));
    memcpy(w->gimr_dsigma, in->gimr_dsigma, sizeof(w->gimr_dsigma));
    memcpy(w->gimr_dbeta, in->gimr_dbeta, sizeof(w->gimr_dbeta));
    memcpy(w->gimr_dalpha, in->gimr_dalpha, sizeof(w->gimr_dalpha));
    w->gimr_nlist = in->gimr_nlist;
    memcpy(w->gimr_list, in->gimr_list, sizeof(w->gimr_list));
}

static void add_tof_phase(double *r, const double *f, int n, double dt_phase)
{
    for (int l = 0; l < n; l++) {
        double ph = dt_phase * f[l];
        double rr = r[2 * l], ri = r[2 * l + 1];
        gwo_cmul(rr, ri, cos(ph), sin(ph), &r[2 * l], &r[2 * l + 1]);
    }
}

/* Finite-difference step of one Fisher-basis coordinate: step > 0 is the reference's absolute
 * epsilon (fisher.cpp:158); step < 0 asks for a step relative to the coordinate, |step| * |v|, or
 * |step| where v is exactly zero (tc and phiref at 0). The relative form is the benchmark
 * configuration's "1e-7 relative" step, not a reference behaviour. */
static double param_step(double step, double v)
{
    if (step > 0) return step;
    return v != 0 ? -step * fabs(v) : -step;
}

/* fisher.cpp:148-557 */
int gwo_calculate_derivatives(double *deriv, const double *f, int n, int dim, int det, int ref_det,
                              const char *gen_method, const gwo_params *p, int order, double step)
{
    char lm[128];
    local_method(gen_method, lm);
    int model = model_of(lm);
    double vec[32], pp[32], pm[32], ppp[32], pmm[32];
    int lf[32];
    if (!gwo_unpack_parameters(vec, lf, p, gen_method, dim)) return 0;
    gwo_params wp;
    repack_non_parameter_options(&wp, p);

    if (p->sky_average && has(lm, "IMRPhenomD")) {
        double *amp = (double *)malloc(sizeof(double) * n);
        double *ap = (double *)malloc(sizeof(double) * n), *am = (double *)malloc(sizeof(double) * n);
        double *php = (double *)malloc(sizeof(double) * n), *phm = (double *)malloc(sizeof(double) * n);
        double *app = (double *)malloc(sizeof(double) * n), *amm = (double *)malloc(sizeof(double) * n);
        double *phpp = (double *)malloc(sizeof(double) * n), *phmm = (double *)malloc(sizeof(double) * n);
        double *tmp = (double *)malloc(sizeof(double) * n);
        gwo_fourier_amplitude(f, n, amp, model, p);
        for (int i = 0; i < dim; i++) {
            double epsilon = param_step(step, vec[i]);
            memcpy(pp, vec, sizeof(double) * dim);
            memcpy(pm, vec, sizeof(double) * dim);
            pp[i] = vec[i] + epsilon;
            pm[i] = vec[i] - epsilon;
            int edge = (i == 8 && vec[i] > .25 - epsilon);
            if (edge) pp[i] = vec[i];
            gwo_repack_parameters(pp, &wp, gen_method, dim, p);
            gwo_fourier_amplitude(f, n, ap, model, &wp);
            gwo_fourier_phase(f, n, php, tmp, model, &wp);
            gwo_repack_parameters(pm, &wp, gen_method, dim, p);
            gwo_fourier_amplitude(f, n, am, model, &wp);
            gwo_fourier_phase(f, n, phm, tmp, model, &wp);
            if (order >= 4) {
                memcpy(ppp, vec, sizeof(double) * dim);
                memcpy(pmm, vec, sizeof(double) * dim);
                ppp[i] = vec[i] + 2 * epsilon;
                pmm[i] = vec[i] - 2 * epsilon;
                if (edge) ppp[i] = vec[i];
                gwo_repack_parameters(ppp, &wp, gen_method, dim, p);
                gwo_fourier_amplitude(f, n, app, model, &wp);
                gwo_fourier_phase(f, n, phpp, tmp, model, &wp);
                gwo_repack_parameters(pmm, &wp, gen_method, dim, p);
                gwo_fourier_amplitude(f, n, amm, model, &wp);
                gwo_fourier_phase(f, n, phmm, tmp, model, &wp);
            }
            double *d = deriv + (size_t)2 * n * i;
            for (int l = 0; l < n; l++) {
                double ad, pd;
                if (order == 2) {
                    double den = edge ? epsilon : 2 * epsilon;
                    ad = (ap[l] - am[l]) / den;
                    pd = (php[l] - phm[l]) / den;
                    d[2 * l] = ad;
                    d[2 * l + 1] = pd * amp[l];
                } else {
                    double den = edge ? 6. * epsilon : 12. * epsilon;
                    ad = (-app[l] + 8. * ap[l] - 8. * am[l] + amm[l]) / den;
                    pd = (-phpp[l] + 8. * php[l] - 8. * phm[l] + phmm[l]) / den;
                    d[2 * l] = ad;
                    d[2 * l + 1] = -pd * amp[l];
                }
            }
        }
        free(amp); free(ap); free(am); free(php); free(phm); free(app); free(amm); free(phpp); free(phmm); free(tmp);
    } else {
        double *rp = (double *)malloc(sizeof(double) * 2 * n), *rm = (double *)malloc(sizeof(double) * 2 * n);
        double *rpp = (double *)malloc(sizeof(double) * 2 * n), *rmm = (double *)malloc(sizeof(double) * 2 * n);
        for (int i = 0; i < dim; i++) {
            double epsilon = param_step(step, vec[i]);
            memcpy(pp, vec, sizeof(double) * dim);
            memcpy(pm, vec, sizeof(double) * dim);
            pp[i] = vec[i] + epsilon;
            pm[i] = vec[i] - epsilon;
            int edge = (i == 8 && vec[i] > .25 - epsilon);
            if (edge) pp[i] = vec[i];
            gwo_repack_parameters(pp, &wp, gen_method, dim, p);
            gwo_detector_response(f, n, rp, det, model, &wp);
            if (ref_det != det) add_tof_phase(rp, f, n, -2 * M_PI * gwo_dtoa(wp.RA, wp.DEC, wp.gmst, ref_det, det));
            gwo_repack_parameters(pm, &wp, gen_method, dim, p);
            gwo_detector_response(f, n, rm, det, model, &wp);
            if (ref_det != det) add_tof_phase(rm, f, n, -2 * M_PI * gwo_dtoa(wp.RA, wp.DEC, wp.gmst, ref_det, det));
            if (order >= 4) {
                /* the +-2eps evaluations shift tc instead (fisher.cpp:433-453) */
                memcpy(ppp, vec, sizeof(double) * dim);
                memcpy(pmm, vec, sizeof(double) * dim);
                ppp[i] = vec[i] + 2 * epsilon;
                pmm[i] = vec[i] - 2 * epsilon;
                if (edge) ppp[i] = vec[i];
                gwo_repack_parameters(ppp, &wp, gen_method, dim, p);
                if (ref_det != det) wp.tc -= gwo_dtoa(wp.RA, wp.DEC, wp.gmst, ref_det, det);
                gwo_detector_response(f, n, rpp, det, model, &wp);
                gwo_repack_parameters(pmm, &wp, gen_method, dim, p);
                if (ref_det != det) wp.tc -= gwo_dtoa(wp.RA, wp.DEC, wp.gmst, ref_det, det);
                gwo_detector_response(f, n, rmm, det, model, &wp);
            }
            double *d = deriv + (size_t)2 * n * i;
            for (int l = 0; l < 2 * n; l++) {
                if (order == 2) {
                    d[l] = edge ? (rp[l] - rm[l]) / epsilon : (rp[l] - rm[l]) / (2. * epsilon);
                } else {
                    double num = (-rpp[l] + 8. * rp[l] - 8. * rm[l] + rmm[l]);
                    d[l] = edge ? num / (6. * epsilon) : num / (12. * epsilon);
                }
            }
        }
        free(rp); free(rm); free(rpp); free(rmm);
    }
    for (int l = 0; l < dim; l++) {
        if (!lf[l]) continue;
        double *d = deriv + (size_t)2 * n * l;
        for (int j = 0; j < 2 * n; j++) d[j] *= vec[l];
    }
    return 1;
}

/* fisher.cpp:2704-2781 (SIMPSONS, df = f[1]-f[0]) */
static void fisher_elements(const double *f, int n, int dim, const double *deriv, double *out, const double *psd)
{
    double *integrand = (double *)malloc(sizeof(double) * n);
    double df = f[1] - f[0];
    for (int j = 0; j < dim; j++) {
        for (int k = 0; k <= j; k++) {
            const double *a = deriv + (size_t)2 * n * j, *b = deriv + (size_t)2 * n * k;
            for (int i = 0; i < n; i++) integrand[i] = (a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1]) / psd[i];
            out[j * dim + k] = 4 * gwo_simpsons_sum(df, n, integrand);
            out[k * dim + j] = out[j * dim + k];
        }
    }
    free(integrand);
}

/* calculate_fisher_elements (fisher.cpp:2704-2781) with either rule: Simpson at f[1]-f[0], or the caller's
 * Gauss-Legendre weights (times f ln 10 on a log10 grid); deriv as calculate_derivatives writes it */
void gwo_fisher_elements(const double *f, int n, int dim, const double *deriv, const double *psd, int gaussleg,
                         const double *weights, int log10f, double *out)
{
    if (!gaussleg && !log10f) {
        fisher_elements(f, n, dim, deriv, out, psd);
        return;
    }
    double *integrand = (double *)malloc(sizeof(double) * n);
    for (int j = 0; j < dim; j++) {
        for (int k = 0; k <= j; k++) {
            const double *a = deriv + (size_t)2 * n * j, *b = deriv + (size_t)2 * n * k;
            for (int i = 0; i < n; i++) {
                integrand[i] = (a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1]) / psd[i];
                if (log10f) integrand[i] *= f[i] * GWO_LOG10;
            }
            double v;
            if (gaussleg) {
                double sum = 0;
                for (int i = 0; i < n; i++) sum += weights[i] * integrand[i];
                v = 4 * sum;
            } else {
                v = 4 * gwo_simpsons_sum(f[1] - f[0], n, integrand);
            }
            out[j * dim + k] = v;
            out[k * dim + j] = v;
        }
    }
    free(integrand);
}

/* fisher.cpp:81-145 */
int gwo_fisher_numerical(const double *f, int n, const char *method, int det, int ref_det, double *out, int dim,
                         const gwo_params *p, int order, const double *psd, double epsilon)
{
    double *deriv = (double *)malloc(sizeof(double) * 2 * (size_t)n * dim);
    int ok = gwo_calculate_derivatives(deriv, f, n, dim, det, ref_det, method, p, order, epsilon);
    if (ok) fisher_elements(f, n, dim, deriv, out, psd);
    free(deriv);
    return ok;
}

/* util.cpp:828-855: inverse through LU decomposition with partial pivoting */
int gwo_lu_invert(const double *in, double *out, int dim)
{
    double *a = (double *)malloc(sizeof(double) * dim * dim);
    int *perm = (int *)malloc(sizeof(int) * dim);
    memcpy(a, in, sizeof(double) * dim * dim);
    for (int i = 0; i < dim; i++) perm[i] = i;
    for (int k = 0; k < dim; k++) {
        int piv = k;
        double mx = fabs(a[k * dim + k]);
        for (int i = k + 1; i < dim; i++)
            if (fabs(a[i * dim + k]) > mx) {
                mx = fabs(a[i * dim + k]);
                piv = i;
            }
        if (mx == 0) {
            free(a);
            free(perm);
            return 0;
        }
        if (piv != k) {
            for (int j = 0; j < dim; j++) {
                double t = a[k * dim + j];
                a[k * dim + j] = a[piv * dim + j];
                a[piv * dim + j] = t;
            }
            int t = perm[k];
            perm[k] = perm[piv];
            perm[piv] = t;
        }
        for (int i = k + 1; i < dim; i++) {
            a[i * dim + k] /= a[k * dim + k];
            for (int j = k + 1; j < dim; j++) a[i * dim + j] -= a[i * dim + k] * a[k * dim + j];
        }
    }
    for (int c = 0; c < dim; c++) {
        double *x = (double *)malloc(sizeof(double) * dim);
        for (int i = 0; i < dim; i++) {
            double s = (perm[i] == c) ? 1.0 : 0.0;
            for (int j = 0; j < i; j++) s -= a[i * dim + j] * x[j];
            x[i] = s;
        }
        for (int i = dim - 1; i >= 0; i--) {
            double s = x[i];
            for (int j = i + 1; j < dim; j++) s -= a[i * dim + j] * x[j];
            x[i] = s / a[i * dim + i];
        }
        for (int i = 0; i < dim; i++) out[i * dim + c] = x[i];
        free(x);
    }
    free(a);
    free(perm);
    return 1;
}
