/* Memory and undefined-behaviour check of the oracle (TEST INFRASTRUCTURE): built with
 * -fsanitize=address,undefined by `make sanitize`, it runs every model, the likelihoods, the integration rules
 * and the numerical Fisher matrices of the oracle on small grids and checks that the results are finite.
 * Exit status 0 when everything ran clean; the sanitizers abort on the first finding. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../gwo.h"

static int bad = 0;

static void check(const char *what, const double *v, int n)
{
    for (int i = 0; i < n; i++)
        if (!isfinite(v[i])) {
            fprintf(stderr, "non-finite %s[%d]\n", what, i);
            bad = 1;
            return;
        }
}

static void base(gwo_params *p)
{
    gwo_default_params(p);
    p->mass1 = 36;
    p->mass2 = 29;
    p->spin1[2] = 0.3;
    p->spin2[2] = -0.1;
    p->DL = 410;
    p->f_ref = 20;
    p->incl = 2.9;
    p->RA = 1.37;
    p->DEC = -1.26;
    p->psi = 2.66;
    p->gmst = 0.7;
    p->tc = 0.01;
    p->phiRef = 0.3;
}

int main(void)
{
    enum { N = 801 };
    double f[N], psd[N], w[N];
    for (int i = 0; i < N; i++) {
        f[i] = 20 + 0.5 * i;
        double x = f[i] / 215.;
        psd[i] = 1e-49 * (pow(x, -4.14) - 5 / (x * x) + 111 * (1 - x * x + x * x * x * x / 2) / (1 + x * x / 2));
        w[i] = 0.5;
    }
    static double hp[2 * N], hc[2 * N], resp[3][2 * N], deriv[16 * 2 * N];
    const int dets[3] = {0, 1, 2};
    gwo_params p;
    const int models[] = {GWO_IMRPHENOMD, GWO_IMRPHENOMPV2, GWO_PPE_IMRPHENOMD_INSPIRAL, GWO_PPE_IMRPHENOMD_IMR,
                          GWO_GIMRPHENOMD, GWO_PPE_IMRPHENOMPV2_INSPIRAL, GWO_PPE_IMRPHENOMPV2_IMR,
                          GWO_GIMRPHENOMPV2};
    for (unsigned m = 0; m < sizeof(models) / sizeof(models[0]); m++) {
        const int model = models[m];
        base(&p);
        if (GWO_IS_PV2(model)) {
            p.spin1[0] = 0.3;
            p.spin2[1] = -0.2;
        }
        if (model == GWO_PPE_IMRPHENOMD_INSPIRAL || model == GWO_PPE_IMRPHENOMD_IMR ||
            model == GWO_PPE_IMRPHENOMPV2_INSPIRAL || model == GWO_PPE_IMRPHENOMPV2_IMR) {
            p.ppe_nmod = 2;
            p.ppe_b[0] = -7;
            p.ppe_b[1] = -3;
            p.ppe_beta[0] = 1e-4;
            p.ppe_beta[1] = -2e-3;
        }
        if (model == GWO_GIMRPHENOMD || model == GWO_GIMRPHENOMPV2) {
            for (int i = 0; i < 14; i++) p.gimr_dphi[i] = 0.01 * (i - 6);
            p.gimr_dsigma[1] = 0.05;
            p.gimr_dbeta[0] = -0.03;
            p.gimr_dalpha[2] = 0.02;
        }
        if (!gwo_fourier_waveform(f, N, hp, hc, model, &p)) bad = 1;
        check("hplus", hp, 2 * N);
        check("hcross", hc, 2 * N);
        for (int d = 0; d < 3; d++) {
            if (!gwo_detector_response(f, N, resp[d], dets[d], model, &p)) bad = 1;
            check("response", resp[d], 2 * N);
        }
        const double *dp[3] = {resp[0], resp[1], resp[2]}, *pp[3] = {psd, psd, psd};
        double ll[3] = {gwo_network_loglike(&p, model, f, N, dp, pp, dets, 3),
                        gwo_network_loglike_ex(&p, model, f, N, dp, pp, dets, 3, w, 1, 1),
                        gwo_mcmc_likelihood_extrinsic(&p, model, f, N, dp, pp, dets, 3, 1.0)};
        check("loglike", ll, 3);
        gwo_params batch[8];
        double out[8];
        for (int i = 0; i < 8; i++) {
            batch[i] = p;
            batch[i].mass1 += i;
        }
        gwo_network_loglike_batch(batch, 8, model, f, N, dp, pp, dets, 3, 2, out);
        check("batch loglike", out, 8);
    }
    /* theories on ppE terms */
    for (int th = 1; th < GWO_THEORY_COUNT; th++) {
        base(&p);
        p.ppe_theory = th;
        p.ppe_nmod = 1;
        p.ppe_beta[0] = 1e-22;
        p.ppe_b[0] = -7;
        if (th == GWO_THEORY_EDGB_GHOV1 || th == GWO_THEORY_EDGB_GHOV2 || th == GWO_THEORY_EDGB_GHOV3) {
            p.ppe_nmod = 2; /* two theory parameters */
            p.ppe_beta[1] = 1.0;
        }
        if (th == GWO_THEORY_MODDISPERSION) {
            p.ppe_beta[0] = 3e-15;
            p.ppe_b[0] = 4.5; /* a power with a distance fit */
        }
        gwo_fourier_waveform(f, N, hp, hc, GWO_PPE_IMRPHENOMD_INSPIRAL, &p);
        check("theory hplus", hp, 2 * N);
    }
    /* Fisher matrices: every basis, orders 2 and 4, relative step, derivatives, inversion */
    static const struct {
        const char *method;
        int dim, sky, pv2;
    } bases[] = {{"IMRPhenomD", 7, 1, 0}, {"IMRPhenomD", 11, 0, 0}, {"MCMC_IMRPhenomD", 11, 0, 0},
                 {"IMRPhenomPv2", 13, 0, 1}, {"MCMC_IMRPhenomPv2", 15, 0, 1}};
    for (unsigned b = 0; b < sizeof(bases) / sizeof(bases[0]); b++)
        for (int order = 2; order <= 4; order += 2) {
            base(&p);
            p.sky_average = bases[b].sky;
            if (bases[b].pv2) {
                p.spin1[0] = 0.3;
                p.spin2[1] = -0.2;
                if (bases[b].dim == 13) {
                    p.chip = 0.35;
                    p.phip = 0.8;
                }
            }
            const int dim = bases[b].dim;
            double F[256], Fi[256];
            const double eps = (order == 2) ? 1e-8 : -1e-7;
            if (!gwo_fisher_numerical(f, N, bases[b].method, bases[b].sky ? 0 : 1, 0, F, dim, &p, order, psd, eps)) {
                fprintf(stderr, "fisher refused: %s %d\n", bases[b].method, dim);
                bad = 1;
            }
            check("fisher", F, dim * dim);
            gwo_lu_invert(F, Fi, dim);
            if (!gwo_calculate_derivatives(deriv, f, N, dim, 2, 1, bases[b].method, &p, order, 1e-6)) bad = 1;
            check("derivatives", deriv, 2 * N * dim);
            double v[32];
            int lf[32];
            gwo_params q;
            gwo_unpack_parameters(v, lf, &p, bases[b].method, dim);
            gwo_repack_parameters(v, &q, bases[b].method, dim, &p);
        }
    double z[5];
    for (int c = 0; c < 5; c++) z[c] = gwo_z_from_dl(410, c);
    check("z", z, 5);
    double fe[256];
    gwo_fisher_elements(f, N, 7, deriv, psd, 1, w, 1, fe);
    check("fisher elements", fe, 49);
    if (bad) return 1;
    printf("oracle sanitizer run clean\n");
    return 0;
}
