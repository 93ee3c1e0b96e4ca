/*
 * Frozen FLOP-equivalent (fe) table of the hot path, counted on the oracle (TEST AND MEASUREMENT
 * INFRASTRUCTURE): the reference algorithm as oracle/ restates it runs on the counting scalar of counted.h,
 * and the per-bin work of each frequency region is the difference between calls on n and n - 2 bins of that
 * region (per-call setup cancels).  Weights: add/sub/mul/cmp 1, div/sqrt 4, exp/log 16, pow 32, sin + cos
 * pair 24, atan 20.  Prints one JSON object.
 */
#include "counted.h"
#define double cd
#include "../gwo.h"
#undef double
#include <cstdio>
#include <vector>

OpTally g_ops;

static double fe_of(const OpTally &t)
{
    return 1.0 * (t.add + t.mul + t.cmp + t.misc) + 4.0 * (t.div + t.sqrt) + 16.0 * (t.exp + t.log) +
           32.0 * t.pow + 12.0 * t.sincos + 20.0 * t.atan;
}

static OpTally diff(const OpTally &a, const OpTally &b)
{
    OpTally d;
    d.add = a.add - b.add; d.mul = a.mul - b.mul; d.div = a.div - b.div; d.sqrt = a.sqrt - b.sqrt;
    d.exp = a.exp - b.exp; d.log = a.log - b.log; d.pow = a.pow - b.pow; d.sincos = a.sincos - b.sincos;
    d.atan = a.atan - b.atan; d.cmp = a.cmp - b.cmp; d.misc = a.misc - b.misc;
    return d;
}

static gwo_params injection(int pv2)
{
    gwo_params p;
    memset(&p, 0, sizeof(p));
    p.mass1 = 36.; p.mass2 = 29.;
    p.spin1[2] = 0.3; p.spin2[2] = -0.1;
    if (pv2) { p.spin1[0] = 0.3; p.spin1[1] = 0.1; p.spin2[0] = -0.1; p.spin2[1] = 0.2; }
    p.DL = 410.; p.tc = 0.; p.phiRef = 0.; p.f_ref = 20.; p.incl = 2.9;
    p.RA = 1.37; p.DEC = -1.26; p.psi = 2.66; p.gmst = 0.7;
    p.chip = -1.0; p.phip = -1.0;
    p.shift_time = 1; p.shift_phase = 1;
    return p;
}

static OpTally g_last; /* operations of the last measured call */

/* fe of a waveform call on n bins at frequencies f0, f0 + df, ... */
static double waveform_fe(const gwo_params &p, int model, double f0, double df, int n)
{
    std::vector<cd> f(n > 0 ? n : 1), hp(2 * (n > 0 ? n : 1)), hc(2 * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) f[i] = cd(f0 + i * df);
    OpTally before = g_ops;
    gwo_fourier_waveform(f.data(), n, hp.data(), hc.data(), model, &p);
    g_last = diff(g_ops, before);
    return fe_of(g_last);
}

static void print_ops(const OpTally &a, const OpTally &b, double per)
{
    OpTally d = diff(a, b);
    printf("{\"add\": %g, \"mul\": %g, \"div\": %g, \"sqrt\": %g, \"exp\": %g, \"log\": %g, \"pow\": %g, "
           "\"sin_or_cos\": %g, \"atan\": %g, \"cmp\": %g}",
           d.add / per, d.mul / per, d.div / per, d.sqrt / per, d.exp / per, d.log / per, d.pow / per, d.sincos / per,
           d.atan / per, (d.cmp + d.misc) / per);
}

static double network_fe(const gwo_params &p, int model, double f0, double df, int n, int ndet)
{
    std::vector<cd> f(n);
    std::vector<std::vector<cd>> data(ndet, std::vector<cd>(2 * n, cd(1e-23))), psd(ndet, std::vector<cd>(n, cd(1e-46)));
    std::vector<const cd *> dp(ndet), pp(ndet);
    for (int d = 0; d < ndet; d++) { dp[d] = data[d].data(); pp[d] = psd[d].data(); }
    for (int i = 0; i < n; i++) f[i] = cd(f0 + i * df);
    int dets[3] = {0, 1, 2};
    OpTally before = g_ops;
    gwo_network_loglike(&p, model, f.data(), n, dp.data(), pp.data(), dets, ndet);
    return fe_of(diff(g_ops, before));
}

int main()
{
    const double MSOL = 4.925491025543575903411922162094833998e-6;
    gwo_params pd = injection(0), pp = injection(1);
    const double M = (pd.mass1.v + pd.mass2.v) * MSOL;
    cd summ[32];
    gwo_phenomd_setup_summary(&pd, GWO_IMRPHENOMD, summ);
    /* fRD, fdamp, f1, f3, f1_phase, f2_phase */
    const double f1 = summ[2].v, f3 = summ[3].v, f1p = summ[4].v, f2p = summ[5].v, fcut = 0.2 / M;
    struct { const char *name; double lo, hi; } reg[5] = {
        {"amp_ins_phase_ins", 20., f1}, {"amp_int_phase_ins", f1, f1p}, {"amp_int_phase_int", f1p, f2p},
        {"amp_int_phase_mr", f2p, f3}, {"amp_mr_phase_mr", f3, fcut}};
    const int n = 9;
    printf("{\n \"workload\": \"m1 = 36, m2 = 29 Msun, chi1z = 0.3, chi2z = -0.1 (in-plane 0.3, 0.1 / -0.1, 0.2 for "
           "IMRPhenomPv2); regions f1 = %.3f, f1_phase = %.3f, f2_phase = %.3f, f3 = %.3f, 0.2/M = %.3f Hz\",\n",
           f1, f1p, f2p, f3, fcut);
    printf(" \"weights\": \"add/sub/mul/cmp 1, div/sqrt 4, exp/log 16, pow 32, sin+cos pair 24, atan 20\",\n");
    printf(" \"imrphenomd_waveform_per_bin\": {\n  ");
    double pv2_add = 0;
    for (int r = 0; r < 5; r++) {
        double df = (reg[r].hi - reg[r].lo) / (n + 2);
        double f0 = reg[r].lo + 0.5 * df;
        double a = waveform_fe(pd, GWO_IMRPHENOMD, f0, df, n);
        OpTally ta = g_last;
        a -= waveform_fe(pd, GWO_IMRPHENOMD, f0, df, n - 2);
        OpTally tb = g_last;
        double b = waveform_fe(pp, GWO_IMRPHENOMPV2, f0, df, n) - waveform_fe(pp, GWO_IMRPHENOMPV2, f0, df, n - 2);
        printf("%s\"%s\": {\"fe\": %.1f, \"ops\": ", r ? ", " : "", reg[r].name, a / 2);
        print_ops(ta, tb, 2.0);
        printf("}");
        if (r == 2) pv2_add = (b - a) / 2;
    }
    printf("},\n");
    {
        double df = (fcut * 1.5 - fcut * 1.1) / n, f0 = fcut * 1.1;
        double z = waveform_fe(pd, GWO_IMRPHENOMD, f0, df, n) - waveform_fe(pd, GWO_IMRPHENOMD, f0, df, n - 2);
        printf(" \"imrphenomd_waveform_per_zero_bin\": %.1f,\n", z / 2);
    }
    printf(" \"imrphenompv2_extra_per_bin\": %.1f,\n", pv2_add);
    /* per detector per non-zero bin: the network log-likelihood on 3 vs 2 detectors, n vs n - 2 bins */
    double df = (f2p - f1p) / (n + 2), f0 = f1p + 0.5 * df;
    for (int model : {GWO_IMRPHENOMD, GWO_IMRPHENOMPV2}) {
        const gwo_params &p = model == GWO_IMRPHENOMD ? pd : pp;
        double d3 = network_fe(p, model, f0, df, n, 3) - network_fe(p, model, f0, df, n - 2, 3);
        double d2 = network_fe(p, model, f0, df, n, 2) - network_fe(p, model, f0, df, n - 2, 2);
        printf(" \"%s_per_detector_per_bin\": %.1f,\n", model == GWO_IMRPHENOMD ? "imrphenomd" : "imrphenompv2",
               (d3 - d2) / 2);
    }
    printf(" \"note\": \"per-bin differences of calls on %d and %d bins of one region (setup cancels)\"\n}\n", n,
           n - 2);
    return 0;
}
