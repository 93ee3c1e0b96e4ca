/*
 * Frozen FLOP-equivalent (fe) table of the hot path, counted on the oracle (TEST AND MEASUREMENT
 * INFRASTRUCTURE).  The reference algorithm as oracle/ restates it runs on the counting scalar of counted.h;
 * the per-bin work of a frequency region is the difference between waveform calls on n and n - 2 bins of that
 * region (per-call setup cancels), split into the amplitude part (gwo_build_amp: region of f against f1 =
 * 0.014/M and the amplitude peak f3), the phase part (gwo_build_phase and the powers of f the reference
 * precomputes for it: region of f against 0.018/M and f_RD / 2) and the rest (alignment, complex exponential,
 * polarisations), which the harness checks is the same in every region.  Weights: add/sub/mul/cmp 1,
 * div/sqrt 4, exp/log 16, pow 32, sin + cos pair 24, atan 20.  Prints one JSON object.
 */
#include "counted.h"
#define double cd
#include "../gwo_internal.h"
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

/* fe of a waveform call on n bins at frequencies f0, f0 + df, ... */
static double waveform_fe(const gwo_params &p, int model, double f0, double df, int n)
{
    std::vector<cd> f(n > 0 ? n : 1), hp(2 * (n > 0 ? n : 1)), hc(2 * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) f[i] = cd(f0 + i * df);
    OpTally before = g_ops;
    gwo_fourier_waveform(f.data(), n, hp.data(), hc.data(), model, &p);
    return fe_of(diff(g_ops, before));
}

static double per_bin(const gwo_params &p, int model, double lo, double hi)
{
    const int n = 9;
    double df = (hi - lo) / (n + 2), f0 = lo + 0.5 * df;
    return (waveform_fe(p, model, f0, df, n) - waveform_fe(p, model, f0, df, n - 2)) / 2;
}

/* fe of gwo_build_amp and gwo_build_phase at one frequency (the source prepared as the waveform call does) */
static void parts_fe(const gwo_params &p, double f, double *amp, double *phase)
{
    gwo_src s;
    gwo_populate_source(&s, &p, GWO_IMRPHENOMD);
    gwo_phenomd_state st;
    gwo_phenomd_prepare(&s, &st);
    gwo_pows pw;
    gwo_precalc_powers_pi(&pw);
    gwo_precalc_powers_ins(cd(f), s.M, &pw);
    OpTally b = g_ops;
    gwo_build_amp(cd(f), &s, &st, &pw);
    *amp = fe_of(diff(g_ops, b));
    b = g_ops;
    gwo_build_phase(cd(f), &s, &st, &pw);
    *phase = fe_of(diff(g_ops, b));
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
    const double f1 = summ[2].v, f3 = summ[3].v, f1p = summ[4].v, f2p = summ[5].v, fcut = 0.2 / M;
    /* the (amplitude region, phase region) pairs this source reaches, in ascending frequency */
    struct { int a, p; double lo, hi; } reg[5] = {
        {0, 0, 20., f1}, {1, 0, f1, f1p}, {1, 1, f1p, f2p}, {1, 2, f2p, f3}, {2, 2, f3, fcut}};
    double amp[3] = {0, 0, 0}, phase[3] = {0, 0, 0}, rest[5], total[5];
    for (int r = 0; r < 5; r++) {
        total[r] = per_bin(pd, GWO_IMRPHENOMD, reg[r].lo, reg[r].hi);
        double a, ph;
        parts_fe(pd, 0.5 * (reg[r].lo + reg[r].hi), &a, &ph);
        amp[reg[r].a] = a;
        phase[reg[r].p] = ph;
    }
    /* the rest: powers of f (which depend on the phase region) and the common per-bin work */
    double powers[3] = {0, 0, 0}, common = 0;
    for (int r = 0; r < 5; r++) rest[r] = total[r] - amp[reg[r].a] - phase[reg[r].p];
    /* regions 1 and 2 share the amplitude region and differ in the phase region; the inspiral phase region
     * alone evaluates the full set of powers */
    common = rest[2];
    powers[1] = powers[2] = 0;
    powers[0] = rest[0] - common;
    int additive = 1;
    for (int r = 0; r < 5; r++)
        if (rest[r] != common + powers[reg[r].p]) additive = 0;
    double pv2_add = per_bin(pp, GWO_IMRPHENOMPV2, f1p, f2p) - per_bin(pp, GWO_IMRPHENOMD, f1p, f2p);
    const int n = 9;
    double df = (f2p - f1p) / (n + 2), f0 = f1p + 0.5 * df;
    double det[2];
    for (int m = 0; m < 2; m++) {
        int model = m ? GWO_IMRPHENOMPV2 : GWO_IMRPHENOMD;
        const gwo_params &p = m ? pp : pd;
        double d3 = network_fe(p, model, f0, df, n, 3) - network_fe(p, model, f0, df, n - 2, 3);
        double d2 = network_fe(p, model, f0, df, n, 2) - network_fe(p, model, f0, df, n - 2, 2);
        det[m] = (d3 - d2) / 2;
    }
    printf("{\n \"source\": \"m1 = 36, m2 = 29 Msun, chi1z = 0.3, chi2z = -0.1 (in-plane 0.3, 0.1 / -0.1, 0.2 for "
           "IMRPhenomPv2); f1 = %.3f, f1_phase = %.3f, f2_phase = %.3f, f3 = %.3f, 0.2/M = %.3f Hz\",\n",
           f1, f1p, f2p, f3, fcut);
    printf(" \"weights\": \"add/sub/mul/cmp 1, div/sqrt 4, exp/log 16, pow 32, sin+cos pair 24, atan 20\",\n");
    printf(" \"amp_region_fe\": [%.1f, %.1f, %.1f],\n", amp[0], amp[1], amp[2]);
    printf(" \"phase_region_fe\": [%.1f, %.1f, %.1f],\n", phase[0] + powers[0], phase[1] + powers[1],
           phase[2] + powers[2]);
    printf(" \"common_per_bin_fe\": %.1f,\n", common);
    printf(" \"additive\": %s,\n", additive ? "true" : "false");
    printf(" \"measured_region_totals\": [%.1f, %.1f, %.1f, %.1f, %.1f],\n", total[0], total[1], total[2], total[3],
           total[4]);
    printf(" \"imrphenompv2_extra_per_bin_fe\": %.1f,\n", pv2_add);
    printf(" \"per_detector_per_bin_fe\": %.1f,\n", det[0]);
    printf(" \"imrphenompv2_per_detector_per_bin_fe\": %.1f\n}\n", det[1]);
    return additive ? 0 : 1;
}
