/*
 * gwat_hip_compat.hpp -- the reference's C++ entry points of the hot path, header-only, over the C ABI
 * of gwat_hip.h.
 *
 * The functions take the reference's own parameter object (any type with the field names of
 * gen_params_base<double>, include/gwat/util.h:121-389) and the reference's string arguments, so a
 * translation unit built inside the reference tree can define the reference symbols with one line each
 * (see the interposition unit in the integration notes):
 *
 *     template<> int fourier_detector_response<double>(double *f, int n, std::complex<double> *r,
 *         std::string det, std::string method, gen_params_base<double> *p, double *times)
 *     { return gwh::fourier_detector_response(f, n, r, det, method, p); }
 *
 * Return values follow the reference: 1 on success, and NaN log-likelihoods on failure (the reason is
 * in gwh_last_error()); nothing here exits the process.
 */
#ifndef GWAT_HIP_COMPAT_HPP
#define GWAT_HIP_COMPAT_HPP

#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "gwat_hip.h"

namespace gwh {

/* "IMRPhenomD" / "IMRPhenomPv2", their ppE ("ppE_IMRPhenomD_Inspiral", "ppE_IMRPhenomD_IMR",
 * "ppE_IMRPhenomPv2_Inspiral", "ppE_IMRPhenomPv2_IMR") and gIMR ("gIMRPhenomD", "gIMRPhenomPv2") variants
 * (waveform_generator.cpp:103-294 dispatch); -1 otherwise */
/* A theory method string (check_theory_support, ppE_utilities.cpp:84-135) -> GWH_THEORY_*, and in *model the ppE
 * method assign_mapping (:158-360) maps it to: the _Inspiral variant, or for PNSeries / ppEAlt the variant the
 * string names, ModDispersion the _IMR one.  0: not a theory string; -1: a theory the engine does not map
 * (polarization_test, the NRT / EA variants, another waveform family). */
inline int theory_id(const std::string &m, int *model = nullptr)
{
    auto has = [&](const char *s) { return m.find(s) != std::string::npos; };
    int th;
    bool ins = true;
    if (has("dCS")) th = GWH_THEORY_DCS;
    else if (has("EdGB")) /* EdGB_HO keeps EdGB's terms: the reference's second if-chain overwrites them */
        th = has("EdGB_HO_LO") ? GWH_THEORY_EDGB_HO_LO
           : has("EdGB_GHOv1") ? GWH_THEORY_EDGB_GHOV1
           : has("EdGB_GHOv2") ? GWH_THEORY_EDGB_GHOV2
           : has("EdGB_GHOv3") ? GWH_THEORY_EDGB_GHOV3 : GWH_THEORY_EDGB;
    else if (has("ExtraDimension")) th = GWH_THEORY_EXTRADIMENSION;
    else if (has("BHEvaporation")) th = GWH_THEORY_BHEVAPORATION;
    else if (has("TVG")) th = GWH_THEORY_TVG;
    else if (has("DipRad")) th = GWH_THEORY_DIPRAD;
    else if (has("NonComm")) th = GWH_THEORY_NONCOMM;
    else if (has("ModDispersion")) th = GWH_THEORY_MODDISPERSION, ins = false;
    else if (has("polarization_test")) return -1;
    else if (has("PNSeries_ppE")) th = GWH_THEORY_PNSERIES, ins = has("Inspiral");
    else if (has("ppEAlt")) th = GWH_THEORY_PPEALT, ins = has("Inspiral");
    else return 0;
    if (has("NRT") || has("EA_")) return -1;
    const bool pv2 = has("IMRPhenomPv2");
    if (!pv2 && !has("IMRPhenomD")) return -1;
    if (model)
        *model = pv2 ? (ins ? GWH_PPE_IMRPHENOMPV2_INSPIRAL : GWH_PPE_IMRPHENOMPV2_IMR)
                     : (ins ? GWH_PPE_IMRPHENOMD_INSPIRAL : GWH_PPE_IMRPHENOMD_IMR);
    return th;
}

/* gen_params_base::cosmology (case-insensitive, util.cpp:356-382) -> GWH_COSMO_*; -1 if unknown */
inline int cosmology_id(const std::string &c)
{
    std::string u;
    for (char ch : c) u += (char)toupper((unsigned char)ch);
    const char *names[5] = {"PLANCK15", "PLANCK13", "WMAP9", "WMAP7", "WMAP5"};
    for (int i = 0; i < 5; i++)
        if (u == names[i]) return i;
    return -1;
}
template <class P>
inline auto cosmology_of(const P &g, int) -> decltype(std::string(g.cosmology), int())
{
    return cosmology_id(std::string(g.cosmology));
}
template <class P>
inline int cosmology_of(const P &, long)
{
    return GWH_COSMO_PLANCK15;
}

inline int model_id(const std::string &method)
{
    if (method == "IMRPhenomD") return GWH_IMRPHENOMD;
    if (method == "IMRPhenomPv2") return GWH_IMRPHENOMPV2;
    if (method == "ppE_IMRPhenomD_Inspiral") return GWH_PPE_IMRPHENOMD_INSPIRAL;
    if (method == "ppE_IMRPhenomD_IMR") return GWH_PPE_IMRPHENOMD_IMR;
    if (method == "gIMRPhenomD") return GWH_GIMRPHENOMD;
    if (method == "ppE_IMRPhenomPv2_Inspiral") return GWH_PPE_IMRPHENOMPV2_INSPIRAL;
    if (method == "ppE_IMRPhenomPv2_IMR") return GWH_PPE_IMRPHENOMPV2_IMR;
    if (method == "gIMRPhenomPv2") return GWH_GIMRPHENOMPV2;
    int m;
    if (theory_id(method, &m) > 0) return m;
    return -1;
}

/* gIMR index lists -> the index tables of gwh_params; the first entry of an index wins, as check_list_id
 * (util.cpp:810-817) finds it; indices outside the modified ranges are ignored, as the reference's loops do */
inline void gimr_table(int n, const int *idx, const double *delta, int lo, int hi, double *out)
{
    bool seen[16] = {false};
    for (int j = 0; j < n && idx && delta; j++) {
        const int i = idx[j];
        if (i < lo || i > hi || seen[i - lo]) continue;
        seen[i - lo] = true;
        out[i - lo] = delta[j];
    }
}

/* a gIMR index list appended to the Fisher-basis codes (an index outside the modified range is a parameter the
 * waveform does not read: code -1); the count keeps growing past GWH_GIMR_LIST_MAX so that the caller sees it */
inline void gimr_list(int n, const int *idx, int lo, int hi, int code0, gwh_params &p)
{
    for (int j = 0; j < n && idx; j++) {
        if (p.gimr_nlist < GWH_GIMR_LIST_MAX)
            p.gimr_list[p.gimr_nlist] = (int8_t)((idx[j] < lo || idx[j] > hi) ? -1 : code0 + idx[j]);
        p.gimr_nlist++;
    }
}

/* "SIMPSONS" / "GAUSSLEG" (mcmc_gw.cpp:811-856) */
inline int integration_id(const std::string &method) { return method == "GAUSSLEG" ? GWH_GAUSSLEG : GWH_SIMPSONS; }

/* the GR method under a ppE or gIMR method string: "ppE_IMRPhenomD_Inspiral" -> "IMRPhenomD",
 * "MCMC_ppE_IMRPhenomPv2_IMR" -> "MCMC_IMRPhenomPv2", "gIMRPhenomD" -> "IMRPhenomD" (fisher.cpp:2083-2088,
 * :2128-2158 append the betas / deltas to that basis) */
inline std::string fisher_gr_method(const std::string &method)
{
    std::string m = method;
    if (theory_id(m) > 0) /* a theory's parameters follow the basis of its waveform family */
        return std::string(m.compare(0, 5, "MCMC_") == 0 ? "MCMC_" : "") +
               (m.find("IMRPhenomPv2") != std::string::npos ? "IMRPhenomPv2" : "IMRPhenomD");
    const size_t g = m.find("gIMR");
    if (g != std::string::npos) return m.erase(g, 1);
    const size_t p = m.find("ppE_");
    if (p == std::string::npos) return m;
    m.erase(p, 4);
    for (const char *suf : {"_Inspiral", "_IMR"}) {
        const std::string s(suf);
        if (m.size() > s.size() && m.compare(m.size() - s.size(), s.size(), s) == 0) {
            m.erase(m.size() - s.size());
            break;
        }
    }
    return m;
}

/* Fisher / MCMC bases of fisher_numerical's method strings (fisher.cpp:1841-2161); a ppE method's betas and a
 * gIMR method's listed deltas are appended to the basis of its GR method */
inline int basis_id(const std::string &method_in, bool sky_average)
{
    const std::string method = fisher_gr_method(method_in);
    if (method == "IMRPhenomD") return sky_average ? GWH_BASIS_PHENOMD_SKYAVG : GWH_BASIS_PHENOMD_FULL;
    if (method == "MCMC_IMRPhenomD") return GWH_BASIS_PHENOMD_MCMC;
    if (method == "IMRPhenomPv2") return GWH_BASIS_PV2_FULL;
    if (method == "MCMC_IMRPhenomPv2") return GWH_BASIS_PV2_MCMC;
    return -1;
}

/* the hot-path fields of a gen_params_base<double>-shaped object for the given model; false if it asks
 * for physics the engine does not model (equatorial / horizon orientation, neutron-star flags, more ppE
 * terms than GWH_PPE_MAX).  As in prep_source_parameters (waveform_generator.cpp:1296-1313), the ppE
 * fields (Nmod, bppe, betappe) are read for the ppE models only and the gIMR lists for the gIMR models only.
 * theory (theory_id of the method string): the ppE fields hold that theory's parameters, mapped on the device. */
template <class Params>
inline bool to_params(const Params &g, gwh_params &p, int model = GWH_IMRPHENOMD, int theory = GWH_THEORY_NONE)
{
    p = gwh_params();
    p.mass1 = g.mass1;
    p.mass2 = g.mass2;
    for (int i = 0; i < 3; i++) {
        p.spin1[i] = g.spin1[i];
        p.spin2[i] = g.spin2[i];
    }
    p.DL = g.Luminosity_Distance;
    p.tc = g.tc;
    p.phiRef = g.phiRef;
    p.f_ref = g.f_ref;
    p.incl = g.incl_angle;
    p.RA = g.RA;
    p.DEC = g.DEC;
    p.psi = g.psi;
    p.gmst = g.gmst;
    p.chip = g.chip;
    p.phip = g.phip;
    p.shift_time = g.shift_time;
    p.shift_phase = g.shift_phase;
    p.sky_average = g.sky_average;
    p.dep_postmerger = g.dep_postmerger;
    if (model == GWH_PPE_IMRPHENOMD_INSPIRAL || model == GWH_PPE_IMRPHENOMD_IMR ||
        model == GWH_PPE_IMRPHENOMPV2_INSPIRAL || model == GWH_PPE_IMRPHENOMPV2_IMR) {
        if (g.Nmod < 0 || g.Nmod > GWH_PPE_MAX || (g.Nmod > 0 && ((!g.bppe && theory <= 0) || !g.betappe)))
            return false;
        p.ppe_nmod = g.Nmod;
        for (int i = 0; i < g.Nmod; i++) {
            p.ppe_b[i] = g.bppe ? g.bppe[i] : 0.0; /* the theories other than PNSeries / ppEAlt need no bppe */
            p.ppe_beta[i] = g.betappe[i];
        }
        p.ppe_theory = theory > 0 ? theory : GWH_THEORY_NONE;
        p.cosmology = cosmology_of(g, 0);
        if (theory < 0 || (theory > 0 && p.cosmology < 0)) return false;
    }
    if (model == GWH_GIMRPHENOMD || model == GWH_GIMRPHENOMPV2) {
        gimr_table(g.Nmod_phi, g.phii, g.delta_phi, -4, 9, p.gimr_dphi);
        gimr_table(g.Nmod_sigma, g.sigmai, g.delta_sigma, 2, 4, p.gimr_dsigma);
        gimr_table(g.Nmod_beta, g.betai, g.delta_beta, 2, 3, p.gimr_dbeta);
        gimr_table(g.Nmod_alpha, g.alphai, g.delta_alpha, 2, 5, p.gimr_dalpha);
        /* the lists in order, as the Fisher basis of the gIMR methods (fisher.cpp:2128-2158) */
        gimr_list(g.Nmod_phi, g.phii, -4, 9, GWH_GIMR_CODE_PHI(0), p);
        gimr_list(g.Nmod_sigma, g.sigmai, 2, 4, GWH_GIMR_CODE_SIGMA(0), p);
        gimr_list(g.Nmod_beta, g.betai, 2, 3, GWH_GIMR_CODE_BETA(0), p);
        gimr_list(g.Nmod_alpha, g.alphai, 2, 5, GWH_GIMR_CODE_ALPHA(0), p);
        if (p.gimr_nlist > GWH_GIMR_LIST_MAX) p.gimr_nlist = -1; /* more than the record holds: no Fisher basis */
    }
    return !(g.equatorial_orientation || g.horizon_coord || g.NSflag1 || g.NSflag2);
}

/* write the basis fields back into the caller's object (repack_parameters, fisher.cpp:2162-2511) */
template <class Params>
inline void from_params(const gwh_params &p, Params &g)
{
    g.mass1 = p.mass1;
    g.mass2 = p.mass2;
    for (int i = 0; i < 3; i++) {
        g.spin1[i] = p.spin1[i];
        g.spin2[i] = p.spin2[i];
    }
    g.Luminosity_Distance = p.DL;
    g.tc = p.tc;
    g.phiRef = p.phiRef;
    g.incl_angle = p.incl;
    g.RA = p.RA;
    g.DEC = p.DEC;
    g.psi = p.psi;
    g.chip = p.chip;
    g.phip = p.phip;
}

inline double nan_with(int code) { return code == GWH_OK ? 0.0 : std::nan(""); }

/* fourier_waveform (waveform_generator.h:43-49): h+ and hx into caller arrays of n bins */
template <class Params>
inline int fourier_waveform(const double *f, int n, std::complex<double> *hplus, std::complex<double> *hcross,
                            const std::string &method, const Params *g)
{
    gwh_params p;
    int model = model_id(method);
    if (model < 0 || !to_params(*g, p, model, theory_id(method))) return GWH_ERR_UNSUPPORTED;
    return gwh_fourier_waveform(f, n, reinterpret_cast<double *>(hplus), reinterpret_cast<double *>(hcross), model,
                                &p);
}

/* fourier_waveform (waveform_generator.h:70-74, waveform_generator.cpp:407-509): the bare (2,2) waveform
 * A e^{-i(Psi - tc (f - f_ref) - phic)} of IMRPhenomD and its two ppE variants, without polarisation or inclination
 * factors (the reference's overload reads no theory mapping and builds no other model; other method strings are
 * refused here, where the reference returns 1 and leaves the array unwritten) */
template <class Params>
inline int fourier_waveform(const double *f, int n, std::complex<double> *waveform, const std::string &method,
                            const Params *g)
{
    gwh_params p;
    int model = model_id(method);
    if ((model != GWH_IMRPHENOMD && model != GWH_PPE_IMRPHENOMD_INSPIRAL && model != GWH_PPE_IMRPHENOMD_IMR) ||
        !to_params(*g, p, model, GWH_THEORY_NONE))
        return GWH_ERR_UNSUPPORTED;
    p.incl = 0; /* h+ = h (1 + cos^2 0) / 2 = h, exactly */
    std::vector<double> hcross(2 * (size_t)(n > 0 ? n : 1));
    return gwh_fourier_waveform(f, n, reinterpret_cast<double *>(waveform), hcross.data(), model, &p);
}

/* fourier_detector_response (waveform_util.h:472-479), terrestrial detectors */
template <class Params>
inline int fourier_detector_response(const double *f, int n, std::complex<double> *response, const std::string &det,
                                     const std::string &method, const Params *g)
{
    gwh_params p;
    int model = model_id(method), d = gwh_detector_id(det.c_str());
    if (model < 0 || d < 0 || !to_params(*g, p, model, theory_id(method))) return GWH_ERR_UNSUPPORTED;
    return gwh_fourier_detector_response(f, n, reinterpret_cast<double *>(response), d, model, &p);
}

/* Log_Likelihood (mcmc_gw.h:210-217) */
template <class Params>
inline double Log_Likelihood(const std::complex<double> *data, const double *psd, const double *f, size_t n,
                             const Params *g, const std::string &det, const std::string &method)
{
    gwh_params p;
    int model = model_id(method), d = gwh_detector_id(det.c_str());
    if (model < 0 || d < 0 || !to_params(*g, p, model, theory_id(method))) return nan_with(GWH_ERR_UNSUPPORTED);
    return gwh_log_likelihood(reinterpret_cast<const double *>(data), psd, f, n, &p, d, model);
}

/* Log_Likelihood_internal (mcmc_gw.h:276-284) */
inline double Log_Likelihood_internal(const std::complex<double> *data, const double *psd, const double *f,
                                      const double *weights, const std::complex<double> *response, int n, bool log10F,
                                      const std::string &integration_method)
{
    return gwh_log_likelihood_internal(reinterpret_cast<const double *>(data), psd, f, weights,
                                       reinterpret_cast<const double *>(response), n, log10F,
                                       integration_id(integration_method));
}

/* MCMC_likelihood_extrinsic (mcmc_gw.h:477-490).  T_eff replaces the reference's pointer-difference
 * time offset (mcmc_gw.cpp:2466): the waveform uses t_c = T_eff - tc. */
template <class Params>
inline double MCMC_likelihood_extrinsic(bool save_waveform, const Params *g, const std::string &method,
                                        const int *data_length, double **frequencies,
                                        std::complex<double> **data, double **psd, double **weights,
                                        const std::string &integration_method, bool log10F,
                                        const std::string *detectors, int num_detectors, double T_eff = 0.0)
{
    gwh_params p;
    int model = model_id(method);
    if (model < 0 || !to_params(*g, p, model, theory_id(method)) || num_detectors < 1 || num_detectors > 4)
        return nan_with(GWH_ERR_UNSUPPORTED);
    int dets[4];
    std::vector<const double *> dp(num_detectors);
    for (int i = 0; i < num_detectors; i++) {
        dets[i] = gwh_detector_id(detectors[i].c_str());
        if (dets[i] < 0) return nan_with(GWH_ERR_UNSUPPORTED);
        dp[i] = reinterpret_cast<const double *>(data[i]);
    }
    return gwh_mcmc_likelihood_extrinsic(save_waveform, &p, model, data_length, frequencies, dp.data(), psd, weights,
                                         integration_id(integration_method), log10F, dets, num_detectors, T_eff);
}

/* fisher_numerical (fisher.h:35-48): output is the caller's dim x dim double** */
template <class Params>
inline void fisher_numerical(const double *f, int n, const std::string &method, const std::string &det,
                             const std::string &ref_det, double **output, int dim, const Params *g, int order,
                             const double *psd, double epsilon = 1e-8)
{
    gwh_params p;
    const std::string mm = method.compare(0, 5, "MCMC_") == 0 ? method.substr(5) : method;
    int basis = basis_id(method, g->sky_average), d = gwh_detector_id(det.c_str()),
        r = gwh_detector_id(ref_det.c_str()), model = model_id(mm), theory = theory_id(mm);
    std::vector<double> out((size_t)dim * dim, std::nan(""));
    if (basis >= 0 && model >= 0 && d >= 0 && r >= 0 && to_params(*g, p, model, theory))
        gwh_fisher_numerical_model(f, n, model, basis, d, r, out.data(), dim, &p, order, psd, epsilon);
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) output[i][j] = out[(size_t)i * dim + j];
}

/* repack_parameters (fisher.cpp:2162-2511): basis vector -> the caller's parameter object */
template <class Params>
inline void repack_parameters(const double *v, Params *g, const std::string &method, int dim)
{
    gwh_params p;
    int basis = basis_id(method, g->sky_average);
    to_params(*g, p);
    if (basis < 0 || gwh_repack_parameters(v, basis, &p) != GWH_OK) return;
    from_params(p, *g);
    if (method.find("ppE") != std::string::npos) /* the betas follow the GR basis (fisher.cpp:2424-2429) */
        for (int i = 0, base = dim - g->Nmod; i < g->Nmod; i++) g->betappe[i] = v[base + i];
    if (method.find("gIMR") != std::string::npos) { /* the deltas in list order (fisher.cpp:2473-2505) */
        int k = dim - (g->Nmod_phi + g->Nmod_sigma + g->Nmod_beta + g->Nmod_alpha);
        for (int i = 0; i < g->Nmod_phi; i++) g->delta_phi[i] = v[k++];
        for (int i = 0; i < g->Nmod_sigma; i++) g->delta_sigma[i] = v[k++];
        for (int i = 0; i < g->Nmod_beta; i++) g->delta_beta[i] = v[k++];
        for (int i = 0; i < g->Nmod_alpha; i++) g->delta_alpha[i] = v[k++];
    }
}

} // namespace gwh

#endif
