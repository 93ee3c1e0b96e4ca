/*
 * gwh_gwatpy.cpp -- the gwatpy C ABI of the hot path (include/gwat/gwatpy_wrapping.h) over the engine.
 *
 * gwatpy binds libgwat with ctypes and treats gen_params_base<double>* as an opaque handle made by
 * gen_params_base_py.  Exporting the same names with the same argument order lets those ctypes
 * declarations (gwatpy/gwatpy/waveform_generator.py, mcmc_routines.py) load libgwat_hip.so instead,
 * for the functions on the hot path: parameter handles, fourier_waveform_py,
 * fourier_detector_response_py, MCMC_likelihood_extrinsic_py, repack_parameters_py and the two
 * geometry helpers.  The handle here is the engine's own record; parameters the engine does not model
 * (equatorial or horizon orientation, neutron-star flags, more than GWH_PPE_MAX ppE terms) are recorded
 * and reported as GWH_ERR_UNSUPPORTED when the handle is used, never silently dropped.  ppE terms are
 * used by the ppE_IMRPhenomD methods and gIMR terms by gIMRPhenomD; like the reference, other methods
 * ignore them.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "gwat_hip.h"
#include "gwat_hip_compat.hpp"
#include "gwatpy_abi.h"

/* the engine's error recorder (gwh_engine.hip) */
int gwh_set_error(int code, const std::string &msg);

namespace {

const uint64_t kMagic = 0x6777617470793031ull; /* "gwatpy01" */

struct PyParams {
    uint64_t magic;
    gwh_params p;
    std::string unsupported; /* empty: every field is modelled */
};

/* live handles: a pointer is dereferenced only after it is found here */
std::mutex g_handles_mu;
std::unordered_set<void *> g_handles;

PyParams *handle(void *h)
{
    std::lock_guard<std::mutex> lk(g_handles_mu);
    if (!h || !g_handles.count(h)) return nullptr;
    PyParams *q = static_cast<PyParams *>(h);
    return q->magic == kMagic ? q : nullptr;
}

int model_of(const char *method)
{
    if (!method) return -1;
    if (!strcmp(method, "IMRPhenomD")) return GWH_IMRPHENOMD;
    if (!strcmp(method, "IMRPhenomPv2")) return GWH_IMRPHENOMPV2;
    if (!strcmp(method, "ppE_IMRPhenomD_Inspiral")) return GWH_PPE_IMRPHENOMD_INSPIRAL;
    if (!strcmp(method, "ppE_IMRPhenomD_IMR")) return GWH_PPE_IMRPHENOMD_IMR;
    if (!strcmp(method, "gIMRPhenomD")) return GWH_GIMRPHENOMD;
    if (!strcmp(method, "ppE_IMRPhenomPv2_Inspiral")) return GWH_PPE_IMRPHENOMPV2_INSPIRAL;
    if (!strcmp(method, "ppE_IMRPhenomPv2_IMR")) return GWH_PPE_IMRPHENOMPV2_IMR;
    if (!strcmp(method, "gIMRPhenomPv2")) return GWH_GIMRPHENOMPV2;
    int m;
    if (gwh::theory_id(method, &m) > 0) return m; /* a theory rides on the ppE method it maps to */
    return -1;
}

/* the handle's record as a method string reads it: a theory string maps the ppE fields as that theory's
 * parameters (prep_source_parameters, waveform_generator.cpp:1379-1404) */
gwh_params for_method(const gwh_params &p, const char *method)
{
    gwh_params r = p;
    const int th = method ? gwh::theory_id(method) : 0;
    r.ppe_theory = th > 0 ? th : GWH_THEORY_NONE;
    return r;
}

/* Fisher / MCMC basis of a generation method string and dimension (fisher.cpp:1841-2161) */
int basis_of(const char *method, int dim, int sky_average)
{
    if (!method) return -1;
    std::string m(method);
    if (m == "IMRPhenomD") return (sky_average && dim == 7) ? GWH_BASIS_PHENOMD_SKYAVG : (dim == 11 ? GWH_BASIS_PHENOMD_FULL : -1);
    if (m == "MCMC_IMRPhenomD") return dim == 11 ? GWH_BASIS_PHENOMD_MCMC : -1;
    if (m == "IMRPhenomPv2") return dim == 13 ? GWH_BASIS_PV2_FULL : -1;
    if (m == "MCMC_IMRPhenomPv2") return dim == 15 ? GWH_BASIS_PV2_MCMC : -1;
    return -1;
}

/* failures found here are recorded like the engine's own (gwh_last_error) */
int gwh_report(int code, const char *msg) { return gwh_set_error(code, msg); }

/* MCMC_modification_struct_py's handle: the modification lists a sampler's parameter vector carries */
const uint64_t kModMagic = 0x6777617470793032ull; /* "gwatpy02" */
struct PyMod {
    uint64_t magic;
    int ppe_nmod = 0;
    std::vector<double> bppe;
    std::vector<int> phii, sigmai, betai, alphai;
    bool ns = false;
};
std::unordered_set<void *> g_mods; /* guarded by g_handles_mu */

/* the gIMR index lists in basis order (fisher.cpp:2128-2158) as table codes; more than 16: no basis (-1) */
void fill_gimr_list(gwh_params &p, const std::vector<int> &phii, const std::vector<int> &sigmai,
                    const std::vector<int> &betai, const std::vector<int> &alphai)
{
    auto list = [&](const std::vector<int> &idx, int lo, int hi, int code0) {
        for (int i : idx) {
            if (p.gimr_nlist < GWH_GIMR_LIST_MAX) p.gimr_list[p.gimr_nlist] = (int8_t)((i < lo || i > hi) ? -1 : code0 + i);
            p.gimr_nlist++;
        }
    };
    list(phii, -4, 9, GWH_GIMR_CODE_PHI(0));
    list(sigmai, 2, 4, GWH_GIMR_CODE_SIGMA(0));
    list(betai, 2, 3, GWH_GIMR_CODE_BETA(0));
    list(alphai, 2, 5, GWH_GIMR_CODE_ALPHA(0));
    if (p.gimr_nlist > GWH_GIMR_LIST_MAX) p.gimr_nlist = -1;
}

/* the modification parameters after a GR basis block: the ppE betas, or the listed gIMR deltas (the first entry
 * of an index is the one the waveform reads, as check_list_id finds it) */
void set_mod_params(gwh_params &p, const double *v, int nmod, bool ppe, bool gimr)
{
    for (int i = 0; i < nmod && ppe; i++) p.ppe_beta[i] = v[i];
    for (int i = 0; i < nmod && gimr; i++) {
        const int c = p.gimr_list[i];
        bool first = c >= 0;
        for (int j = 0; j < i; j++) first &= (p.gimr_list[j] != c);
        if (!first) continue;
        if (c < GWH_GIMR_CODE_SIGMA(2)) p.gimr_dphi[c] = v[i];
        else if (c < GWH_GIMR_CODE_BETA(2)) p.gimr_dsigma[c - GWH_GIMR_CODE_SIGMA(2)] = v[i];
        else if (c < GWH_GIMR_CODE_ALPHA(2)) p.gimr_dbeta[c - GWH_GIMR_CODE_BETA(2)] = v[i];
        else p.gimr_dalpha[c - GWH_GIMR_CODE_ALPHA(2)] = v[i];
    }
}

/* the GR family of a method string: IMRPhenomD or IMRPhenomPv2 (ppE, gIMR and theory variants alike) */
std::string gr_family(const char *method)
{
    return (method && strstr(method, "IMRPhenomPv2")) ? "IMRPhenomPv2" : "IMRPhenomD";
}

} // namespace

extern "C" {

void *gen_params_base_py(double mass1, double mass2, double *spin1, double *spin2, double Luminosity_Distance,
                         double incl_angle, double RA, double DEC, double psi, double gmst, double tc, double phiRef,
                         double f_ref, double theta_l, double phi_l, double theta, double phi, char *cosmology,
                         bool equatorial_orientation, bool horizon_coord, bool NSflag1, bool NSflag2,
                         bool dep_postmerger, bool shift_time, bool shift_phase, bool sky_average, double LISA_alpha0,
                         double LISA_phi0, int Nmod_phi, int Nmod_sigma, int Nmod_beta, int Nmod_alpha, int *phii,
                         int *sigmai, int *betai, int *alphai, double *delta_phi, double *delta_sigma,
                         double *delta_beta, double *delta_alpha, int Nmod, double *bppe, double *betappe)
{
    (void)theta_l, (void)phi_l, (void)theta, (void)phi, (void)LISA_alpha0, (void)LISA_phi0;

    PyParams *q = new PyParams;
    q->magic = kMagic;
    gwh_params &p = q->p;
    memset(&p, 0, sizeof(p));
    p.mass1 = mass1;
    p.mass2 = mass2;
    for (int i = 0; i < 3; i++) {
        p.spin1[i] = spin1 ? spin1[i] : 0.0;
        p.spin2[i] = spin2 ? spin2[i] : 0.0;
    }
    p.DL = Luminosity_Distance;
    p.incl = incl_angle;
    p.RA = RA;
    p.DEC = DEC;
    p.psi = psi;
    p.gmst = gmst;
    p.tc = tc;
    p.phiRef = phiRef;
    p.f_ref = f_ref;
    p.chip = -1; /* gen_params_base defaults (util.h:231-236): spins from the vectors */
    p.phip = -1;
    p.dep_postmerger = dep_postmerger;
    p.shift_time = shift_time;
    p.shift_phase = shift_phase;
    p.sky_average = sky_average;
    if (equatorial_orientation) q->unsupported = "equatorial_orientation";
    else if (horizon_coord) q->unsupported = "horizon_coord";
    else if (NSflag1 || NSflag2) q->unsupported = "NSflag1/NSflag2";
    else if (Nmod > GWH_PPE_MAX)
        q->unsupported = "more than " + std::to_string(GWH_PPE_MAX) + " ppE terms";
    else if (Nmod > 0 && !betappe)
        q->unsupported = "ppE terms without a betappe array"; /* bppe may be absent: the theories fix it */
    /* ppE terms are read by the ppE methods only, gIMR terms by the gIMR methods only; the other methods ignore
     * them, as prep_source_parameters does (waveform_generator.cpp:1296-1313).  gIMR index lists become
     * index -> delta tables (the first entry of an index wins, as check_list_id finds it). */
    auto table = [](int n, const int *idx, const double *delta, int lo, int hi, double *out) {
        bool seen[16] = {false};
        for (int j = 0; j < n && idx && delta; j++) {
            const int i = idx[j];
            if (i < lo || i > hi || seen[i - lo]) continue;
            seen[i - lo] = true;
            out[i - lo] = delta[j];
        }
    };
    table(Nmod_phi, phii, delta_phi, -4, 9, p.gimr_dphi);
    table(Nmod_sigma, sigmai, delta_sigma, 2, 4, p.gimr_dsigma);
    table(Nmod_beta, betai, delta_beta, 2, 3, p.gimr_dbeta);
    table(Nmod_alpha, alphai, delta_alpha, 2, 5, p.gimr_dalpha);
    /* the lists in order, as the Fisher basis of the gIMR methods (fisher.cpp:2128-2158); an index outside the
     * modified range is a parameter the waveform does not read (code -1); more than 16 entries: no basis */
    auto list = [&](int n, const int *idx, int lo, int hi, int code0) {
        for (int j = 0; j < n && idx; j++) {
            if (p.gimr_nlist < GWH_GIMR_LIST_MAX)
                p.gimr_list[p.gimr_nlist] = (int8_t)((idx[j] < lo || idx[j] > hi) ? -1 : code0 + idx[j]);
            p.gimr_nlist++;
        }
    };
    list(Nmod_phi, phii, -4, 9, GWH_GIMR_CODE_PHI(0));
    list(Nmod_sigma, sigmai, 2, 4, GWH_GIMR_CODE_SIGMA(0));
    list(Nmod_beta, betai, 2, 3, GWH_GIMR_CODE_BETA(0));
    list(Nmod_alpha, alphai, 2, 5, GWH_GIMR_CODE_ALPHA(0));
    if (p.gimr_nlist > GWH_GIMR_LIST_MAX) p.gimr_nlist = -1;
    if (q->unsupported.empty() && Nmod > 0) {
        p.ppe_nmod = Nmod;
        for (int i = 0; i < Nmod; i++) {
            p.ppe_b[i] = bppe ? bppe[i] : 0.0; /* a theory's parameters come without exponents */
            p.ppe_beta[i] = betappe[i];
        }
    }
    /* the redshift of the theory maps (gen_params_base::cosmology); unknown names leave a theory unmappable */
    p.cosmology = cosmology ? gwh::cosmology_id(cosmology) : GWH_COSMO_PLANCK15;
    std::lock_guard<std::mutex> lk(g_handles_mu);
    g_handles.insert(q);
    return q;
}

void gen_params_base_py_destructor(void *p)
{
    std::lock_guard<std::mutex> lk(g_handles_mu);
    if (!p || !g_handles.erase(p)) return;
    PyParams *q = static_cast<PyParams *>(p);
    q->magic = 0;
    delete q;
}

static int usable(void *parameters, PyParams **out)
{
    PyParams *q = handle(parameters);
    if (!q) return gwh_report(GWH_ERR_INVALID, "not a parameter handle from gen_params_base_py");
    if (!q->unsupported.empty())
        return gwh_report(GWH_ERR_UNSUPPORTED, ("the engine does not model " + q->unsupported).c_str());
    *out = q;
    return GWH_OK;
}

int fourier_waveform_py(double *frequencies, int length, double *wf_plus_real, double *wf_plus_imaginary,
                        double *wf_cross_real, double *wf_cross_imaginary, char *generation_method, void *parameters)
{
    PyParams *q;
    int rc = usable(parameters, &q);
    if (rc != GWH_OK) return rc;
    int model = model_of(generation_method);
    if (model < 0) return gwh_report(GWH_ERR_UNSUPPORTED, "generation method not on the engine's path");
    if (length < 1 || !wf_plus_real || !wf_plus_imaginary || !wf_cross_real || !wf_cross_imaginary)
        return gwh_report(GWH_ERR_INVALID, "null output or empty grid");
    std::vector<double> hp(2 * (size_t)length), hc(2 * (size_t)length);
    const gwh_params pm = for_method(q->p, generation_method);
    rc = gwh_fourier_waveform(frequencies, length, hp.data(), hc.data(), model, &pm);
    if (rc != GWH_OK) return rc;
    for (int i = 0; i < length; i++) {
        wf_plus_real[i] = hp[2 * i];
        wf_plus_imaginary[i] = hp[2 * i + 1];
        wf_cross_real[i] = hc[2 * i];
        wf_cross_imaginary[i] = hc[2 * i + 1];
    }
    return rc;
}

int fourier_detector_response_py(double *frequencies, int length, double *response_real, double *response_imaginary,
                                 char *detector, char *generation_method, void *parameters)
{
    PyParams *q;
    int rc = usable(parameters, &q);
    if (rc != GWH_OK) return rc;
    int model = model_of(generation_method);
    if (model < 0) return gwh_report(GWH_ERR_UNSUPPORTED, "generation method not on the engine's path");
    int det = gwh_detector_id(detector);
    if (det < 0) return gwh_report(GWH_ERR_UNSUPPORTED, "unsupported detector");
    if (length < 1 || !response_real || !response_imaginary)
        return gwh_report(GWH_ERR_INVALID, "null output or empty grid");
    std::vector<double> r(2 * (size_t)length);
    const gwh_params pm = for_method(q->p, generation_method);
    rc = gwh_fourier_detector_response(frequencies, length, r.data(), det, model, &pm);
    if (rc != GWH_OK) return rc;
    for (int i = 0; i < length; i++) {
        response_real[i] = r[2 * i];
        response_imaginary[i] = r[2 * i + 1];
    }
    return rc;
}

/* gwatpy_wrapping.cpp:184-241.  Detector letters follow the wrapper's own matching: the defaults
 * Hanford, Livingston, Virgo, replaced only where the suffix starting at letter i is exactly one of
 * "H", "L", "V", "K", "C" (so in practice only the last letter is ever matched).  The wrapper's T is a
 * pointer difference (mcmc_gw.cpp:2466); the engine uses T_eff = 0, i.e. t_c enters as -tc. */
/* the body shared by the two wrappers: flattened arrays, detector letters, the engine's network likelihood */
static double extrinsic_flat(bool save_waveform, const gwh_params &params, char *generation_method, int *data_length,
                             double *frequencies, double *dataREAL, double *dataIMAG, double *psd, double *weights,
                             char *integration_method, bool log10F, char *detectors, int num_detectors)
{
    const double nan = std::nan("");
    int model = model_of(generation_method);
    if (model < 0) {
        gwh_report(GWH_ERR_UNSUPPORTED, "generation method not on the engine's path");
        return nan;
    }
    if (!data_length || !frequencies || !dataREAL || !dataIMAG || !psd || !detectors || num_detectors < 1 ||
        num_detectors > 4) {
        gwh_report(GWH_ERR_INVALID, "bad arguments");
        return nan;
    }
    int method = (integration_method && !strcmp(integration_method, "GAUSSLEG")) ? GWH_GAUSSLEG : GWH_SIMPSONS;
    if (method == GWH_GAUSSLEG && !weights) {
        gwh_report(GWH_ERR_INVALID, "GAUSSLEG needs weights");
        return nan;
    }
    static const char *defaults[3] = {"Hanford", "Livingston", "Virgo"};
    int dets[4];
    size_t nlet = strlen(detectors);
    for (int i = 0; i < num_detectors; i++) {
        const char *name = i < 3 ? defaults[i] : nullptr;
        if ((size_t)i < nlet) {
            const char *suffix = detectors + i;
            if (!strcmp(suffix, "H")) name = "Hanford";
            else if (!strcmp(suffix, "L")) name = "Livingston";
            else if (!strcmp(suffix, "V")) name = "Virgo";
            else if (!strcmp(suffix, "K")) name = "KAGRA"; /* not a detector_util name: rejected below */
            else if (!strcmp(suffix, "C")) name = "CE";
        }
        dets[i] = name ? gwh_detector_id(name) : -1;
        if (dets[i] < 0) {
            gwh_report(GWH_ERR_UNSUPPORTED, "unsupported detector");
            return nan;
        }
    }
    /* flattened [detector][bin] arrays at i * data_length[i] + j (gwatpy_wrapping.cpp:211-216) */
    std::vector<std::vector<double>> dat(num_detectors);
    std::vector<const double *> fp(num_detectors), dp(num_detectors), pp(num_detectors), wp(num_detectors);
    for (int i = 0; i < num_detectors; i++) {
        int n = data_length[i];
        if (n < 2) {
            gwh_report(GWH_ERR_INVALID, "need at least 2 bins");
            return nan;
        }
        size_t off = (size_t)i * n;
        dat[i].resize(2 * (size_t)n);
        for (int j = 0; j < n; j++) {
            dat[i][2 * j] = dataREAL[off + j];
            dat[i][2 * j + 1] = dataIMAG[off + j];
        }
        fp[i] = frequencies + off;
        dp[i] = dat[i].data();
        pp[i] = psd + off;
        wp[i] = weights ? weights + off : nullptr;
    }
    const gwh_params pm = for_method(params, generation_method);
    return gwh_mcmc_likelihood_extrinsic(save_waveform, &pm, model, data_length, fp.data(), dp.data(), pp.data(),
                                         weights ? wp.data() : nullptr, method, log10F, dets, num_detectors, 0.0);
}

double MCMC_likelihood_extrinsic_py(bool save_waveform, void *parameters, char *generation_method, int *data_length,
                                    double *frequencies, double *dataREAL, double *dataIMAG, double *psd,
                                    double *weights, char *integration_method, bool log10F, char *detectors,
                                    int num_detectors)
{
    PyParams *q;
    if (usable(parameters, &q) != GWH_OK) return std::nan("");
    return extrinsic_flat(save_waveform, q->p, generation_method, data_length, frequencies, dataREAL, dataIMAG, psd,
                          weights, integration_method, log10F, detectors, num_detectors);
}

void *MCMC_modification_struct_py(int ppE_Nmod, double *bppe, int gIMR_Nmod_phi, int *gIMR_phii, int gIMR_Nmod_sigma,
                                  int *gIMR_sigmai, int gIMR_Nmod_beta, int *gIMR_betai, int gIMR_Nmod_alpha,
                                  int *gIMR_alphai, bool NSflag1, bool NSflag2)
{
    PyMod *m = new PyMod;
    m->magic = kModMagic;
    m->ppe_nmod = ppE_Nmod;
    if (bppe && ppE_Nmod > 0) m->bppe.assign(bppe, bppe + ppE_Nmod);
    auto copy = [](int n, const int *v, std::vector<int> &out) {
        if (v && n > 0) out.assign(v, v + n);
    };
    copy(gIMR_Nmod_phi, gIMR_phii, m->phii);
    copy(gIMR_Nmod_sigma, gIMR_sigmai, m->sigmai);
    copy(gIMR_Nmod_beta, gIMR_betai, m->betai);
    copy(gIMR_Nmod_alpha, gIMR_alphai, m->alphai);
    m->ns = NSflag1 || NSflag2;
    std::lock_guard<std::mutex> lk(g_handles_mu);
    g_mods.insert(m);
    return m;
}

void MCMC_modification_struct_py_destructor(void *mod_struct)
{
    std::lock_guard<std::mutex> lk(g_handles_mu);
    if (!mod_struct || !g_mods.erase(mod_struct)) return;
    PyMod *m = static_cast<PyMod *>(mod_struct);
    m->magic = 0;
    delete m;
}

double MCMC_likelihood_extrinsic_pyv2(bool save_waveform, double *parameters, void *mod_struct, int dimension,
                                      char *generation_method, int *data_length, double *frequencies,
                                      double *dataREAL, double *dataIMAG, double *psd, double *weights,
                                      char *integration_method, bool log10F, char *detectors, int num_detectors)
{
    const double nan = std::nan("");
    PyMod *m = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_handles_mu);
        if (mod_struct && g_mods.count(mod_struct) && static_cast<PyMod *>(mod_struct)->magic == kModMagic)
            m = static_cast<PyMod *>(mod_struct);
    }
    if (!m || !parameters || !generation_method) {
        gwh_report(GWH_ERR_INVALID, "not a modification handle from MCMC_modification_struct_py");
        return nan;
    }
    if (m->ns) {
        gwh_report(GWH_ERR_UNSUPPORTED, "the engine does not model NSflag1/NSflag2");
        return nan;
    }
    /* MCMC_prep_params (mcmc_gw.cpp:2492-2567): f_ref 20 Hz, both alignments on, the sampler's gmst (the
     * module's mcmc_gmst, zero until a sampler sets it), then the MCMC basis vector with the modification
     * parameters after the GR block */
    gwh_params p;
    memset(&p, 0, sizeof(p));
    p.f_ref = 20.;
    p.shift_time = 1;
    p.shift_phase = 1;
    p.chip = -1;
    p.phip = -1;
    p.cosmology = GWH_COSMO_PLANCK15;
    const int theory = gwh::theory_id(generation_method);
    const bool ppe = strstr(generation_method, "ppE_") || theory > 0;
    const bool gimr = !strncmp(generation_method, "gIMR", 4);
    int nmod = 0;
    if (ppe) {
        if (m->ppe_nmod > GWH_PPE_MAX) {
            gwh_report(GWH_ERR_UNSUPPORTED, "more than GWH_PPE_MAX ppE terms");
            return nan;
        }
        p.ppe_nmod = nmod = m->ppe_nmod;
        for (int i = 0; i < nmod; i++) p.ppe_b[i] = i < (int)m->bppe.size() ? m->bppe[i] : 0.0;
    } else if (gimr) {
        fill_gimr_list(p, m->phii, m->sigmai, m->betai, m->alphai);
        if (p.gimr_nlist < 0) {
            gwh_report(GWH_ERR_UNSUPPORTED, "at most 16 gIMR deltas in a parameter vector");
            return nan;
        }
        nmod = p.gimr_nlist;
    }
    std::vector<double> v(parameters, parameters + dimension);
    const int base = dimension - nmod;
    /* dCS and EdGB sample the length scale in km: the theory's parameter is (l / (c / 1000))^4 */
    if (nmod > 0 && base >= 0 && (strstr(generation_method, "dCS") || strstr(generation_method, "EdGB"))) {
        const double x = v[base] / (299792458. / 1000.);
        v[base] = x * x * x * x;
    }
    std::string gr = std::string("MCMC_") + gr_family(generation_method);
    const int basis = basis_of(gr.c_str(), base, 0);
    if (basis < 0) {
        gwh_report(GWH_ERR_UNSUPPORTED, "generation method / dimension not on the engine's path");
        return nan;
    }
    gwh_repack_parameters(v.data(), basis, &p);
    set_mod_params(p, v.data() + base, nmod, ppe, gimr);
    return extrinsic_flat(save_waveform, p, generation_method, data_length, frequencies, dataREAL, dataIMAG, psd,
                          weights, integration_method, log10F, detectors, num_detectors);
}

void repack_parameters_py(double *parameters, void *gen_param, char *generation_method, int dim)
{
    PyParams *q = handle(gen_param);
    if (!q || !parameters) {
        gwh_report(GWH_ERR_INVALID, "not a parameter handle from gen_params_base_py");
        return;
    }
    /* a ppE method's betas, a gIMR method's listed deltas follow its GR basis (fisher.cpp:2083-2088, :2128-2158,
     * :2424-2429, :2473-2505) */
    const int theory = generation_method ? gwh::theory_id(generation_method) : 0;
    const bool ppe = generation_method && (strstr(generation_method, "ppE_") || theory > 0);
    const bool gimr = generation_method && !strncmp(generation_method, "gIMR", 4);
    if (gimr && q->p.gimr_nlist < 0) {
        gwh_report(GWH_ERR_UNSUPPORTED, "at most 16 gIMR deltas in a parameter vector");
        return;
    }
    const int nmod = ppe ? q->p.ppe_nmod : (gimr ? q->p.gimr_nlist : 0);
    std::string gr = generation_method ? generation_method : "";
    if (gimr) gr.erase(0, 1);
    if (theory > 0) /* a theory's parameters follow its family's GR basis, as betappe does */
        gr = gwh::fisher_gr_method(gr);
    else if (ppe) {
        gr.erase(gr.find("ppE_"), 4);
        for (const char *suf : {"_Inspiral", "_IMR"}) {
            const std::string s(suf);
            if (gr.size() > s.size() && gr.compare(gr.size() - s.size(), s.size(), s) == 0) {
                gr.erase(gr.size() - s.size());
                break;
            }
        }
    }
    int basis = basis_of(gr.c_str(), dim - nmod, q->p.sky_average);
    if (basis < 0) {
        gwh_report(GWH_ERR_UNSUPPORTED, "generation method / dimension not on the engine's path");
        return;
    }
    gwh_repack_parameters(parameters, basis, &q->p);
    set_mod_params(q->p, parameters + dim - nmod, nmod, ppe, gimr);
}

double gps_to_GMST_radian_py(double gps) { return gwh_gps_to_gmst_radian(gps); }

double DTOA_DETECTOR_py(double RA, double DEC, double GMST_rad, char *det1, char *det2)
{
    int a = gwh_detector_id(det1), b = gwh_detector_id(det2);
    if (a < 0 || b < 0) {
        gwh_report(GWH_ERR_UNSUPPORTED, "unsupported detector");
        return std::nan("");
    }
    return gwh_dtoa(RA, DEC, GMST_rad, a, b);
}

} /* extern "C" */
