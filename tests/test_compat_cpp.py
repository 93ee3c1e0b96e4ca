"""The header-only C++ layer (include/gwat_hip_compat.hpp) compiles against a parameter object shaped like
the reference's gen_params_base<double> and maps its fields and strings onto the C ABI (host-side
checks only; no device work)."""
import os
import shutil
import subprocess

import pytest

import gw_analysis_tools_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include <cstdio>
#include <cmath>
#include "gwat_hip_compat.hpp"

/* a stand-in with the field names (not the layout) of gen_params_base<double> */
struct params_like {
    double mass1 = 36, mass2 = 29, Luminosity_Distance = 410, spin1[3] = {0.1, 0.2, 0.3},
           spin2[3] = {-0.1, 0, -0.2}, tc = 0.01, phiRef = 0.4, f_ref = 20, incl_angle = 2.9, RA = 1.37,
           DEC = -1.26, psi = 2.66, gmst = 2.45, chip = -1, phip = -1;
    bool shift_time = true, shift_phase = true, sky_average = false, dep_postmerger = false,
         equatorial_orientation = false, horizon_coord = false, NSflag1 = false, NSflag2 = false;
    int Nmod = 0, Nmod_phi = 0, Nmod_sigma = 0, Nmod_beta = 0, Nmod_alpha = 0;
    double *bppe = nullptr, *betappe = nullptr;
    int *phii = nullptr, *sigmai = nullptr, *betai = nullptr, *alphai = nullptr;
    double *delta_phi = nullptr, *delta_sigma = nullptr, *delta_beta = nullptr, *delta_alpha = nullptr;
};

int main()
{
    params_like g;
    gwh_params p;
    int ok = gwh::to_params(g, p);
    printf("ok %d %.17g %.17g %.17g %.17g %.17g %d\n", ok, p.mass1, p.DL, p.spin2[2], p.incl, p.gmst, p.shift_time);
    g.NSflag2 = true;
    printf("ns %d\n", (int)gwh::to_params(g, p));
    g.NSflag2 = false;
    printf("models %d %d %d %d %d\n", gwh::model_id("IMRPhenomD"), gwh::model_id("IMRPhenomPv2"),
           gwh::model_id("PhenomD"), gwh::model_id("ppE_IMRPhenomD_Inspiral"), gwh::model_id("ppE_IMRPhenomD_IMR"));
    /* ppE fields: read for the ppE models only (a GR method ignores them, as prep_source_parameters does) */
    double b[2] = {-7, -3}, be[2] = {1e-5, -2e-3};
    g.Nmod = 2;
    g.bppe = b;
    g.betappe = be;
    int okgr = gwh::to_params(g, p, GWH_IMRPHENOMD);
    int ngr = p.ppe_nmod;
    int okppe = gwh::to_params(g, p, GWH_PPE_IMRPHENOMD_IMR);
    printf("ppe %d %d %d %d %.17g %.17g\n", okgr, ngr, okppe, p.ppe_nmod, p.ppe_b[0], p.ppe_beta[1]);
    g.Nmod = GWH_PPE_MAX + 1;
    printf("ppemax %d\n", (int)gwh::to_params(g, p, GWH_PPE_IMRPHENOMD_INSPIRAL));
    g.Nmod = 0;
    /* gIMR lists -> tables: first entry of an index wins, out-of-range indices ignored */
    int ph[4] = {2, -3, 2, 12};
    double dph[4] = {0.1, 0.2, 0.3, 0.4};
    int al[1] = {5};
    double dal[1] = {-0.5};
    g.Nmod_phi = 4;
    g.phii = ph;
    g.delta_phi = dph;
    g.Nmod_alpha = 1;
    g.alphai = al;
    g.delta_alpha = dal;
    gwh::to_params(g, p, GWH_GIMRPHENOMD);
    printf("gimr %d %g %g %g %g\n", gwh::model_id("gIMRPhenomD"), p.gimr_dphi[2 + 4], p.gimr_dphi[-3 + 4],
           p.gimr_dalpha[5 - 2], p.gimr_dphi[0]);
    gwh::to_params(g, p, GWH_IMRPHENOMD);
    printf("gimrgr %g\n", p.gimr_dphi[2 + 4]);
    g.Nmod_phi = g.Nmod_alpha = 0;
    printf("bases %d %d %d %d %d\n", gwh::basis_id("IMRPhenomD", true), gwh::basis_id("IMRPhenomD", false),
           gwh::basis_id("MCMC_IMRPhenomD", false), gwh::basis_id("IMRPhenomPv2", false),
           gwh::basis_id("MCMC_IMRPhenomPv2", false));
    /* repack through the C ABI: the MCMC basis vector of these parameters, bumped, written back */
    double v[11];
    gwh::to_params(g, p);
    gwh_unpack_parameters(&p, GWH_BASIS_PHENOMD_MCMC, v, nullptr);
    v[6] += 0.1; /* ln DL */
    gwh::repack_parameters(v, &g, "MCMC_IMRPhenomD", 11);
    printf("repack %.17g %.17g\n", g.Luminosity_Distance, g.mass1);
    /* the gIMR lists as a Fisher basis: codes in list order (-1: an index the model does not modify), and a
     * basis vector written back into the caller's delta arrays entry by entry */
    g.Nmod_phi = 4;
    g.Nmod_alpha = 1;
    gwh::to_params(g, p, GWH_GIMRPHENOMD);
    printf("gimrlist %d %d %d %d %d %d %d\n", p.gimr_nlist, p.gimr_list[0], p.gimr_list[1], p.gimr_list[2],
           p.gimr_list[3], p.gimr_list[4], gwh::basis_id("MCMC_gIMRPhenomD", false));
    double w[16];
    gwh_unpack_parameters(&p, GWH_BASIS_PHENOMD_FULL, w, nullptr);
    for (int i = 0; i < 5; i++) w[11 + i] = 0.5 + 0.1 * i;
    gwh::repack_parameters(w, &g, "gIMRPhenomD", 16);
    printf("gimrrepack %g %g %g %g %g\n", dph[0], dph[1], dph[2], dph[3], dal[0]);
    /* theory method strings: the theory id and the ppE method the reference maps each to */
    const char *th[8] = {"dCS_IMRPhenomD", "EdGB_HO_IMRPhenomPv2", "EdGB_HO_LO_IMRPhenomD", "EdGB_GHOv2_IMRPhenomD",
                         "PNSeries_ppE_IMRPhenomD_IMR", "ppEAlt_ppE_IMRPhenomPv2_Inspiral", "ModDispersion_IMRPhenomD",
                         "dCS_IMRPhenomD_NRT"};
    printf("theories");
    for (int i = 0; i < 8; i++) {
        int m = -9;
        int t = gwh::theory_id(th[i], &m);
        printf(" %d:%d", t, m);
    }
    printf(" %d %s %d\n", gwh::model_id("TVG_IMRPhenomD"), gwh::fisher_gr_method("MCMC_dCS_IMRPhenomPv2").c_str(),
           gwh::cosmology_id("wmap9"));
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_compat_header_with_reference_shaped_params(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(SRC)
    exe = tmp_path / "t"
    libdir = os.path.dirname(G.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", libdir, "-lgwat_hip", "-Wl,-rpath," + libdir, "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.check_call(cmd)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    f = out[0].split()
    assert f[:2] == ["ok", "1"] and float(f[2]) == 36 and float(f[3]) == 410 and float(f[4]) == -0.2
    assert float(f[5]) == 2.9 and float(f[6]) == 2.45 and f[7] == "1"
    assert out[1] == "ns 0"
    assert out[2] == "models 0 1 -1 %d %d" % (G.PPE_IMRPHENOMD_INSPIRAL, G.PPE_IMRPHENOMD_IMR)
    assert out[3].split() == ["ppe", "1", "0", "1", "2", "-7", "-0.002"]
    assert out[4] == "ppemax 0"
    assert out[5].split() == ["gimr", str(G.GIMRPHENOMD), "0.1", "0.2", "-0.5", "0"]
    assert out[6] == "gimrgr 0"
    out = out[:3] + out[7:]
    assert out[3] == "bases %d %d %d %d %d" % (G.BASIS_PHENOMD_SKYAVG, G.BASIS_PHENOMD_FULL, G.BASIS_PHENOMD_MCMC,
                                             G.BASIS_PV2_FULL, G.BASIS_PV2_MCMC)
    r = out[4].split()
    assert abs(float(r[1]) - 410 * 2.718281828459045 ** 0.1) < 1e-9 and abs(float(r[2]) - 36) < 1e-9
    # phi 2, -3, 2, 12 (out of range), alpha 5 -> codes 6, 1, 6, -1, 22
    assert out[5].split() == ["gimrlist", "5", "6", "1", "6", "-1", "22", str(G.BASIS_PHENOMD_MCMC)]
    assert out[6].split() == ["gimrrepack", "0.5", "0.6", "0.7", "0.8", "0.9"]
    D, DI, PI = G.PPE_IMRPHENOMD_INSPIRAL, G.PPE_IMRPHENOMD_IMR, G.PPE_IMRPHENOMPV2_INSPIRAL
    assert out[7].split() == ["theories", "%d:%d" % (G.THEORY_DCS, D), "%d:%d" % (G.THEORY_EDGB, PI),
                              "%d:%d" % (G.THEORY_EDGB_HO_LO, D), "%d:%d" % (G.THEORY_EDGB_GHOV2, D),
                              "%d:%d" % (G.THEORY_PNSERIES, DI), "%d:%d" % (G.THEORY_PPEALT, PI),
                              "%d:%d" % (G.THEORY_MODDISPERSION, DI), "-1:-9",
                              str(D), "MCMC_IMRPhenomPv2", str(G.COSMO_WMAP9)]
