"""Regenerate ``gwh_tables.h``: the numeric tables the waveform and detector
models depend on.

The numbers are physical/calibration data, not code: the IMRPhenomD
calibration fits (Khan et al. 2016, arXiv:1508.07253, Table V), the Kerr
quasi-normal-mode ringdown/damping frequencies of the (2,2,0) mode tabulated
against the remnant spin, and the ground-based detector response tensors and
Earth-fixed vertex locations.  Results can only match the reference if these
numbers are bit-identical, so they are read from the reference's headers as
text and re-emitted here in this project's own layout:

  include/gwat/QNM_data.h            QNMData_a / QNMData_fring / QNMData_fdamp
  include/gwat/IMRPhenomD.h:190-270  lambda_num_params[19][11]
  include/gwat/detector_util.h:30-161  *_location[3], *_D[3][3]
  include/gwat/D_Z_Config.h          cosmos, num_segments, interp_degree, boundaries_D, COEFF_VEC_DZ
                                     (redshift from luminosity distance: per cosmology, segments of a
                                     power series in sqrt(D_L / Mpc), fitted to astropy.cosmology)
  include/gwat/D_Z_Config_modified_dispersion.h  MD_alphas, num_MD_segments, interp_MD_degree,
                                     MD_boundaries_Z, MD_COEFF_VEC_ZD (the modified-dispersion distance
                                     D_alpha(z) of Planck15: segments of a series in z^(k/2 - 7/2))

Run from the repository root (only needed when the reference tables change):
    python gw_analysis_tools_amd/csrc/gen_tables.py /root/reference
"""
import os
import re
import sys

# A sign may be separated from its literal by blanks ("- 0.12" is unary minus in C).
NUM = r"[-+]?\s*(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"

# Detector order defines the integer detector ids of the C-ABI.
DETECTORS = [
    ("HANFORD", "H", "Hanford"),
    ("LIVINGSTON", "L", "Livingston"),
    ("VIRGO", "V", "Virgo"),
    ("KAGRA", "K", "Kagra"),
    ("INDIGO", "I", "Indigo"),
    ("CE", "CE", "CE"),
    ("ET1", "ET1", "ET1"),
    ("ET2", "ET2", "ET2"),
    ("ET3", "ET3", "ET3"),
]


def _numbers(text):
    # Strip C comments before extracting literals.
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return [float(re.sub(r"\s+", "", x)) for x in re.findall(NUM, text)]


def _array_body(src, name):
    m = re.search(re.escape(name) + r"\s*(?:\[[^\]]*\])+\s*=\s*\{", src)
    if not m:
        raise RuntimeError("table %s not found" % name)
    i = m.end() - 1
    depth = 0
    for j in range(i, len(src)):
        if src[j] == "{":
            depth += 1
        elif src[j] == "}":
            depth -= 1
            if depth == 0:
                return src[i:j + 1]
    raise RuntimeError("unterminated table %s" % name)


def _strip_comments(src):
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.sub(r"//[^\n]*", "", src)


def _fmt(vals, per_line=4):
    out = []
    for k in range(0, len(vals), per_line):
        out.append("    " + ", ".join(repr(v) for v in vals[k:k + per_line]) + ",")
    return "\n".join(out)


def main(ref):
    inc = os.path.join(ref, "include", "gwat")
    qnm = _strip_comments(open(os.path.join(inc, "QNM_data.h")).read())
    phd = _strip_comments(open(os.path.join(inc, "IMRPhenomD.h")).read())
    det = _strip_comments(open(os.path.join(inc, "detector_util.h")).read())
    dz = _strip_comments(open(os.path.join(inc, "D_Z_Config.h")).read())
    md = _strip_comments(open(os.path.join(inc, "D_Z_Config_modified_dispersion.h")).read())

    n = int(re.search(r"QNMData_length\s*=\s*(\d+)", qnm).group(1))
    spin = _numbers(_array_body(qnm, "QNMData_a"))
    fring = _numbers(_array_body(qnm, "QNMData_fring"))
    fdamp = _numbers(_array_body(qnm, "QNMData_fdamp"))
    assert len(spin) == len(fring) == len(fdamp) == n, (len(spin), len(fring), len(fdamp), n)
    fits = _numbers(_array_body(phd, "lambda_num_params"))
    assert len(fits) == 19 * 11, len(fits)

    tensors, locs = [], []
    for _, pre, dname in DETECTORS:
        t = _numbers(_array_body(det, dname + "_D"))
        l = _numbers(_array_body(det, pre + "_location"))
        assert len(t) == 9 and len(l) == 3, dname
        tensors.append(t)
        locs.append(l)

    cosmos = re.findall(r'"([A-Z0-9_]+)"', _array_body(dz, "cosmos"))
    nc = len(cosmos)
    nseg = [int(v) for v in _numbers(_array_body(dz, "num_segments"))]
    deg = [int(v) for v in _numbers(_array_body(dz, "interp_degree"))]
    bnd = _numbers(_array_body(dz, "boundaries_D"))
    cdz = _numbers(_array_body(dz, "COEFF_VEC_DZ"))
    S, D = max(nseg), max(deg)
    assert len(nseg) == len(deg) == nc and len(bnd) == nc * (S + 1) and len(cdz) == nc * S * D, (nc, len(bnd), len(cdz))

    md_alpha = _numbers(_array_body(md, "MD_alphas"))
    na = len(md_alpha)
    md_nseg = [int(v) for v in _numbers(_array_body(md, "num_MD_segments"))]
    md_deg = [int(v) for v in _numbers(_array_body(md, "interp_MD_degree"))]
    md_bnd = _numbers(_array_body(md, "MD_boundaries_Z"))
    md_c = _numbers(_array_body(md, "MD_COEFF_VEC_ZD"))
    MS, MD = max(md_nseg), max(md_deg)
    assert len(md_nseg) == len(md_deg) == na and len(md_bnd) == na * (MS + 1) and len(md_c) == na * MS * MD

    lines = []
    lines.append("/* Generated by gen_tables.py -- do not edit by hand.")
    lines.append(" * Calibration and detector data (see gen_tables.py for provenance). */")
    lines.append("#ifndef GWH_TABLES_H")
    lines.append("#define GWH_TABLES_H")
    lines.append("")
    lines.append("#define GWH_QNM_N %d" % n)
    lines.append("/* remnant dimensionless spin a, and M*f_ring, M*f_damp of the (2,2,0) mode */")
    lines.append("static const double gwh_qnm_spin[GWH_QNM_N] = {\n%s\n};" % _fmt(spin))
    lines.append("static const double gwh_qnm_fring[GWH_QNM_N] = {\n%s\n};" % _fmt(fring))
    lines.append("static const double gwh_qnm_fdamp[GWH_QNM_N] = {\n%s\n};" % _fmt(fdamp))
    lines.append("")
    lines.append("/* IMRPhenomD fits: row r holds the 11 coefficients of")
    lines.append(" * p0 + p1 eta + x(p2 + p3 eta + p4 eta^2) + x^2(p5..p7) + x^3(p8..p10), x = chi_PN - 1.")
    lines.append(" * Rows: rho1..3, v2, gamma1..3, sigma1..4, beta1..3, alpha1..5. */")
    lines.append("/* a compile-time constant in C++ (device code folds the coefficients into its instructions); const in C */")
    lines.append("#ifdef __cplusplus")
    lines.append("#define GWH_FIT_CONST constexpr")
    lines.append("#else")
    lines.append("#define GWH_FIT_CONST const")
    lines.append("#endif")
    lines.append("static GWH_FIT_CONST double gwh_phenomd_fit[19][11] = {")
    for r in range(19):
        lines.append("  {" + ", ".join(repr(v) for v in fits[11 * r:11 * r + 11]) + "},")
    lines.append("};")
    lines.append("")
    lines.append("#define GWH_NDET_TABLE %d" % len(DETECTORS))
    for k, (enum, _, dname) in enumerate(DETECTORS):
        lines.append("#define GWH_DET_%s %d" % (enum, k))
    lines.append("/* detector response tensor D_ij (row major) */")
    lines.append("static const double gwh_det_tensor[GWH_NDET_TABLE][9] = {")
    for t in tensors:
        lines.append("  {" + ", ".join(repr(v) for v in t) + "},")
    lines.append("};")
    lines.append("/* Earth-centred vertex location, metres */")
    lines.append("static const double gwh_det_location[GWH_NDET_TABLE][3] = {")
    for l in locs:
        lines.append("  {" + ", ".join(repr(v) for v in l) + "},")
    lines.append("};")
    lines.append("")
    lines.append("/* redshift from luminosity distance: cosmology c, segment s (D_L below gwh_cosmo_dl_bound[c][s + 1],")
    lines.append(" * Mpc) holds z = sum_k gwh_cosmo_dz[c][s][k] (D_L / Mpc)^(k/2), k < gwh_cosmo_degree[c] */")
    lines.append("#define GWH_NCOSMO %d" % nc)
    lines.append("#define GWH_COSMO_SEG %d" % S)
    lines.append("#define GWH_COSMO_DEG %d" % D)
    for k, name in enumerate(cosmos):
        lines.append("#define GWH_COSMO_%s %d" % (name, k))
    lines.append("static const int gwh_cosmo_nseg[GWH_NCOSMO] = {%s};" % ", ".join(str(v) for v in nseg))
    lines.append("static const int gwh_cosmo_degree[GWH_NCOSMO] = {%s};" % ", ".join(str(v) for v in deg))
    lines.append("static GWH_FIT_CONST double gwh_cosmo_dl_bound[GWH_NCOSMO][GWH_COSMO_SEG + 1] = {")
    for c in range(nc):
        lines.append("  {" + ", ".join(repr(v) for v in bnd[(S + 1) * c:(S + 1) * (c + 1)]) + "},")
    lines.append("};")
    lines.append("static GWH_FIT_CONST double gwh_cosmo_dz[GWH_NCOSMO][GWH_COSMO_SEG][GWH_COSMO_DEG] = {")
    for c in range(nc):
        lines.append("  {")
        for s in range(S):
            o = (c * S + s) * D
            lines.append("    {" + ", ".join(repr(v) for v in cdz[o:o + D]) + "},")
        lines.append("  },")
    lines.append("};")
    lines.append("")
    lines.append("/* modified-dispersion distance (Planck15): alpha gwh_md_alpha[a], segment s (z below gwh_md_z_bound[a][s + 1])")
    lines.append(" * holds D_alpha / Mpc = sum_k gwh_md_dz[a][s][k] z^(k/2 - 7/2), k < gwh_md_degree[a] */")
    lines.append("#define GWH_MD_NALPHA %d" % na)
    lines.append("#define GWH_MD_SEG %d" % MS)
    lines.append("#define GWH_MD_DEG %d" % MD)
    lines.append("static GWH_FIT_CONST double gwh_md_alpha[GWH_MD_NALPHA] = {%s};" % ", ".join(repr(v) for v in md_alpha))
    lines.append("static const int gwh_md_nseg[GWH_MD_NALPHA] = {%s};" % ", ".join(str(v) for v in md_nseg))
    lines.append("static const int gwh_md_degree[GWH_MD_NALPHA] = {%s};" % ", ".join(str(v) for v in md_deg))
    lines.append("static GWH_FIT_CONST double gwh_md_z_bound[GWH_MD_NALPHA][GWH_MD_SEG + 1] = {")
    for a in range(na):
        lines.append("  {" + ", ".join(repr(v) for v in md_bnd[(MS + 1) * a:(MS + 1) * (a + 1)]) + "},")
    lines.append("};")
    lines.append("static GWH_FIT_CONST double gwh_md_dz[GWH_MD_NALPHA][GWH_MD_SEG][GWH_MD_DEG] = {")
    for a in range(na):
        lines.append("  {")
        for s in range(MS):
            o = (a * MS + s) * MD
            lines.append("    {" + ", ".join(repr(v) for v in md_c[o:o + MD]) + "},")
        lines.append("  },")
    lines.append("};")
    lines.append("")
    lines.append("#endif")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gwh_tables.h")
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
