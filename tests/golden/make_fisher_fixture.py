"""Extract a subset of the reference's own Fisher fixtures into tests/golden.

Source: testing/fisher_data/{parameters,numerical,autodiff}.csv of the
reference (written by testing/fisher_comparison.cpp:34-233): sky-averaged 7x7
IMRPhenomD Fisher matrices on 3000 bins, 15-1000 Hz, Hanford_O1_fitted PSD,
basis (ln A0, phi_c, t_c, ln Mc, ln eta, chi_s, chi_a).  Only data is copied.

    python tests/golden/make_fisher_fixture.py /root/reference
"""
import os
import sys

import numpy as np

ROWS = 60


def main(ref):
    d = os.path.join(ref, "testing", "fisher_data")
    par = np.loadtxt(os.path.join(d, "parameters.csv"), delimiter=",")[:ROWS]
    num = np.loadtxt(os.path.join(d, "numerical.csv"), delimiter=",")[:ROWS]
    aut = np.loadtxt(os.path.join(d, "autodiff.csv"), delimiter=",")[:ROWS]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fisher_data_subset.npz")
    np.savez_compressed(out, parameters=par, numerical=num.reshape(-1, 7, 7), autodiff=aut.reshape(-1, 7, 7))
    print("wrote", out, par.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
