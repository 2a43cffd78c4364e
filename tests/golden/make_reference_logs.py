"""Extract the parameter printouts and time-averaged results that the
reference keeps in its example logs into reference_logs.json.

Reads the reference's text logs (run once, where /root/reference exists):
  examples/Piel2017/Figure1/consistent/outputPiel2017Fig1Consistent
  examples/Piel2017/Figure1/oml/outputPiel2017Fig1OML
  examples/Piel2017/Figure1/given/outputPiel2017Fig1Given
  examples/Patacchini2008/Figure3/outputPatacchini2008
The inputs that produced each log are transcribed from the matching launch
scripts (launchPiel2017Figure1*.py:13-59, launchPatacchini2008Figure3.py:66-107).
"""

import json
import os

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))


def block(path, start, stop):
    with open(os.path.join(REF, path)) as fh:
        lines = fh.read().splitlines()
    return lines[start - 1:stop]


def means(path):
    with open(os.path.join(REF, path)) as fh:
        lines = fh.read().splitlines()
    out = {}
    for ln in lines:
        for key, name in (("Mean force from ions orbits = ", "orbit"),
                          ("Mean force from collected ions = ", "collect"),
                          ("Mean force from ions = ", "total")):
            if ln.startswith(key):
                out[name] = float(ln[len(key):])
    return out


data = {
    "piel2017_fig1_consistent": {
        "source": "examples/Piel2017/Figure1/consistent/outputPiel2017Fig1Consistent:1-30",
        "inputs": {"T_i": 290.11, "T_e": 29011, "n_inf": 1e14, "M": 1, "m_i": 6.6335209e-26,
                   "R_over_lambda": 3, "H_over_lambda": 6, "N": 65536, "n": 3000, "d_t": 3.5148240854e-09,
                   "dust_r": 58.8e-6, "dust_z_over_lambda": -1, "charge": "consistent", "q_over_e": 0},
        "printout": block("Piel2017/Figure1/consistent/outputPiel2017Fig1Consistent", 1, 30),
        "means_avg_i_gt_1000": means("Piel2017/Figure1/consistent/outputPiel2017Fig1Consistent"),
    },
    "piel2017_fig1_oml": {
        "source": "examples/Piel2017/Figure1/oml/outputPiel2017Fig1OML",
        "means_avg_i_gt_1000": means("Piel2017/Figure1/oml/outputPiel2017Fig1OML"),
    },
    "piel2017_fig1_given": {
        "source": "examples/Piel2017/Figure1/given/outputPiel2017Fig1Given (N = 2^19, q = 392500 e)",
        "means_avg_i_gt_1000": means("Piel2017/Figure1/given/outputPiel2017Fig1Given"),
    },
    "patacchini2008_fig3_point1": {
        "source": "examples/Patacchini2008/Figure3/outputPatacchini2008:1-32",
        "inputs": {"T_n": 300, "T_e": 30000, "n_inf": 3.57167962497e15, "m_i": 1.673557e-27,
                   "w_c": 1.57319303e04, "E": 2.58520000e00, "R_over_lambda": 3, "H_over_lambda": 6,
                   "N": 131072, "n": 200000, "d_t": 1e-11},
        "printout": block("Patacchini2008/Figure3/outputPatacchini2008", 1, 32),
    },
}

with open(os.path.join(HERE, "reference_logs.json"), "w") as fh:
    json.dump(data, fh, indent=1)
print("wrote reference_logs.json")
