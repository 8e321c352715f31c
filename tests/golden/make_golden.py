"""Copies the reference's own model fixtures (notebooks/true_layers.H@, smooth_layers.H@:
VpVsRho, n1=101 z, n2=201 x, n3=3, d=10 m, XDR float32) into tests/golden as data, with a
fresh header, and records the reference's known-answer values used by the CPU tests.

Run from the repo root in the build container (the reference is not on the GPU box):
    python tests/golden/make_golden.py /root/reference
"""
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    nb = os.path.join(ref, "elastic_iso_lib", "notebooks")
    for name in ("true_layers", "smooth_layers"):
        shutil.copyfile(os.path.join(nb, name + ".H@"), os.path.join(HERE, name + ".H@"))
        with open(os.path.join(HERE, name + ".H"), "w") as f:
            f.write("n1=101 o1=0 d1=10\nn2=201 o2=0 d2=10\nn3=3 o3=0 d3=10\n"
                    'esize=4\ndata_format="xdr_float"\nin=./%s.H@\n' % name)
    kat = {
        # NB/Nonlinear.ipynb cell 10 (padding program output)
        "pad_marmousi": {"nz": 701, "nx": 3401, "zPad": 100, "xPad": 100,
                         "zPadPlus": 111, "xPadPlus": 115, "nzNew": 920, "nxNew": 3624},
        # NB/Elastic FWI example.ipynb cells 15/23 and NB/EFWI_test.p
        "pad_layers": {"nz": 101, "nx": 201, "zPad": 100, "xPad": 100,
                       "zPadPlus": 103, "xPadPlus": 115, "nzNew": 312, "nxNew": 424},
        # NB/Nonlinear.ipynb cell 23 (fdParam info): sub=10 -> fine sub 20, ntw 39981, dtw 0.0002
        "time_axes_marmousi": {"nts": 2000, "dts": 0.004, "sub": 10, "sub_fd": 20, "ntw": 39981,
                               "dtw": 0.0002},
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
