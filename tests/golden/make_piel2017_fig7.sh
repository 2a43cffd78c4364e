#!/bin/sh
# Extracts the per-point mean forces that examples/Piel2017/Figure7/launchPiel2017Figure7.py
# printed (d_x / r_D_e, <F1x>, <F1z>, <F2x>, <F2z> in N, means over steps i > 2500)
# from the reference's output file.  Usage: sh tests/golden/make_piel2017_fig7.sh /root/reference
grep -v "^Step\|^$\|#\|=" "$1/examples/Piel2017/Figure7/outputPiel2017Fig7" > "$(dirname "$0")/piel2017_fig7_reference.txt"
