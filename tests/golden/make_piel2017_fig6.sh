#!/bin/sh
# Extracts the per-point mean forces that examples/Piel2017/Figure6/launchPiel2017Figure6.py
# printed (Mach number, <F1x>, <F1z>, <F2x>, <F2z> in N, orbit + dust-dust, means over
# steps i > 10000) from the reference's output file.
# Usage: sh tests/golden/make_piel2017_fig6.sh /root/reference
grep -v "^Step\|^$\|#\|=" "$1/examples/Piel2017/Figure6/outputPiel2017Fig6" > "$(dirname "$0")/piel2017_fig6_reference.txt"
