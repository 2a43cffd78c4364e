#!/bin/sh
# Extracts the per-point mean forces that examples/Piel2017/Figure5/launchPiel2017Figure5.py
# printed (Mach number, <F_orbit,z>, <F_collect,z> in N, means over steps i > 3000 of 5000)
# from the reference's output file.
# Usage: sh tests/golden/make_piel2017_fig5.sh /root/reference
grep -v "^Step\|^$\|#\|=" "$1/examples/Piel2017/Figure5/outputPiel2017Fig5" > "$(dirname "$0")/piel2017_fig5_reference.txt"
