#!/bin/sh
# Joins the per-point means that examples/Patacchini2008/Figure3/launchPatacchini2008Figure3.py
# appended to charge.txt and force.txt (w_c in 1/s, <q> in C, <F_orbit,z>, <F_collect,z> in N,
# means over steps i > 150000) into one table.  Usage: sh tests/golden/make_patacchini2008_fig3.sh /root/reference
d="$1/examples/Patacchini2008/Figure3"
paste "$d/charge.txt" "$d/force.txt" | awk -F'\t' '{print $1 "\t" $2 "\t" $4 "\t" $5}' > "$(dirname "$0")/patacchini2008_fig3_reference.txt"
