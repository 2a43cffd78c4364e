"""Writes xyz_frame_head.txt: the first 1000 ion lines of the reference's XYZ
fixture examples/DustDynamics/TwoParticles/data_dynamics/trajectory0.xyz
(a 2^16-ion frame written by OpenDust's output path, opendust.py:1080-1088).
Run in a checkout that has the reference next to this repository:

    python tests/golden/make_xyz_fixture.py /root/reference
"""

import os
import sys

src = os.path.join(sys.argv[1], "examples/DustDynamics/TwoParticles/data_dynamics/trajectory0.xyz")
with open(src) as fh:
    lines = [next(fh) for _ in range(1002)]
assert lines[0].strip() == "65536" and lines[1].strip() == "0"
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "xyz_frame_head.txt"), "w") as out:
    out.writelines(lines[2:])
