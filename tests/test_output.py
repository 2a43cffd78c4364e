"""Native frame writers (od_write_xyz / od_write_csv / od_write_restart) against
the reference's own output: Python's "{}".format of every number
(opendust.py:1080-1115) and the XYZ frame the reference wrote
(tests/golden/xyz_frame_head.txt, from
examples/DustDynamics/TwoParticles/data_dynamics/trajectory0.xyz).  Host code
only: these run without a GPU."""

import csv
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def eng():
    from opendust_amd import engine

    engine.load_library()
    return engine


def _values(rng, n):
    mant = rng.uniform(1.0, 10.0, n)
    expo = rng.integers(-330, 308, n)
    v = mant * 10.0 ** expo.astype(np.float64)
    v[rng.random(n) < 0.5] *= -1
    special = [0.0, -0.0, 1.0, -1.0, 0.1, 1e16, 1e15, 9999999999999998.0, 1e-4, 1e-5, 0.0001, 0.00001234,
               123456789012345678.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
               0.30000000000000004, 100.0, 1e22, 1e100, 1.5e-7, 12345.678, float("inf"), float("-inf")]
    v[: len(special)] = special
    # values the positions actually take: nm-scale doubles and float32 values
    v[len(special): len(special) + 2000] = rng.normal(0, 3e6, 2000)
    v[len(special) + 2000: len(special) + 4000] = rng.normal(0, 3e6, 2000).astype(np.float32)
    return v


def test_restart_matches_python_format(eng, tmp_path):
    rng = np.random.default_rng(1)
    v = _values(rng, 60000).reshape(-1, 6)
    p = tmp_path / "RESTART"
    eng.write_restart(p, v[:, :3], v[:, 3:])
    expect = "".join("{}\t{}\t{}\t{}\t{}\t{}\n".format(*row) for row in v.tolist())
    assert p.read_bytes() == expect.encode()


def test_nan_is_written_as_python_does(eng, tmp_path):
    v = np.array([[float("nan"), -float("nan"), 1.0]])
    p = tmp_path / "f.xyz"
    eng.write_xyz(p, v, 0)
    assert p.read_text() == "1\n0\nAr\t{}\t{}\t{}\n".format(*v[0].tolist())


def test_xyz_frame_reproduces_reference_file(eng, tmp_path):
    """Parse the reference's XYZ lines and write them back: byte-identical."""
    with open(os.path.join(HERE, "golden", "xyz_frame_head.txt")) as fh:
        lines = fh.read()
    pos = np.array([[float(t) for t in l.split("\t")[1:]] for l in lines.splitlines()])
    p = tmp_path / "t.xyz"
    eng.write_xyz(p, pos, 0)
    eng.write_xyz(p, pos[:10], 1)  # frames are appended
    text = p.read_text()
    head = "{}\n{}\n".format(len(pos), 0)
    assert text.startswith(head)
    assert text[len(head): len(head) + len(lines)] == lines
    tail = text[len(head) + len(lines):]
    assert tail == "10\n1\n" + "".join(lines.splitlines(keepends=True)[:10])


def test_csv_matches_csv_writer(eng, tmp_path):
    rng = np.random.default_rng(2)
    pos = rng.normal(0, 2e6, (5000, 3))
    pos[:3] = [[0.0, -0.0, 1e-5], [1e16, 3.0, -2.5], [float("inf"), 1e300, 5e-324]]
    dens, mass = 3.57e14 * 1e-27, 64.37617112
    p = tmp_path / "traj0.csv"
    eng.write_csv(p, pos, dens, mass)
    q = tmp_path / "ref.csv"
    with open(q, "w") as fh:
        w = csv.writer(fh)
        w.writerow(["x", "y", "z", "density", "mass"])
        for a, b, c in pos.tolist():
            w.writerow([a, b, c, dens, mass])
    assert p.read_bytes() == q.read_bytes()


def test_writer_errors_are_reported(eng, tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        eng.write_xyz(tmp_path / "no_such_dir" / "f.xyz", np.zeros((2, 3)), 0)
    with pytest.raises(ValueError):
        eng.write_restart(tmp_path / "R", np.zeros((2, 3)), np.zeros((3, 3)))
