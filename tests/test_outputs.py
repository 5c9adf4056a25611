"""`.pdr.fin` writer (pdrchem/outputs.py) against the layout of the
reference's own output (the first lines of Benchmarks/B01.pdr.fin, a PSEUDO_1D
reduced-network run, kept as a fixture): parsing a reference line and writing
it back reproduces it character for character."""
import os

import numpy as np

from pdrchem import outputs

HERE = os.path.dirname(os.path.abspath(__file__))


def test_pseudo_1d_line_round_trip():
    lines = open(os.path.join(HERE, "golden", "B01_pdr_fin_head.txt")).read().splitlines()
    for line in lines:
        f = line.split()
        pid, x, av6, tg, td, et, rho, uv = int(f[0]), *map(float, f[1:5]), int(f[5]), float(f[6]), float(f[7])
        ab = [float(v) for v in f[8:]]
        assert len(ab) == 33                                   # reduced network
        av = np.zeros(12)
        av[6] = av6
        out = outputs.pdr_fin_line(pid, (x, 0.0, 0.0), tg, td, et, rho, uv, ab, av, pseudo_1d=True)
        assert out == line.rstrip("\n").rstrip() or out == line.rstrip("\n")


def test_3d_line_layout(tmp_path):
    xyz = np.array([[0.0, 0.5, 1.25], [1e-3, 2e-3, 3e-3]])
    ab = np.array([[1.0, 1e-30, 0.0], [0.5, 2.5e-12, 1e-120]])
    av = np.array([[0.0] * 12, [1.5] * 12])
    p = tmp_path / "x.pdr.fin"
    outputs.write_pdr_fin(str(p), xyz, [10.0, 20.0], [20.0, 20.0], [1e3, 2e3], [1.0, 0.5], ab, av)
    l0, l1 = p.read_text().splitlines()
    assert len(l0) == 7 + 5 * 11 + 5 + 11 * (2 + 3 + 12)
    assert l0[:7] == "      1" and l1[:7] == "      2"
    assert l1[7:18] == "  1.000E-03" and l1[62:67] == "    1"
    assert "0.100-119" in l1 or "1.000-120" in l1              # three-digit exponent as Fortran prints it
