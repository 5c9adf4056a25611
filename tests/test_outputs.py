"""Output writers (pdrchem/outputs.py) against the reference's own output
files (the first and last lines of each Benchmarks/B01.*.fin, a PSEUDO_1D
reduced-network run, kept as fixtures): parsing a reference line and writing
it back reproduces it character for character, and the row order and id
column of each file follow the reference's convention for its last two
points, as B01's last rows show it."""
import os

import numpy as np

from pdrchem import model, outputs

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


def fin_rows():
    rows, kind = {}, None
    for line in open(os.path.join(HERE, "golden", "B01_fin_rows.txt")).read().splitlines():
        if line.startswith("# "):
            kind = line.split()[1].split(".")[1]
            rows[kind] = []
        else:
            rows[kind].append(line)
    return rows


def test_thermal_files_round_trip(tmp_path):
    rows = fin_rows()
    writers = {"cool": (7, 2 + 5), "heat": (7, 2 + 12), "line": (9, 2 + 17), "spop": (9, 1 + 56),
               "opdp": (9, 1 + 17)}
    for kind, (w, nval) in writers.items():
        for line in rows[kind]:
            f = line.split()
            vals = [float(v) for v in f[1:]]
            assert len(vals) == nval, (kind, len(vals))
            assert outputs._row(int(f[0]), w, vals) == line, kind
    # and through the file writers: the first two rows of each file
    x = np.array([float(r.split()[1]) for r in rows["cool"][:2]])
    av6 = np.array([float(r.split()[2]) for r in rows["cool"][:2]])
    xyz = np.stack([x, np.zeros(2), np.zeros(2)], 1)
    av = np.zeros((2, 12))
    av[:, 6] = av6
    vals = lambda k, a: np.array([[float(v) for v in r.split()[a:]] for r in rows[k][:2]])
    outputs.write_cool_fin(str(tmp_path / "c"), xyz, av, vals("cool", 3), pseudo_1d=True)
    outputs.write_heat_fin(str(tmp_path / "h"), xyz, av, vals("heat", 3), pseudo_1d=True)
    outputs.write_spop_fin(str(tmp_path / "s"), av, vals("spop", 2))
    # B01.opdp.fin (2017) has Av(6) in its second column; the reference's
    # source writes x there (3DPDR.F90:1490), as the writer does
    x_o = np.array([float(r.split()[1]) for r in rows["opdp"][:2]])
    outputs.write_opdp_fin(str(tmp_path / "o"), np.stack([x_o, np.zeros(2), np.zeros(2)], 1), vals("opdp", 2))
    for k, fn in (("cool", "c"), ("heat", "h"), ("spop", "s"), ("opdp", "o")):
        assert (tmp_path / fn).read_text().splitlines() == rows[k][:2], k


def test_row_order_of_the_last_two_points(tmp_path):
    """a 5-point run in internal order (input order, last two swapped):
    .pdr.fin and .line.fin rows in input order with ids ..., 5, 4 (pdr) and
    indices ..., 5, 4 (line); .cool/.heat/.spop rows in internal order with
    ids 1..5 -- as B01's last rows: pdr ids 292, 294, 293 with x rising, cool
    ids 292, 293, 294 with x not, line indices 292, 294, 293."""
    n = 5
    x_in = np.array([0.1, 0.2, 0.3, 0.4, 0.5])
    order = outputs.internal_order(n)
    assert list(order) == [0, 1, 2, 4, 3]
    x = x_in[order]
    xyz = np.stack([x, np.zeros(n), np.zeros(n)], 1)
    av = np.tile(x[:, None] * 10, (1, 12))
    T = 10.0 * x
    outputs.write_reference_pdr_fin(str(tmp_path / "p"), xyz, T, T, np.ones(n), np.ones(n), np.ones((n, 2)), av,
                                    pseudo_1d=True)
    outputs.write_cool_fin(str(tmp_path / "c"), xyz, av, np.ones((n, 5)), pseudo_1d=True)
    outputs.write_line_fin(str(tmp_path / "l"), xyz, av, np.ones((n, 17)), pseudo_1d=True)
    outputs.write_spop_fin(str(tmp_path / "s"), av, np.ones((n, 3)))
    col = lambda fn, i: [float(r.split()[i]) for r in (tmp_path / fn).read_text().splitlines()]
    assert col("p", 0) == [1, 2, 3, 5, 4] and col("p", 1) == list(x_in)
    assert col("p", 3) == list(10.0 * x_in)                      # each row's own data
    assert col("c", 0) == [1, 2, 3, 4, 5] and col("c", 1) == list(x)
    assert col("l", 0) == [1, 2, 3, 5, 4] and col("l", 1) == list(x_in)
    assert col("s", 0) == [1, 2, 3, 4, 5] and col("s", 1) == list(10 * x)
    rows = fin_rows()
    ids = lambda k: [int(r.split()[0]) for r in rows[k][-3:]]
    xs = lambda k: [float(r.split()[1]) for r in rows[k][-3:]]
    assert ids("pdr") == [292, 294, 293] and xs("pdr") == sorted(xs("pdr"))
    assert ids("line") == [292, 294, 293] and xs("line") == sorted(xs("line"))
    assert ids("cool") == [292, 293, 294] and xs("cool") != sorted(xs("cool"))
    assert ids("heat") == [292, 293, 294] and xs("heat") == xs("cool")
    assert ids("spop") == [292, 293, 294]
    # the spop Av column of row 293 is the deepest point (input row 294)
    av_s = [float(r.split()[1]) for r in rows["spop"][-3:]]
    assert av_s[1] == max(av_s)


def test_non_finite_values_as_fortran_writes_them():
    assert outputs._es(float("nan")) == "        NaN"
    assert outputs._es(float("inf")) == "   Infinity"
    assert outputs._es(-float("inf")) == "  -Infinity"
    assert outputs._es(-4.763e-27) == " -4.763E-27"


def test_params_reader_round_trip(tmp_path):
    """read_params reads the reference's params.dat layout (values first on
    each line, comments after)"""
    text = ["=" * 10, "header", "=" * 10, "1Dn40.dat  !in", "B01  !out", "0", "1.3", "8", "6000", "0.9", "1e12",
            "1.3E-17", "1.0", "20.0", "1e7", "100.0", "1.0", "0.42", "1.0E-7", "=", "x", "=", "12co.dat", "12c+.dat",
            "12c.dat", "16o.dat", "=", "x", "=", "10.0", "10.0", "8000.0", "10.0", "30000.0", "0.005", "0.01", "=",
            "x", "=", "UNI", "1"]
    (tmp_path / "params.dat").write_text("\n".join(text) + "\n")
    p = model.read_params(str(tmp_path / "params.dat"))
    assert (p.input, p.output, p.itertot, p.zeta, p.Tmin, p.fieldchoice, p.Gext) == \
        ("1Dn40.dat", "B01", 6000, 1.3e-17, 10.0, "UNI", 1.0)
