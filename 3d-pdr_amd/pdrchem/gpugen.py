"""HIP code generator for the batched BDF kernels.

Reads a compiled network (networks/<name>.json, written by netgen.py from the
reference's species_*.d / rates_*.d) and emits ``csrc/gen/net_<name>.h``: the
constants and lane tables that specialise the wavefront-per-system BDF kernel
(csrc/bdf_wave.hip) to that network.  This replaces the MakeRates-generated
odes_*.c / jacobian_*.c of the reference (odes_reduced.c:26-148,
jacobian_reduced.c:26-...) with table-driven device functions laid out for a
64-lane wavefront:

* Rate evaluation (calc_reac_rates.F90:16-401): per-reaction parameters.
* RHS: every reaction with a non-zero net stoichiometry owns a flux slot
  s = round*64 + lane; flux_s = k_s * X[a_s] * X[b_s] (* n_H) is computed
  lane-parallel.  ydot_i = sum_s S_is flux_s is split into per-lane chunks:
  every species row receives a number of lanes proportional to its term
  count (greedy min-max), each lane sums one contiguous chunk of one row, and
  the row total is the sum of that row's lane partials.
* Jacobian (exact, with d x_e / d x_j = charge_j, which jacobian_*.c omits):
  the wave layout holds row i in lanes i and i+32 (columns 0-15 and 16-31).
  Each lane accumulates its 16 entries from per-column term lists over the
  flux partials D[2s] = dflux_s/dX[a_s], D[2s+1] = dflux_s/dX[b_s]; a 17th
  pseudo-column accumulates g_i = d ydot_i / d x_e, applied as g_i*charge_j.
"""
from __future__ import annotations

import math
import struct
import os
import sys
from typing import List

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from pdrchem.netgen import load_network, Network, T_XR  # type: ignore
else:
    from .netgen import load_network, Network, T_XR

WAVE = 64


def _fmt_list(vals, per_line=16):
    out = []
    for i in range(0, len(vals), per_line):
        out.append("  " + ", ".join(vals[i:i + per_line]) + ",")
    return "\n".join(out)


def _dbl(v: float) -> str:
    if v == 0.0:
        return "0.0"
    return repr(float(v))


def emit_rate_params(L: List[str], net: Network) -> None:
    """Per-reaction rate parameters (calc_reac_rates.F90) for the device rate
    kernel; appended to the network header inside its namespace."""
    rs = net.reactions
    L.append("// rate evaluation parameters (calc_reac_rates.F90)")
    for nm, get in (("alpha", lambda r: r.alpha), ("beta", lambda r: r.beta),
                    ("gamma", lambda r: r.gamma), ("rtmin", lambda r: r.rtmin),
                    ("rtmax", lambda r: r.rtmax)):
        L.append(f"__device__ const double {nm}[NREAC] = {{")
        L.append(_fmt_list([_dbl(get(r)) for r in rs], 6))
        L.append("};")
    L.append("__device__ const uint8_t rtype[NREAC] = {")
    L.append(_fmt_list([str(r.rtype) for r in rs], 24))
    L.append("};")
    L.append("__device__ const uint8_t dup[NREAC] = {")
    L.append(_fmt_list([str(min(r.duplicate, 255)) for r in rs], 24))
    L.append("};")
    L.append("__device__ const uint8_t gflag[NREAC] = {")
    L.append(_fmt_list([str(r.gflag) for r in rs], 24))
    L.append("};")
    # Duplicate-window selection as a per-member test: the Fortran loop
    # (calc_reac_rates.F90:113-144) picks the first member with
    # T <= RTMAX, else the last member.  Member j is therefore selected iff
    # T > max(RTMAX of earlier members) and (T <= RTMAX_j or j is last).
    prev_max, last = [], []
    for idx, r in enumerate(rs):
        if r.duplicate == 0:
            prev_max.append("-1.0e300")
            last.append("1")
            continue
        lead = idx - (r.duplicate - 1)
        pm = max((rs[m].rtmax for m in range(lead, idx)), default=None)
        prev_max.append("-1.0e300" if pm is None else _dbl(pm))
        nxt = rs[idx + 1].duplicate if idx + 1 < len(rs) else 0
        last.append("1" if nxt < r.duplicate else "0")
    L.append("__device__ const double dup_prev_max[NREAC] = {")
    L.append(_fmt_list(prev_max, 8))
    L.append("};")
    L.append("__device__ const uint8_t dup_last[NREAC] = {")
    L.append(_fmt_list(last, 24))
    L.append("};")


def coef_hi16(c: int) -> int:
    """Top 16 bits of the FP64 high word of a small integer coefficient; its
    other 48 bits are zero for |c| <= 8, so (hi16 << 48) is the double."""
    bits = struct.unpack("<Q", struct.pack("<d", float(c)))[0]
    if bits & ((1 << 48) - 1):
        raise ValueError(f"coefficient {c} is not exact in 16 bits")
    return bits >> 48


def rhs_word(slot: int, c: int) -> int:
    """RHS term word: LDS byte offset of flux slot `slot` in the low half,
    the coefficient as coef_hi16 in the high half."""
    return (8 * slot) | (coef_hi16(c) << 16)


class WaveTables:
    """Lane tables of the wavefront-per-system layout (N = 32 rows)."""

    def __init__(self, net: Network):
        self.net = net
        neq = net.neq
        if neq > 32:
            raise ValueError(f"{net.name}: wavefront layout supports up to 32 equations, got {neq}")
        self.N = 32
        # ---- flux slots: reactions with terms
        self.slot_reac: List[int] = [r.index for r in net.reactions if r.stoich]
        nact = len(self.slot_reac)
        self.rounds = max(1, math.ceil(nact / WAVE))
        self.nslot = self.rounds * WAVE
        # the RHS reads fluxes F[0..nslot) (FP64) and the zero F[nslot]; the
        # Jacobian reads partials D[0..2*nslot) (matrix type) and the zero
        # D[2*nslot], in the same LDS region
        self.zero_slot = self.nslot
        reac_slot = {r: s for s, r in enumerate(self.slot_reac)}
        self.slot_fab = []
        for s in range(self.nslot):
            if s < nact:
                r = net.reactions[self.slot_reac[s]]
                self.slot_fab.append(r.fa | (r.fb << 8) | (r.nh << 16) | (1 << 17))
            else:
                self.slot_fab.append((neq + 1) | ((neq + 1) << 8))  # 1*1, rate 0
        # ---- RHS rows -> (slot, coef) term lists
        rows = [[] for _ in range(neq)]
        for r in net.reactions:
            if not r.stoich:
                continue
            s = reac_slot[r.index]
            for k, v in sorted(r.stoich.items()):
                rows[k].append((s, v))
        self.row_terms = rows
        # greedy lane allocation (each row at least one lane)
        nl = [1] * neq
        extra = WAVE - neq
        load = lambda i: math.ceil(len(rows[i]) / nl[i]) if rows[i] else 0
        for _ in range(extra):
            i = max(range(neq), key=load)
            if load(i) <= 1:
                break
            nl[i] += 1
        self.row_nlanes = nl
        base, b = [], 0
        for i in range(neq):
            base.append(b)
            b += nl[i]
        self.row_base = base
        self.lanes_used = b
        self.tm = max(load(i) for i in range(neq))
        self.pm = max(nl)
        # per-lane chunk lists
        self.rhs_term = [[(self.zero_slot, 0)] * WAVE for _ in range(self.tm)]
        for i in range(neq):
            terms = rows[i]
            k = nl[i]
            csz = math.ceil(len(terms) / k) if terms else 0
            for u in range(k):
                chunk = terms[u * csz:(u + 1) * csz]
                lane = base[i] + u
                for t, (s, c) in enumerate(chunk):
                    self.rhs_term[t][lane] = (s, c)
        # ---- Jacobian term lists per (row i, column c) and electron column
        ncol = 16
        self.jzero = 2 * self.nslot           # D array has one trailing zero
        jlists = {}                           # (lane, j) -> [(dslot, coef)]
        for lane in range(WAVE):
            for j in range(ncol + 1):
                jlists[(lane, j)] = []
        for r in net.reactions:
            if not r.stoich:
                continue
            s = reac_slot[r.index]
            for i, coef in r.stoich.items():
                for which, f in ((0, r.fa), (1, r.fb)):
                    d = 2 * s + which
                    if f < neq:
                        lane = i + 32 * (f // 16)
                        jlists[(lane, f % 16)].append((d, coef))
                    elif f == neq:  # electron factor -> pseudo column, both halves
                        jlists[(i, ncol)].append((d, coef))
                        jlists[(i + 32, ncol)].append((d, coef))
        self.jt = []
        for j in range(ncol + 1):
            self.jt.append(max(len(jlists[(lane, j)]) for lane in range(WAVE)))
        self.jac_term = []  # flattened [j][t][lane]
        for j in range(ncol + 1):
            for t in range(self.jt[j]):
                row = []
                for lane in range(WAVE):
                    lst = jlists[(lane, j)]
                    row.append(lst[t] if t < len(lst) else (self.jzero, 0))
                self.jac_term.append(row)
        self.charges = net.charges() + [0] * (self.N - neq)

    def emit(self, path: str) -> None:
        net = self.net
        ns = f"pdrnet_{net.name}"
        L = []
        L.append(f"// Generated by pdrchem/gpugen.py from networks/{net.name}.json; do not edit.")
        L.append("// Network constants and lane tables for the wavefront-per-system BDF kernel.")
        L.append("#pragma once")
        L.append("#include <stdint.h>")
        L.append(f"namespace {ns} {{")
        L.append(f"constexpr int NSPEC = {net.nspec};")
        L.append(f"constexpr int NREAC = {net.nreac};")
        L.append(f"constexpr int NEQ = {net.neq};")
        L.append(f"constexpr int N = {self.N};")
        L.append(f"constexpr int ROUNDS = {self.rounds};      // flux slot rounds of 64 lanes")
        L.append(f"constexpr int NSLOT = {self.nslot};")
        L.append(f"constexpr int NACTIVE = {len(self.slot_reac)};")
        L.append(f"constexpr int TM = {self.tm};          // RHS terms per lane")
        L.append(f"constexpr int PM = {self.pm};           // lanes per species row (max)")
        L.append(f"constexpr int XLAMBDA_SPLINE = {net.xlambda_spline};")
        L.append("constexpr int JT[17] = {" + ", ".join(str(x) for x in self.jt) + "};")
        L.append(f"constexpr int JT_TOTAL = {sum(self.jt)};")
        L.append("// slot -> reaction index (-1: padding)")
        L.append(f"__device__ const int16_t slot_reac[NSLOT] = {{")
        L.append(_fmt_list([str(x) for x in self.slot_reac] + ["-1"] * (self.nslot - len(self.slot_reac))))
        L.append("};")
        L.append("// slot -> fa | fb<<8 | nH<<16 | active<<17 (factor indices into X)")
        L.append(f"__device__ const uint32_t slot_fab[NSLOT] = {{")
        L.append(_fmt_list([f"0x{x:x}u" for x in self.slot_fab], 8))
        L.append("};")
        L.append("// [t][lane]: byte offset of the flux slot (8 * slot) | top 16 bits of the")
        L.append("// coefficient's FP64 high word << 16 (the rest of the double is zero)")
        flat = []
        for t in range(self.tm):
            for lane in range(WAVE):
                flat.append(f"0x{rhs_word(*self.rhs_term[t][lane]):x}u")
        L.append(f"__device__ const uint32_t rhs_term[TM * 64] = {{")
        L.append(_fmt_list(flat, 8))
        L.append("};")
        rb = [self.row_base[i % 32] if (i % 32) < net.neq else 0 for i in range(WAVE)]
        rn = [self.row_nlanes[i % 32] if (i % 32) < net.neq else 0 for i in range(WAVE)]
        L.append("// row layout lane (i = lane & 31): first partial lane and count of row i")
        L.append("__device__ const uint8_t row_base[64] = {")
        L.append(_fmt_list([str(x) for x in rb]))
        L.append("};")
        L.append("__device__ const uint8_t row_nlanes[64] = {")
        L.append(_fmt_list([str(x) for x in rn]))
        L.append("};")
        flat = []
        for row in self.jac_term:
            for (d, c) in row:
                flat.append(f"0x{(d & 0xffff) | ((c & 0xffff) << 16):x}u")
        cols = [j for j in range(len(self.jt)) for _ in range(self.jt[j])]
        L.append("// column (16 = d/dx_e) of each round of jac_term")
        L.append(f"constexpr int8_t jac_col[(JT_TOTAL > 0 ? JT_TOTAL : 1)] = {{")
        L.append(_fmt_list([str(x) for x in cols] if cols else ["0"]))
        L.append("};")
        L.append("// [column j (16 = d/dx_e)][t][lane]: D index | (int16 coefficient) << 16")
        L.append(f"__device__ const uint32_t jac_term[(JT_TOTAL > 0 ? JT_TOTAL : 1) * 64] = {{")
        L.append(_fmt_list(flat if flat else ["0u"] * 64, 8))
        L.append("};")
        L.append("__device__ const int8_t charge[N] = {")
        L.append(_fmt_list([str(x) for x in self.charges]))
        L.append("};")
        emit_rate_params(L, net)
        L.append(f"}}  // namespace {ns}")
        with open(path, "w") as fh:
            fh.write("\n".join(L) + "\n")

    def summary(self) -> str:
        return (f"{self.net.name}: slots={len(self.slot_reac)}/{self.nslot} TM={self.tm} PM={self.pm} "
                f"lanes={self.lanes_used} JT={self.jt} JT_TOTAL={sum(self.jt)}")


# Shielding / extinction tables (constant data, identical by necessity to the
# reference's): Savage & Mathis tau(lambda)/tau(V) (modules.F90:354-367) and
# van Dishoeck & Black log10 CO shielding (3DPDR.F90:1532-1537).
L_GRID = [910.0, 950.0, 1000.0, 1050.0, 1110.0, 1180.0, 1250.0, 1390.0, 1490.0, 1600.0,
          1700.0, 1800.0, 1900.0, 2000.0, 2100.0, 2190.0, 2300.0, 2400.0, 2500.0, 2740.0,
          3440.0, 4000.0, 4400.0, 5500.0, 7000.0, 9000.0, 12500.0, 22000.0, 34000.0, 1.0e9]
X_GRID = [5.76, 5.18, 4.65, 4.16, 3.73, 3.40, 3.11, 2.74, 2.63, 2.62,
          2.54, 2.50, 2.58, 2.78, 3.01, 3.12, 2.86, 2.58, 2.35, 2.00,
          1.58, 1.42, 1.32, 1.00, 0.75, 0.48, 0.28, 0.12, 0.05, 0.00]
NCO_GRID = [12.0, 13.0, 14.0, 15.0, 16.0, 17.0, 18.0, 19.0]
NH2_GRID = [18.0, 19.0, 20.0, 21.0, 22.0, 23.0]
SCO_BY_NH2 = [  # SCO_GRID(1:8, j) for j = 1..6
    [0.000e+00, -1.408e-02, -1.099e-01, -4.400e-01, -1.154e+00, -1.888e+00, -2.760e+00, -4.001e+00],
    [-8.539e-02, -1.015e-01, -2.104e-01, -5.608e-01, -1.272e+00, -1.973e+00, -2.818e+00, -4.055e+00],
    [-1.451e-01, -1.612e-01, -2.708e-01, -6.273e-01, -1.355e+00, -2.057e+00, -2.902e+00, -4.122e+00],
    [-4.559e-01, -4.666e-01, -5.432e-01, -8.665e-01, -1.602e+00, -2.303e+00, -3.146e+00, -4.421e+00],
    [-1.303e+00, -1.312e+00, -1.367e+00, -1.676e+00, -2.305e+00, -3.034e+00, -3.758e+00, -5.077e+00],
    [-3.883e+00, -3.888e+00, -3.936e+00, -4.197e+00, -4.739e+00, -5.165e+00, -5.441e+00, -6.446e+00],
]


def natural_spline_m(x, y):
    """Second derivatives of the natural cubic spline (tridiagonal
    elimination in the normalised form s*m[i-1] + 2 m[i] + (1-s) m[i+1])."""
    n = len(x)
    c, d, m = [0.0] * n, [0.0] * n, [0.0] * n
    for i in range(1, n - 1):
        s = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        piv = s * c[i - 1] + 2.0
        c[i] = (s - 1.0) / piv
        rhs = 6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) / (
            x[i + 1] - x[i - 1])
        d[i] = (rhs - s * d[i - 1]) / piv
    for i in range(n - 2, -1, -1):
        m[i] = c[i] * m[i + 1] + d[i]
    return m


def emit_shield_tables(path: str) -> None:
    sco = [[SCO_BY_NH2[j][i] for j in range(6)] for i in range(8)]  # [nco][nh2]
    xd = natural_spline_m(L_GRID, X_GRID)
    scod = [natural_spline_m(NH2_GRID, row) for row in sco]
    L = ["// Generated by pdrchem/gpugen.py; do not edit.",
         "// Shielding tables for the device rate evaluation (shield.F90).",
         "#pragma once", "namespace shieldtab {"]
    L.append("__device__ const double L_GRID[30] = {\n" + _fmt_list([_dbl(v) for v in L_GRID], 6) + "\n};")
    L.append("__device__ const double X_GRID[30] = {\n" + _fmt_list([_dbl(v) for v in X_GRID], 6) + "\n};")
    L.append("__device__ const double X_DERIV[30] = {\n" + _fmt_list([_dbl(v) for v in xd], 4) + "\n};")
    L.append("__device__ const double X_ZERO[30] = {0.0};")
    L.append("__device__ const double NCO_GRID[8] = {\n" + _fmt_list([_dbl(v) for v in NCO_GRID]) + "\n};")
    L.append("__device__ const double NH2_GRID[6] = {\n" + _fmt_list([_dbl(v) for v in NH2_GRID]) + "\n};")
    L.append("__device__ const double SCO_GRID[8][6] = {\n" +
             "\n".join("  {" + ", ".join(_dbl(v) for v in row) + "}," for row in sco) + "\n};")
    L.append("__device__ const double SCO_DERIV[8][6] = {\n" +
             "\n".join("  {" + ", ".join(_dbl(v) for v in row) + "}," for row in scod) + "\n};")
    L.append("}  // namespace shieldtab")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def generate(names=("reduced",), net_dir=None, out_dir=None) -> List[WaveTables]:
    base = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    net_dir = net_dir or os.path.join(base, "networks")
    out_dir = out_dir or os.path.join(base, "csrc", "gen")
    os.makedirs(out_dir, exist_ok=True)
    emit_shield_tables(os.path.join(out_dir, "shield_tables.h"))
    out = []
    for n in names:
        wt = WaveTables(load_network(n, net_dir))
        wt.emit(os.path.join(out_dir, f"net_{n}.h"))
        out.append(wt)
    return out


if __name__ == "__main__":
    for wt in generate():
        print(wt.summary())
