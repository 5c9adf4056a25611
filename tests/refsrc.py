"""Read-only text access to the reference tree for CPU-side structure checks.

Nothing here compiles or imports reference code: the generated RHS files
(odes_*.c) are parsed as text into term lists so the network compiler can be
compared with the MakeRates output term by term.  Tests that use this module
skip when the reference tree is not mounted.
"""
import os
import re
from collections import defaultdict

REF_DIR = os.environ.get("PDR_REFERENCE_DIR", "/root/reference")


def have_reference() -> bool:
    return os.path.isdir(REF_DIR) and os.path.exists(os.path.join(REF_DIR, "odes_reduced.c"))


_TERM = re.compile(r"([+-])((?:\d+\*)?rate\[\d+\](?:\*(?:x\[\d+\]|x_e|n_H|zeta_H2|zeta_He|zeta_H))*)")


def _parse_sum(expr: str):
    """'-rate[1]*x[2]*n_H+2*rate[5]*x_e*n_H' -> list of (coef, r, factors, nH, zeta)."""
    expr = expr.strip()
    out = []
    pos = 0
    for m in _TERM.finditer(expr):
        if m.start() != pos:
            raise ValueError(f"unparsed text {expr[pos:m.start()]!r}")
        pos = m.end()
        sign = -1 if m.group(1) == "-" else 1
        parts = m.group(2).split("*")
        coef = 1
        if parts[0].isdigit():
            coef = int(parts[0])
            parts = parts[1:]
        r = int(parts[0][5:-1])
        facs, nh, zeta = [], 0, None
        for p in parts[1:]:
            if p == "n_H":
                nh += 1
            elif p == "x_e":
                facs.append("e")
            elif p.startswith("zeta"):
                zeta = p
            else:
                facs.append(int(p[2:-1]))
        out.append((sign * coef, r, facs, nh, zeta))
    if pos != len(expr):
        raise ValueError(f"trailing text {expr[pos:]!r}")
    return out


def parse_odes(net: str):
    """Return (form, loss, xe) for odes_<net>.c.

    form[i] / loss[i]: dict (r, factors-tuple, nH) -> signed coefficient, where
    loss terms are multiplied back by x[i] (ode[i] = form + x[i]*loss).
    Terms that carry an X-ray zeta factor are returned separately in zeta_r.
    xe: dict species index -> coefficient of the charge sum.
    """
    with open(os.path.join(REF_DIR, f"odes_{net}.c")) as fh:
        src = fh.read()
    body = src[src.index("The ODEs created by MakeRates begin here"):]
    xe_line = re.search(r"x_e = ([^;]*);", src).group(1)
    xe = defaultdict(int)
    for m in re.finditer(r"([+-]?)(\d+\*)?x\[(\d+)\]", xe_line):
        c = int(m.group(2)[:-1]) if m.group(2) else 1
        xe[int(m.group(3))] += -c if m.group(1) == "-" else c
    form, loss = defaultdict(dict), defaultdict(dict)
    zeta_r = set()
    cur_loss, cur_form = None, None
    for line in body.splitlines():
        line = line.strip()
        if line.startswith("loss = "):
            cur_loss = _parse_sum(line[len("loss = "):-1])
        elif line.startswith("form = "):
            cur_form = _parse_sum(line[len("form = "):-1])
        elif line.startswith("ode["):
            i = int(line[4:line.index("]")])
            assert line.endswith(f"= form+x[{i}]*loss;"), line
            for (c, r, facs, nh, z) in cur_form:
                if z:
                    zeta_r.add(r)
                    continue
                key = (r, tuple(sorted(facs, key=str)), nh)
                form[i][key] = form[i].get(key, 0) + c
            for (c, r, facs, nh, z) in cur_loss:
                if z:
                    zeta_r.add(r)
                    continue
                key = (r, tuple(sorted(facs + [i], key=str)), nh)
                loss[i][key] = loss[i].get(key, 0) + c
            cur_loss = cur_form = None
    return dict(form), dict(loss), dict(xe), zeta_r
