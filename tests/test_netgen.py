"""Network compiler checks.

* Against the reference's own generated RHS (odes_*.c, read as text only):
  every form and loss term, its coefficient and its factors must match
  MakeRates' output for all three networks.  The shipped reduced-network RHS
  was generated for a species order that differs from species_reduced.d in
  slots 3-5 (CH5+, O2, H2+ instead of H2+, CH5+, O2); the comparison maps
  those slots by identity and also checks that the unpermuted comparison
  fails, so the permutation is a fact and not a fudge.
* Against the reader semantics of read_rates.F90 / read_species.F90 on
  synthetic files (duplicates, null fields, H2 renormalisation, e- init).
* Element and charge conservation of every reaction that has ODE terms.
"""
import os
import re
import tempfile
from collections import Counter

import pytest

from pdrchem import netgen
from pdrchem.netgen import load_network
import refsrc

NETS = ["reduced", "full", "mynetwork"]

# reference ODE slot -> species-file slot for the shipped reduced network
REDUCED_PERM = {3: 4, 4: 5, 5: 3}


def my_terms(net, perm_inv=None):
    """form/loss dicts in the reference's gross (uncancelled) convention."""
    neq = net.neq
    idx = {s.name: s.index for s in net.species}
    form, loss = {}, {}

    def fx(k):
        if k == neq:
            return "e"
        if k == neq + 1:
            return None
        return perm_inv.get(k, k) if perm_inv else k

    for r in net.reactions:
        if r.rtype == netgen.T_XR:
            continue
        facs = [f for f in (fx(r.fa), fx(r.fb)) if f is not None]
        key = (r.index, tuple(sorted(facs, key=str)), r.nh)
        react = [x for x in (r.reactants[:2] if r.reactants[2] == "#" else r.reactants)
                 if x in idx and x != "e-"]
        for name, cnt in Counter(react).items():
            i = fx(idx[name])
            loss.setdefault(i, {})
            loss[i][key] = loss[i].get(key, 0) - cnt
        prods = [x for x in r.products if x in idx and x != "e-"]
        for name, cnt in Counter(prods).items():
            i = fx(idx[name])
            form.setdefault(i, {})
            form[i][key] = form[i].get(key, 0) + cnt
    return form, loss


needs_ref = pytest.mark.skipif(not refsrc.have_reference(), reason="reference tree not mounted")


@needs_ref
@pytest.mark.parametrize("name", NETS)
def test_terms_match_makerates_output(name):
    net = load_network(name)
    rf, rl, rxe, zeta_r = refsrc.parse_odes(name)
    inv = {v: k for k, v in REDUCED_PERM.items()} if name == "reduced" else None
    mf, ml = my_terms(net, inv)
    assert set(rf) | set(rl) == set(range(net.neq))
    # X-ray reactions (rate identically 0, calc_reac_rates.F90:265-268) carry
    # no terms in the compiled network; drop them from the reference side
    xr = {r.index for r in net.reactions if r.rtype == netgen.T_XR}
    for d in (rf, rl):
        for i in d:
            d[i] = {k: v for k, v in d[i].items() if k[0] not in xr}
    for i in range(net.neq):
        assert mf.get(i, {}) == rf.get(i, {}), f"{name} form of ode[{i}]"
        assert ml.get(i, {}) == rl.get(i, {}), f"{name} loss of ode[{i}]"
    # every zeta-weighted term belongs to an X-ray reaction (rate == 0)
    assert zeta_r <= xr
    # charge sum
    mine = {(inv or {}).get(i, i): c for i, c in enumerate(net.charges()) if c}
    assert mine == {k: v for k, v in rxe.items() if v}


@needs_ref
def test_reduced_permutation_is_real():
    """Of the 6 orderings of species-file slots 3-5, exactly one reproduces
    the shipped reduced RHS, and it is not the identity."""
    import itertools
    net = load_network("reduced")
    rf, rl, _, _ = refsrc.parse_odes("reduced")
    hits = []
    for perm in itertools.permutations([3, 4, 5]):
        ref_to_file = dict(zip([3, 4, 5], perm))
        inv = {v: k for k, v in ref_to_file.items()}
        mf, ml = my_terms(net, inv)
        if all(mf.get(i, {}) == rf.get(i, {}) and ml.get(i, {}) == rl.get(i, {}) for i in range(net.neq)):
            hits.append(ref_to_file)
    assert hits == [REDUCED_PERM]


@needs_ref
@pytest.mark.parametrize("name", NETS)
def test_tables_regenerate_identically(name, tmp_path):
    """The committed network tables are exactly what the compiler produces
    from the reference's species/rate files."""
    netgen.generate_all(refsrc.REF_DIR, str(tmp_path))
    committed = os.path.join(os.path.dirname(netgen.__file__), "..", "networks", f"{name}.tab")
    assert open(tmp_path / f"{name}.tab").read() == open(committed).read()


def _write(d, fname, text):
    p = os.path.join(d, fname)
    with open(p, "w") as fh:
        fh.write(text)
    return p


def test_reader_semantics_synthetic():
    with tempfile.TemporaryDirectory() as d:
        sp = _write(d, "s.d", "1,H,0.5,1.0\n2,H2,0.1,2.0\n3,H+,0.0,1.0\n4,C+,1.0E-4,12.0\n"
                              "5,Mg+,2.0D-5,24\n6,C,0,12\n7,e-,0.00E+00,0.0\n")
        species, labels = netgen.read_species(sp)
        ab = {s.name: s.abundance for s in species}
        # H + H+ + 2 H2 = 0.7 != 1  ->  H2 = (1 - 0.5 - 0)/2
        assert ab["H2"] == 0.25
        assert ab["e-"] == 1.0e-4 + 2.0e-5
        assert labels["NH"] == 1 and labels["NELECT"] == 7 and labels["NMGx"] == 5
        assert [s.charge for s in species] == [0, 0, 1, 1, 1, 0, 0]
        rt = _write(d, "r.d",
                    "1,H,C+,,C,H+,,,1.0E-10,0.5,10.0,L,10,100,A,\n"
                    "2,H,C+,,C,H+,,,2.0E-10,0.0,20.0,L,50,300,A,\n"
                    "3,H,C+,,C,H+,,,3.0E-10,0.0,30.0,L,200,1000,A,\"1/2\"\n"
                    "4,C,PHOTON,,C+,e-,,,3.0E-10,,3.0,,,,\n"
                    "5,H,H,#,H2,#,,,5.2E-17,0.5,0.0,,10,41000,,\n")
        rs = netgen.read_rates(rt)
        assert [r.duplicate for r in rs] == [1, 2, 3, 0, 0]
        assert [r.rtmin for r in rs[:3]] == [10.0, 100.0, 300.0]
        assert rs[3].beta == 0.0 and rs[3].rtmin == 0.0 and rs[3].rtmax == 0.0
        assert rs[3].reactants == ["C", "PHOTON", ""]
        for r in rs:
            r.rtype = netgen.classify(r)
        assert [r.rtype for r in rs] == [netgen.T_THERMAL] * 3 + [netgen.T_PHOTO_CI, netgen.T_H2FORM]
        netgen.build_ode_terms(species, rs)
        assert rs[4].stoich == {0: -2, 1: 1} and rs[4].nh == 1 and rs[4].fb == 7
        assert rs[3].stoich == {5: -1, 3: 1} and rs[3].nh == 0 and rs[3].fa == 5


_ELEMENTS = ["He", "Mg", "Na", "Si", "Fe", "Cl", "Ca", "H", "D", "C", "N", "O", "S", "F", "P"]


def composition(name):
    """Element counts of a species label (e.g. 'CH3OH2+' -> C1 H5 O1)."""
    s = name.rstrip("+-")
    if s.startswith("#"):
        s = s[1:]
    s = s.replace("*", "")
    out = Counter()
    pos = 0
    while pos < len(s):
        for el in sorted(_ELEMENTS, key=len, reverse=True):
            if s.startswith(el, pos) or s.upper().startswith(el.upper(), pos) and len(el) == 2 and el != "Cl" and s[pos:pos + 2] in ("HE", "MG", "NA", "SI", "FE", "CL", "CA"):
                pos += len(el)
                m = re.match(r"\d+", s[pos:])
                n = 1
                if m:
                    n = int(m.group(0))
                    pos += len(m.group(0))
                out[el] += n
                break
        else:
            return None
    return out


@pytest.mark.parametrize("name", NETS)
def test_element_and_charge_conservation(name):
    net = load_network(name)
    comp = {}
    for s in net.species:
        if s.name.startswith("PAH") or s.name == "e-":
            continue
        c = composition(s.name)
        if c is not None:
            comp[s.index] = c
    checked = 0
    for r in net.reactions:
        if r.rtype == netgen.T_XR or not r.stoich:
            continue
        if any(k not in comp for k in r.stoich):
            continue
        # element balance through the net stoichiometry
        tot = Counter()
        for k, v in r.stoich.items():
            for el, n in comp[k].items():
                tot[el] += v * n
        assert all(v == 0 for v in tot.values()), (r.number, r.reactants, r.products, tot)
        # charge balance: species charges plus electrons produced/consumed
        q = sum(v * net.species[k].charge for k, v in r.stoich.items())
        ne = r.products.count("e-") - r.reactants.count("e-")
        assert q - ne == 0, (r.number, r.reactants, r.products)
        checked += 1
    assert checked > 0.9 * sum(1 for r in net.reactions if r.rtype != netgen.T_XR)
