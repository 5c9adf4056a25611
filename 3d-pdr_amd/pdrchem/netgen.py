"""Network compiler: reads a 3D-PDR species file and rate file and produces the
network description used by the CPU oracle and the HIP solver.

Reader semantics restated from the reference:
  * species file  -> READ_SPECIES  (read_species.F90:91-190)
  * rate file     -> READ_RATES    (read_rates.F90:53-91)
  * reaction-type dispatch for rate evaluation follows
    CALCULATE_REACTION_RATES (calc_reac_rates.F90:53-67, 75-100, 151-230)
  * ODE term conventions follow the MakeRates output shipped as odes_*.c
    (odes_reduced.c:26-148): two-body flux k*x_a*x_b*n_H, first-order flux
    k*x_a for PHOTON/CRP/CRPHOT-type reactions, grain H2 formation
    H + H + # -> H2 written as k*x_H*n_H, electrons enter through the
    charge sum x_e = sum(charge_i * x_i) and are not integrated.

The output is a plain text table (``<name>.tab``) read by the C oracle and by
the HIP code generator, plus a JSON description for the Python host.
"""
from __future__ import annotations

import dataclasses
import json
import os
from typing import Dict, List, Optional, Sequence

# --------------------------------------------------------------------------
# Reaction type codes (shared with oracle/pdr_oracle.h and csrc/pdr_types.h)
# --------------------------------------------------------------------------
T_THERMAL = 0   # Arrhenius with duplicate T windows  calc_reac_rates.F90:99-144
T_H2FORM = 1    # grain H2 formation                  calc_reac_rates.F90:75-87
T_PAH = 2       # PAH reactions                       calc_reac_rates.F90:91-96
T_PHOTO = 3     # generic photo reaction               calc_reac_rates.F90:214-239
T_PHOTO_H2 = 4  # H2 photodissociation (H2PDRATE)      calc_reac_rates.F90:151-159
T_PHOTO_HD = 5  # HD photodissociation (H2PDRATE)      calc_reac_rates.F90:163-171
T_PHOTO_CO = 6  # CO photodissociation (COPDRATE)      calc_reac_rates.F90:175-185
T_PHOTO_CI = 7  # C photoionisation (CIPDRATE)         calc_reac_rates.F90:189-199
T_PHOTO_SI = 8  # S photoionisation (SIPDRATE)         calc_reac_rates.F90:203-212
T_CRP = 9       # cosmic-ray ionisation                calc_reac_rates.F90:246-263
T_CRPHOT = 10   # CR-induced photons                   calc_reac_rates.F90:274-294
T_FREEZE = 11   # freeze-out of neutrals               calc_reac_rates.F90:300-310
T_ELFRZE = 12   # freeze-out of ions                   calc_reac_rates.F90:316-326
T_CRH = 13      # CR-heating desorption                calc_reac_rates.F90:335-342
T_PHOTD = 14    # photodesorption                      calc_reac_rates.F90:348-364
T_THERM = 15    # thermal desorption                   calc_reac_rates.F90:369-371
T_GRAIN = 16    # grain mantle reaction (R2 = #...)    calc_reac_rates.F90:377
T_XR = 17       # X-ray types, rate always 0           calc_reac_rates.F90:265-268

TYPE_NAMES = {
    T_THERMAL: "THERMAL", T_H2FORM: "H2FORM", T_PAH: "PAH", T_PHOTO: "PHOTO",
    T_PHOTO_H2: "PHOTO_H2", T_PHOTO_HD: "PHOTO_HD", T_PHOTO_CO: "PHOTO_CO",
    T_PHOTO_CI: "PHOTO_CI", T_PHOTO_SI: "PHOTO_SI", T_CRP: "CRP",
    T_CRPHOT: "CRPHOT", T_FREEZE: "FREEZE", T_ELFRZE: "ELFRZE", T_CRH: "CRH",
    T_PHOTD: "PHOTD", T_THERM: "THERM", T_GRAIN: "GRAIN", T_XR: "XR",
}

# reaction types whose rate is chosen among duplicates by temperature window
DUPLICATE_AWARE = (T_THERMAL, T_PHOTO, T_CRP, T_CRPHOT)

NAME_LEN = 10  # CHARACTER(len=10) in read_species.F90 / read_rates.F90


class NetworkError(ValueError):
    pass


def _fortran_fields(line: str, nitems: int) -> List[Optional[str]]:
    """Split one record the way a Fortran list-directed READ of ``nitems``
    items does for these CSV files: values separated by commas (blanks
    around them ignored), an empty field is a null value (the variable keeps
    its previous value), items beyond ``nitems`` are ignored.  Records that
    would need the r*c repeat form, quotes or a slash terminator before the
    last wanted item are rejected, because the shipped files use none."""
    raw = line.rstrip("\r\n").split(",")
    out: List[Optional[str]] = []
    for f in raw[:nitems]:
        v = f.strip()
        if v == "":
            out.append(None)
            continue
        if any(c in v for c in " \t'\"/"):
            raise NetworkError(f"unsupported list-directed token {v!r}")
        head = v.split("*", 1)[0]
        if "*" in v and head.isdigit():
            raise NetworkError(f"repeat count form not supported: {v!r}")
        out.append(v)
    while len(out) < nitems:
        out.append(None)
    return out


def _fortran_real(tok: Optional[str], prev: float) -> float:
    if tok is None:
        return prev
    return float(tok.replace("D", "E").replace("d", "e"))


def _name(tok: Optional[str], prev: str = "") -> str:
    if tok is None:
        return prev
    return tok[:NAME_LEN]


def species_charge(name: str) -> int:
    """Charge from the species label: one per '+' minus one per '-'."""
    return name.count("+") - name.count("-")


@dataclasses.dataclass
class Species:
    index: int
    name: str
    abundance: float
    mass: float
    charge: int


@dataclasses.dataclass
class Reaction:
    index: int               # 0-based position in the rate file
    number: int              # the N column
    reactants: List[str]     # 3 names, '' for blank
    products: List[str]      # 4 names
    alpha: float
    beta: float
    gamma: float
    rtmin: float
    rtmax: float
    duplicate: int
    rtype: int = T_THERMAL
    gflag: int = 0           # reactant 1 starts with 'G': no clamp at 1
    fa: int = 0              # flux factor indices into X (see Network.x_index)
    fb: int = 0
    nh: int = 0              # 1: flux carries one power of n_H
    stoich: Dict[int, int] = dataclasses.field(default_factory=dict)


@dataclasses.dataclass
class Network:
    name: str
    species: List[Species]
    reactions: List[Reaction]
    labels: Dict[str, int]    # 1-based species labels (0 = absent) as READ_SPECIES
    nr_indices: Dict[str, int]  # NRGR, NRH2, ... (1-based, 0 = absent)
    xlambda_spline: int = 0   # see xlambda_mode()

    @property
    def nspec(self) -> int:
        return len(self.species)

    @property
    def nreac(self) -> int:
        return len(self.reactions)

    @property
    def neq(self) -> int:
        return self.nspec - 1

    # X array used by the flux tables: X[0..neq-1] = abundances,
    # X[neq] = x_e, X[neq+1] = 1.0
    @property
    def x_electron(self) -> int:
        return self.neq

    @property
    def x_one(self) -> int:
        return self.neq + 1

    def charges(self) -> List[int]:
        return [s.charge for s in self.species[: self.neq]]


# --------------------------------------------------------------------------
# the reference's own slot numbering
# --------------------------------------------------------------------------

# The reference's shipped reduced-network ODE code (odes_reduced.c,
# jacobian_reduced.c) numbers species slots 3-5 as CH5+, O2, H2+ where
# species_reduced.d lists H2+, CH5+, O2: the loss terms of its ode[4] are O2's
# reactions and its electron sum counts x[5], not x[4].  Its arrays therefore
# hold, at the position of each species-file slot, the abundance of the
# species-file slot mapped here; every reader of those arrays by label (the
# heating's ABUNDANCE(NH2x), the output files' columns) sees that species.
REFERENCE_SLOTS = {"reduced": {3: 4, 4: 5, 5: 3}}


def reference_slot(network: str, slot: int) -> int:
    """the species-file slot (0-based) whose abundance the reference's
    arrays hold at position `slot`"""
    return REFERENCE_SLOTS.get(network, {}).get(slot, slot)


# --------------------------------------------------------------------------
# readers
# --------------------------------------------------------------------------

_LABELS = {
    "H": "NH", "D": "ND", "H2": "NH2", "HD": "NHD", "H2+": "NH2x", "H3+": "NH3x",
    "H+": "NPROTON", "C": "NC", "C+": "NCx", "O": "NO", "O+": "NOx", "N": "NN",
    "N+": "NNx", "S": "NS", "S+": "NSx", "He": "NHE", "HE": "NHE", "He+": "NHEx",
    "HE+": "NHEx", "Na": "NNA", "NA": "NNA", "Na+": "NNAx", "NA+": "NNAx",
    "Mg": "NMG", "MG": "NMG", "Mg+": "NMGx", "MG+": "NMGx", "Si": "NSI",
    "SI": "NSI", "Si+": "NSIx", "SI+": "NSIx", "Fe": "NFE", "FE": "NFE",
    "Fe+": "NFEx", "FE+": "NFEx", "Cl": "NCL", "CL": "NCL", "Cl+": "NCLx",
    "CL+": "NCLx", "Ca": "NCA", "CA": "NCA", "Ca+": "NCAx", "CA+": "NCAx",
    "Ca++": "NCAxx", "CA++": "NCAxx", "CO": "NCO", "CH": "NCH", "CH2": "NCH2",
    "OH": "NOH", "O2": "NO2", "CS": "NCS", "H2O": "NH2O", "H3O+": "NH3Ox",
    "HCO+": "NHCOx", "e-": "NELECT", "ELECTR": "NELECT",
}
_LABEL_NAMES = sorted(set(_LABELS.values()))


def read_species(path: str):
    """READ_SPECIES (read_species.F90:15-190): returns species list (with the
    H2 renormalisation and initial electron abundance applied) and the
    1-based index labels."""
    labels = {k: 0 for k in _LABEL_NAMES}
    recs = []
    with open(path) as fh:
        for line in fh:
            if not line.strip():
                continue
            idx, name, ab, ms = _fortran_fields(line, 4)
            recs.append([int(idx), _name(name), _fortran_real(ab, 0.0), _fortran_real(ms, 0.0)])
    for i, (_, name, _, _) in enumerate(recs, start=1):
        lab = _LABELS.get(name)
        if lab is not None:
            labels[lab] = i
    ab = [r[2] for r in recs]
    nh, npr, nh2, nel = labels["NH"], labels["NPROTON"], labels["NH2"], labels["NELECT"]
    if nh == 0 or npr == 0 or nh2 == 0 or nel == 0:
        raise NetworkError("species file lacks one of H, H+, H2, e-")
    # read_species.F90:172-174 (exact floating comparison, as in Fortran)
    if (ab[nh - 1] + ab[npr - 1] + 2.0 * ab[nh2 - 1]) != 1.0:
        ab[nh2 - 1] = 0.5 * (1.0 - ab[nh - 1] - ab[npr - 1])
    # read_species.F90:178-190
    if ab[nel - 1] <= 0.0:
        s = 0.0
        for lab, mult in (("NCx", 1.0), ("NSx", 1.0), ("NNAx", 1.0), ("NMGx", 1.0),
                          ("NSIx", 1.0), ("NFEx", 1.0), ("NCLx", 1.0), ("NCAx", 1.0),
                          ("NCAxx", 2.0), ("NPROTON", 1.0)):
            if labels[lab] != 0:
                s = s + mult * ab[labels[lab] - 1]
        ab[nel - 1] = s
    species = [Species(i, r[1], ab[i], r[3], 0 if r[1] == "e-" else species_charge(r[1]))
               for i, r in enumerate(recs)]
    return species, labels


def read_rates(path: str) -> List[Reaction]:
    """READ_RATES (read_rates.F90:9-91) including duplicate grouping and the
    RTMIN adjustment that makes duplicate temperature windows adjacent."""
    out: List[Reaction] = []
    with open(path) as fh:
        lines = [l for l in fh if l.strip()]
    for i, line in enumerate(lines):
        f = _fortran_fields(line, 14)
        num = int(f[0])
        reac = [_name(x) for x in f[1:4]]
        prod = [_name(x) for x in f[4:8]]
        r = Reaction(index=i, number=num, reactants=reac, products=prod,
                     alpha=_fortran_real(f[8], 0.0), beta=_fortran_real(f[9], 0.0),
                     gamma=_fortran_real(f[10], 0.0), rtmin=_fortran_real(f[12], 0.0),
                     rtmax=_fortran_real(f[13], 0.0), duplicate=0)
        if i > 0:
            p = out[-1]
            if p.reactants == r.reactants and p.products == r.products:
                if p.duplicate == 0:
                    p.duplicate = 1
                r.duplicate = p.duplicate + 1
                r.rtmin = p.rtmax
        out.append(r)
    return out


# --------------------------------------------------------------------------
# classification and ODE structure
# --------------------------------------------------------------------------

def classify(r: Reaction) -> int:
    """Reaction-type dispatch of calc_reac_rates.F90:53-67 and the thermal /
    photo sub-branches (:75-96, :151-212)."""
    r1, r2, r3 = r.reactants
    p = r.products
    if r2 == "PHOTON":
        if r1 == "H2" and r3 == "":
            return T_PHOTO_H2
        if r1 == "HD" and r3 == "":
            return T_PHOTO_HD
        if r1 == "CO" and r3 == "" and "C" in p and "O" in p:
            return T_PHOTO_CO
        if r1 == "C" and r3 == "" and ((p[0] == "C+" and p[1] == "e-") or (p[0] == "e-" and p[1] == "C+")):
            return T_PHOTO_CI
        if r1 == "S" and r3 == "" and ((p[0] == "S+" and p[1] == "e-") or (p[0] == "e-" and p[1] == "S+")):
            return T_PHOTO_SI
        return T_PHOTO
    simple = {"CRP": T_CRP, "CRPHOT": T_CRPHOT, "FREEZE": T_FREEZE, "ELFRZE": T_ELFRZE,
              "CRH": T_CRH, "PHOTD": T_PHOTD, "THERM": T_THERM}
    if r2 in simple:
        return simple[r2]
    if r2[:1] == "#":
        return T_GRAIN
    if r2 in ("XRAY", "XRSEC", "XRLYA", "XRPHOT"):
        return T_XR
    if (r1 == "H" and r2 == "H" and r3 in ("", "#") and p[0] == "H2" and p[1] in ("", "#")):
        return T_H2FORM
    if any(x in ("PAH", "PAH0", "PAH+", "PAH-") for x in r.reactants):
        return T_PAH
    return T_THERMAL


def build_ode_terms(net_species: Sequence[Species], reactions: Sequence[Reaction]) -> None:
    """Fill fa/fb/nh/stoich for every reaction (MakeRates conventions of
    odes_*.c; see module docstring).  X-ray reactions get no terms: their
    rate coefficient is identically zero (calc_reac_rates.F90:265-268)."""
    nspec = len(net_species)
    neq = nspec - 1
    index = {s.name: s.index for s in net_species}
    if net_species[-1].name != "e-":
        raise NetworkError("the electron must be the last species (read_species.F90:164)")
    XE, ONE = neq, neq + 1

    def xidx(name):
        return XE if name == "e-" else index[name]

    for r in reactions:
        r.stoich = {}
        if r.rtype == T_XR:
            r.fa, r.fb, r.nh = ONE, ONE, 0
            continue
        r1, r2, r3 = r.reactants
        if r3 == "#":
            # grain-catalysed H + H -> H2 : k * x_H * n_H, both H consumed
            if r1 not in index or r2 not in index:
                raise NetworkError(f"grain reaction {r.number} with unknown reactants")
            r.fa, r.fb, r.nh = xidx(r1), ONE, 1
            react = [r1, r2]
        else:
            react = [x for x in (r1, r2, r3) if x in index]
            if len(react) == 2:
                r.fa, r.fb, r.nh = xidx(react[0]), xidx(react[1]), 1
            elif len(react) == 1:
                r.fa, r.fb, r.nh = xidx(react[0]), ONE, 0
            else:
                raise NetworkError(f"reaction {r.number}: {len(react)} species reactants")
        for x in react:
            if x != "e-":
                r.stoich[index[x]] = r.stoich.get(index[x], 0) - 1
        for x in r.products:
            if x in index and x != "e-":
                r.stoich[index[x]] = r.stoich.get(index[x], 0) + 1
        r.stoich = {k: v for k, v in r.stoich.items() if v != 0}


def nr_indices(reactions: Sequence[Reaction]) -> Dict[str, int]:
    """NRGR/NRH2/NRHD/NRCO/NRCI/NRSI as set by calc_reac_rates.F90 (last
    matching reaction wins; 1-based; 0 = none)."""
    out = {"NRGR": 0, "NRH2": 0, "NRHD": 0, "NRCO": 0, "NRCI": 0, "NRSI": 0}
    for r in reactions:
        n = r.index + 1
        if r.rtype == T_H2FORM:
            out["NRGR"] = n
        elif r.rtype == T_PHOTO_H2 and r.products[0] == "H" and r.products[1] == "H":
            out["NRH2"] = n
        elif r.rtype == T_PHOTO_HD and "H" in r.products and "D" in r.products:
            out["NRHD"] = n
        elif r.rtype == T_PHOTO_CO:
            out["NRCO"] = n
        elif r.rtype == T_PHOTO_CI:
            out["NRCI"] = n
        elif r.rtype == T_PHOTO_SI:
            out["NRSI"] = n
    return out


def xlambda_mode(reactions: Sequence[Reaction]) -> int:
    """Whether XLAMBDA (shield.F90:525-549) interpolates tau(lambda)/tau(V)
    with the cubic spline (1) or, effectively, piecewise linearly (0).

    XLAMBDA computes its spline second derivatives only while the shared
    START flag is set (shield.F90:543); COSHIELD clears START the first time
    it runs (shield.F90:398-401), and COPDRATE calls COSHIELD before SCATTER
    -> XLAMBDA (shield.F90:104).  So if the first reaction whose rate calls
    XLAMBDA is a CO photodissociation, the derivative table stays zero for
    the whole run and the spline degenerates to linear interpolation; if an
    H2/HD photodissociation (H2PDRATE -> SCATTER, shield.F90:51) comes first,
    the table is filled.  All three shipped networks list CO before H2, and
    the amdflang-built reference rate code confirms the linear behaviour."""
    for r in reactions:
        if r.rtype in (T_PHOTO_H2, T_PHOTO_HD):
            return 1
        if r.rtype == T_PHOTO_CO:
            return 0
    return 1


def compile_network(name: str, species_path: str, rates_path: str) -> Network:
    species, labels = read_species(species_path)
    reactions = read_rates(rates_path)
    for r in reactions:
        r.rtype = classify(r)
        r.gflag = 1 if r.reactants[0][:1] == "G" else 0
    build_ode_terms(species, reactions)
    return Network(name, species, reactions, labels, nr_indices(reactions), xlambda_mode(reactions))


# --------------------------------------------------------------------------
# serialisation
# --------------------------------------------------------------------------

def write_tab(net: Network, path: str) -> None:
    """Numeric table read by oracle/pdr_oracle.c (pdr_net_load) and by the
    HIP code generator.  Doubles are written with 17 significant digits so
    the round trip is exact."""
    with open(path, "w") as fh:
        fh.write(f"# pdrchem network table v1: {net.name}\n")
        fh.write(f"N {net.nspec} {net.nreac}\n")
        fh.write(f"F {net.xlambda_spline}\n")
        for s in net.species:
            fh.write(f"S {s.index} {s.charge} {s.abundance:.17g} {s.mass:.17g} {s.name}\n")
        for r in net.reactions:
            st = " ".join(f"{k} {v}" for k, v in sorted(r.stoich.items()))
            fh.write(f"R {r.index} {r.rtype} {r.duplicate} {r.gflag} {r.alpha:.17g} {r.beta:.17g} "
                     f"{r.gamma:.17g} {r.rtmin:.17g} {r.rtmax:.17g} {r.fa} {r.fb} {r.nh} "
                     f"{len(r.stoich)} {st}\n".replace("  \n", " \n"))
        # elimination order of the bordered iteration matrix (pdrchem.sparse)
        from .sparse import analyse
        a = analyse(net)
        fh.write(f"O {a.nb} {a.nhead} " + " ".join(str(v) for v in a.order) + "\n")
        fh.write("E\n")


def to_json(net: Network) -> dict:
    return {
        "name": net.name, "nspec": net.nspec, "nreac": net.nreac, "neq": net.neq,
        "species": [dataclasses.asdict(s) for s in net.species],
        "reactions": [
            {**{k: v for k, v in dataclasses.asdict(r).items() if k != "stoich"},
             "stoich": {str(k): v for k, v in sorted(r.stoich.items())},
             "type_name": TYPE_NAMES[r.rtype]}
            for r in net.reactions
        ],
        "labels": net.labels,
        "nr_indices": net.nr_indices,
        "xlambda_spline": net.xlambda_spline,
    }


def from_json(d: dict) -> Network:
    species = [Species(**s) for s in d["species"]]
    reactions = []
    for rd in d["reactions"]:
        rd = dict(rd)
        rd.pop("type_name", None)
        st = {int(k): v for k, v in rd.pop("stoich").items()}
        reactions.append(Reaction(**rd, stoich=st))
    return Network(d["name"], species, reactions, d["labels"], d["nr_indices"],
                   d.get("xlambda_spline", 1))


NETWORK_FILES = {
    "reduced": ("species_reduced.d", "rates_reduced.d"),
    "full": ("species_full.d", "rates_full.d"),
    "mynetwork": ("species_mynetwork.d", "rates_mynetwork.d"),
}


def generate_all(src_dir: str, out_dir: str) -> List[Network]:
    """Compile every network found in ``src_dir`` into ``out_dir``."""
    os.makedirs(out_dir, exist_ok=True)
    nets = []
    for name, (sp, rt) in NETWORK_FILES.items():
        sp_path, rt_path = os.path.join(src_dir, sp), os.path.join(src_dir, rt)
        if not (os.path.exists(sp_path) and os.path.exists(rt_path)):
            continue
        net = compile_network(name, sp_path, rt_path)
        write_tab(net, os.path.join(out_dir, f"{name}.tab"))
        with open(os.path.join(out_dir, f"{name}.json"), "w") as fh:
            json.dump(to_json(net), fh, indent=0)
        nets.append(net)
    return nets


def load_network(name: str, net_dir: Optional[str] = None) -> Network:
    net_dir = net_dir or os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "networks")
    with open(os.path.join(net_dir, f"{name}.json")) as fh:
        return from_json(json.load(fh))


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("src_dir", help="directory holding species_*.d / rates_*.d")
    ap.add_argument("out_dir")
    a = ap.parse_args()
    for n in generate_all(a.src_dir, a.out_dir):
        print(f"{n.name}: nspec={n.nspec} nreac={n.nreac}")
