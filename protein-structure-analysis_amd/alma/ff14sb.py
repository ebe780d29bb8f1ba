"""ff14SB per-atom parameters for the 20 standard amino acids (+ HID/HIE/HIP).

Values are the ones ALMA assigns through ``amber_params.get_atom_params``:
  * R* / epsilon by atom type      -- python/amber_params.py:57-98
  * atom type by (residue, atom)   -- python/amber_params.py:110-325
  * RESP partial charges           -- python/amber_params.py:348-595
  * element fallback for names with no type (terminal H1/H2/H3)
                                   -- python/amber_params.py:971-1001

The layout here is this package's own: one record per residue listing atoms
in *storage order*.  The storage order is chosen so that every rotatable bond
of ``BondTopology.build`` has a j-side that is one contiguous index run
(N, H, CA, HA, side chain depth-first with each atom's hydrogens right after
it and the non-rotating branch first, then C, O).  The force-field numbers are
identical to the reference tables; only their arrangement differs.
"""
from __future__ import annotations

# atom type -> (R*, epsilon)            python/amber_params.py:57-98
VDW = {
    "C": (1.9080, 0.0860), "CA": (1.9080, 0.0860), "CB": (1.9080, 0.0860),
    "CC": (1.9080, 0.0860), "CN": (1.9080, 0.0860), "CR": (1.9080, 0.0860),
    "CT": (1.9080, 0.1094), "CX": (1.9080, 0.1094), "C8": (1.9080, 0.1094),
    "2C": (1.9080, 0.1094), "3C": (1.9080, 0.1094), "C*": (1.9080, 0.0860),
    "CV": (1.9080, 0.0860), "CW": (1.9080, 0.0860),
    "N": (1.8240, 0.1700), "N2": (1.8240, 0.1700), "N3": (1.8750, 0.1700),
    "NA": (1.8240, 0.1700), "NB": (1.8240, 0.1700), "NC": (1.8240, 0.1700),
    "O": (1.6612, 0.2100), "O2": (1.6612, 0.2100), "OH": (1.7210, 0.2104),
    "S": (2.0000, 0.2500), "SH": (2.0000, 0.2500),
    "H": (0.6000, 0.0157), "H1": (1.3870, 0.0157), "H4": (1.4090, 0.0150),
    "H5": (1.3590, 0.0150), "HA": (1.4590, 0.0150), "HC": (1.4870, 0.0157),
    "HO": (0.3000, 0.0000), "HP": (1.1000, 0.0157), "HS": (0.6000, 0.0157),
}
VDW_FALLBACK = (1.9000, 0.1000)
# element fallback for atoms without a type          python/amber_params.py:981-991
ELEMENT_FALLBACK = {"C": (1.9080, 0.1094), "N": (1.8240, 0.1700),
                    "O": (1.6612, 0.2100), "S": (2.0000, 0.2500),
                    "H": (0.6000, 0.0157)}

# Backbone charges (N, H, C, O) default            python/amber_params.py:350-356
_BBQ = (-0.4157, 0.2719, 0.5973, -0.5679)
_BBQ_CHARGED_NEG = (-0.5163, 0.2936, 0.5366, -0.5816)     # ASP, GLU
_BBQ_CHARGED_POS = (-0.3479, 0.2747, 0.7341, -0.5894)     # LYS, ARG


def _res(bbq, ca, ha, side, ha_names=("HA",), ha_type="H1", n_has_h=True):
    """Assemble one residue record: list of (atom, type, charge) in storage order."""
    n, h, c, o = bbq
    atoms = [("N", "N", n)]
    if n_has_h:
        atoms.append(("H", "H", h))
    atoms.append(("CA", "CX", ca[0] if isinstance(ca, tuple) else ca))
    for nm, q in zip(ha_names, ha if isinstance(ha, tuple) else (ha,)):
        atoms.append((nm, ha_type, q))
    atoms.extend(side)
    atoms.append(("C", "C", c))
    atoms.append(("O", "O", o))
    return atoms


# Side chains in storage order: (atom, ff type, charge).  Charges from
# python/amber_params.py:363-595, types from :124-325.
RESIDUES: dict[str, list[tuple[str, str, float]]] = {
    "GLY": _res(_BBQ, -0.0252, (0.0698, 0.0698), [], ha_names=("HA2", "HA3")),
    "ALA": _res(_BBQ, 0.0337, 0.0823, [
        ("CB", "3C", -0.1825), ("HB1", "HC", 0.0603), ("HB2", "HC", 0.0603),
        ("HB3", "HC", 0.0603)]),
    "VAL": _res(_BBQ, -0.0875, 0.0969, [
        ("CB", "C8", 0.2985), ("HB", "HC", -0.0297),
        ("CG1", "3C", -0.3192), ("HG11", "HC", 0.0791), ("HG12", "HC", 0.0791),
        ("HG13", "HC", 0.0791),
        ("CG2", "3C", -0.3192), ("HG21", "HC", 0.0791), ("HG22", "HC", 0.0791),
        ("HG23", "HC", 0.0791)]),
    "LEU": _res(_BBQ, -0.0518, 0.0922, [
        ("CB", "2C", -0.1102), ("HB2", "HC", 0.0457), ("HB3", "HC", 0.0457),
        ("CG", "CT", 0.3531), ("HG", "HC", -0.0361),
        ("CD1", "3C", -0.4121), ("HD11", "HC", 0.1000), ("HD12", "HC", 0.1000),
        ("HD13", "HC", 0.1000),
        ("CD2", "3C", -0.4121), ("HD21", "HC", 0.1000), ("HD22", "HC", 0.1000),
        ("HD23", "HC", 0.1000)]),
    "ILE": _res(_BBQ, -0.0597, 0.0869, [
        ("CB", "C8", 0.1303), ("HB", "HC", 0.0187),
        ("CG2", "3C", -0.3204), ("HG21", "HC", 0.0882), ("HG22", "HC", 0.0882),
        ("HG23", "HC", 0.0882),
        ("CG1", "2C", -0.0430), ("HG12", "HC", 0.0236), ("HG13", "HC", 0.0236),
        ("CD1", "3C", -0.0660), ("HD11", "HC", 0.0186), ("HD12", "HC", 0.0186),
        ("HD13", "HC", 0.0186)]),
    "PRO": _res((-0.2548, 0.0, 0.5896, -0.5748), -0.0266, 0.0641, [
        ("CB", "CT", -0.0070), ("HB2", "HC", 0.0253), ("HB3", "HC", 0.0253),
        ("CG", "CT", 0.0189), ("HG2", "HC", 0.0213), ("HG3", "HC", 0.0213),
        ("CD", "CT", 0.0192), ("HD2", "HC", 0.0391), ("HD3", "HC", 0.0391)],
        n_has_h=False),
    "PHE": _res(_BBQ, -0.0024, 0.0978, [
        ("CB", "2C", -0.0343), ("HB2", "HC", 0.0295), ("HB3", "HC", 0.0295),
        ("CG", "CA", 0.0118),
        ("CD1", "CA", -0.1256), ("HD1", "HA", 0.1330),
        ("CE1", "CA", -0.1704), ("HE1", "HA", 0.1430),
        ("CZ", "CA", -0.1072), ("HZ", "HA", 0.1297),
        ("CE2", "CA", -0.1704), ("HE2", "HA", 0.1430),
        ("CD2", "CA", -0.1256), ("HD2", "HA", 0.1330)]),
    "TRP": _res(_BBQ, -0.0275, 0.1123, [
        ("CB", "2C", -0.0050), ("HB2", "HC", 0.0339), ("HB3", "HC", 0.0339),
        ("CG", "C*", -0.1415),
        ("CD1", "CW", -0.1638), ("HD1", "H4", 0.2062),
        ("NE1", "NA", -0.3418), ("HE1", "H", 0.3412),
        ("CE2", "CN", 0.0390),
        ("CZ2", "CW", -0.1928), ("HZ2", "HA", 0.1999),
        ("CH2", "CA", -0.1134), ("HH2", "HA", 0.1417),
        ("CZ3", "CA", -0.2387), ("HZ3", "HA", 0.1936),
        ("CE3", "CA", -0.2601), ("HE3", "HA", 0.1730),
        ("CD2", "CB", 0.1243)]),
    "SER": _res(_BBQ, -0.0249, 0.0843, [
        ("CB", "2C", 0.2117), ("HB2", "H1", 0.0352), ("HB3", "H1", 0.0352),
        ("OG", "OH", -0.6546), ("HG", "HO", 0.4275)]),
    "THR": _res(_BBQ, -0.0389, 0.1007, [
        ("CB", "C8", 0.3654), ("HB", "H1", 0.0043),
        ("OG1", "OH", -0.6761), ("HG1", "HO", 0.4102),
        ("CG2", "3C", -0.2438), ("HG21", "HC", 0.0642), ("HG22", "HC", 0.0642),
        ("HG23", "HC", 0.0642)]),
    "CYS": _res(_BBQ, 0.0213, 0.1124, [
        ("CB", "2C", -0.1231), ("HB2", "H1", 0.1112), ("HB3", "H1", 0.1112),
        ("SG", "SH", -0.3119), ("HG", "HS", 0.1933)]),
    "MET": _res(_BBQ, -0.0237, 0.0880, [
        ("CB", "2C", 0.0342), ("HB2", "HC", 0.0241), ("HB3", "HC", 0.0241),
        ("CG", "2C", 0.0018), ("HG2", "HC", 0.0440), ("HG3", "HC", 0.0440),
        ("SD", "S", -0.2737),
        ("CE", "3C", -0.0536), ("HE1", "HC", 0.0684), ("HE2", "HC", 0.0684),
        ("HE3", "HC", 0.0684)]),
    "ASP": _res(_BBQ_CHARGED_NEG, 0.0381, 0.0880, [
        ("CB", "2C", -0.0303), ("HB2", "HC", -0.0122), ("HB3", "HC", -0.0122),
        ("CG", "C", 0.7994), ("OD1", "O2", -0.8014), ("OD2", "O2", -0.8014)]),
    "ASN": _res(_BBQ, 0.0143, 0.1048, [
        ("CB", "2C", -0.2041), ("HB2", "HC", 0.0797), ("HB3", "HC", 0.0797),
        ("CG", "C", 0.7130), ("OD1", "O", -0.5931),
        ("ND2", "N2", -0.9191), ("HD21", "H", 0.4196), ("HD22", "H", 0.4196)]),
    "GLU": _res(_BBQ_CHARGED_NEG, 0.0397, 0.1105, [
        ("CB", "2C", 0.0560), ("HB2", "HC", -0.0173), ("HB3", "HC", -0.0173),
        ("CG", "2C", 0.0136), ("HG2", "HC", -0.0425), ("HG3", "HC", -0.0425),
        ("CD", "C", 0.8054), ("OE1", "O2", -0.8188), ("OE2", "O2", -0.8188)]),
    "GLN": _res(_BBQ, -0.0031, 0.1048, [
        ("CB", "2C", -0.0036), ("HB2", "HC", 0.0171), ("HB3", "HC", 0.0171),
        ("CG", "2C", -0.0645), ("HG2", "HC", 0.0352), ("HG3", "HC", 0.0352),
        ("CD", "C", 0.6951), ("OE1", "O", -0.6086),
        ("NE2", "N2", -0.9407), ("HE21", "H", 0.4152), ("HE22", "H", 0.4152)]),
    "LYS": _res(_BBQ_CHARGED_POS, -0.2400, 0.1426, [
        ("CB", "2C", -0.0094), ("HB2", "HC", 0.0362), ("HB3", "HC", 0.0362),
        ("CG", "2C", 0.0187), ("HG2", "HC", 0.0103), ("HG3", "HC", 0.0103),
        ("CD", "2C", -0.0479), ("HD2", "HC", 0.0621), ("HD3", "HC", 0.0621),
        ("CE", "2C", -0.0143), ("HE2", "HP", 0.1135), ("HE3", "HP", 0.1135),
        ("NZ", "N3", -0.3854), ("HZ1", "H", 0.3400), ("HZ2", "H", 0.3400),
        ("HZ3", "H", 0.3400)]),
    "ARG": _res(_BBQ_CHARGED_POS, -0.2637, 0.1560, [
        ("CB", "2C", -0.0007), ("HB2", "HC", 0.0327), ("HB3", "HC", 0.0327),
        ("CG", "2C", 0.0390), ("HG2", "HC", 0.0285), ("HG3", "HC", 0.0285),
        ("CD", "2C", 0.0486), ("HD2", "HC", 0.0687), ("HD3", "HC", 0.0687),
        ("NE", "N2", -0.5295), ("HE", "H", 0.3456),
        ("CZ", "CA", 0.8076),
        ("NH1", "N2", -0.8627), ("HH11", "H", 0.4478), ("HH12", "H", 0.4478),
        ("NH2", "N2", -0.8627), ("HH21", "H", 0.4478), ("HH22", "H", 0.4478)]),
    # HIS is stored as the HID tautomer (python/amber_params.py:299-302,
    # src/physics_engine.cpp:1381).
    "HIS": _res(_BBQ, -0.0581, 0.1360, [
        ("CB", "2C", -0.0074), ("HB2", "HC", 0.0367), ("HB3", "HC", 0.0367),
        ("CG", "CC", 0.0543),
        ("ND1", "NA", -0.3811), ("HD1", "H", 0.3649),
        ("CE1", "CR", 0.2057), ("HE1", "H5", 0.1635),
        ("NE2", "NB", -0.4874),
        ("CD2", "CW", -0.1452), ("HD2", "H4", 0.1958)]),
    # HIE / HIP tautomers (python/amber_params.py:302-322, charges :553-576)
    "HIE": _res(_BBQ, -0.0581, 0.1360, [
        ("CB", "2C", -0.0074), ("HB2", "HC", 0.0367), ("HB3", "HC", 0.0367),
        ("CG", "CC", 0.1421),
        ("ND1", "NB", -0.5432),
        ("CE1", "CR", 0.1635), ("HE1", "H5", 0.1435),
        ("NE2", "NA", -0.2779), ("HE2", "H", 0.3339),
        ("CD2", "CC", -0.2207), ("HD2", "H4", 0.2293)]),
    "HIP": _res((-0.3479, 0.2747, 0.5185, -0.5894), -0.1354, 0.1212, [
        ("CB", "2C", -0.0414), ("HB2", "HC", 0.0810), ("HB3", "HC", 0.0810),
        ("CG", "CC", 0.1880),
        ("ND1", "NA", -0.1513), ("HD1", "H", 0.3866),
        ("CE1", "CR", -0.0012), ("HE1", "H5", 0.2681),
        ("NE2", "NA", -0.1707), ("HE2", "H", 0.3911),
        ("CD2", "CW", -0.1046), ("HD2", "H4", 0.2317)]),
    "TYR": _res(_BBQ, -0.0014, 0.0876, [
        ("CB", "2C", -0.0152), ("HB2", "HC", 0.0295), ("HB3", "HC", 0.0295),
        ("CG", "CA", -0.0011),
        ("CD1", "CA", -0.1906), ("HD1", "HA", 0.1699),
        ("CE1", "CA", -0.2341), ("HE1", "HA", 0.1680),
        ("CZ", "C", 0.3226), ("OH", "OH", -0.5579), ("HH", "HO", 0.3992),
        ("CE2", "CA", -0.2341), ("HE2", "HA", 0.1680),
        ("CD2", "CA", -0.1906), ("HD2", "HA", 0.1699)]),
}

RESIDUES["HID"] = RESIDUES["HIS"]

STANDARD_20 = ("ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS",
               "ILE", "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP",
               "TYR", "VAL")

_TABLE = {(res, a): (t, q) for res, rec in RESIDUES.items() for a, t, q in rec}


def atom_params(resname: str, atomname: str) -> tuple[float, float, float]:
    """(charge, R*, epsilon) for one atom of the tabulated residues (the
    synthetic-structure path): get_atom_params below, which it equals on
    these residues."""
    return get_atom_params(resname, atomname)


# ── PDB input: the rest of get_atom_params (python/amber_params.py:928-1001) ──
# Backbone types every residue registered with the reference's _add() carries
# (python/amber_params.py:110-122): the table entries above, plus OXT (O2) and
# an HA on glycine (H1, no charge).  Proline is not registered that way
# (:163-170): its N-terminal H and OXT take the element fallback.
_BB_TYPES = {"N": "N", "H": "H", "CA": "CX", "HA": "H1", "C": "C", "O": "O", "OXT": "O2"}

# Ions by residue name: (charge e, R* A, epsilon), python/amber_params.py:839-868
ION_PARAMS = {
    "MG": (2.0, 1.185, 0.8947), "CA": (2.0, 1.713, 0.4598), "ZN": (2.0, 1.213, 0.0125),
    "FE": (2.0, 1.275, 0.0130), "FE2": (2.0, 1.275, 0.0130), "FE3": (3.0, 1.206, 0.0130),
    "MN": (2.0, 1.258, 0.0150), "NI": (2.0, 1.163, 0.0150), "CU": (2.0, 1.178, 0.0150),
    "CO": (2.0, 1.227, 0.0150), "NA": (1.0, 1.364, 0.0874), "CL": (-1.0, 2.513, 0.0356),
    "K": (1.0, 1.764, 0.0004), "MO": (0.0, 1.500, 0.0200),
}

# Element fallback for atoms without a type, keyed by the element read off
# the atom name (python/amber_params.py:981-1001; its table lists CL twice,
# the later ion-like entry is the one in effect)
ELEMENT_FALLBACK_ALL = {
    "C": (1.9080, 0.1094), "N": (1.8240, 0.1700), "O": (1.6612, 0.2100), "S": (2.0000, 0.2500),
    "H": (0.6000, 0.0157), "P": (2.1000, 0.2000), "F": (1.7500, 0.0610), "BR": (2.2200, 0.3200),
    "I": (2.3500, 0.4000), "MG": (1.1850, 0.8947), "CA": (1.7131, 0.4598), "ZN": (1.2126, 0.0125),
    "FE": (1.2750, 0.0130), "MN": (1.2580, 0.0150), "NA": (1.3638, 0.0874), "K": (1.7638, 0.0004),
    "CL": (2.5130, 0.0356),
}

# Residue names dropped by the PDB reader (python/amber_params.py:887, :925)
WATER_RESNAMES = frozenset({"HOH", "WAT", "H2O", "TIP", "TIP3", "SOL"})
NUCLEOTIDE_RESNAMES = frozenset({"DA", "DC", "DG", "DT", "DI", "A", "C", "G", "U", "I"})

# United-atom RESP charges of HETATM cofactors (hydrogens folded onto their
# heavy atoms), python/amber_params.py:596-838: "atom charge" pairs.  They
# carry no atom types, so their radii come from the element fallback.
_COFACTOR_TEXT = {
    "ATP": (
        "O1G -0.9526 PG 1.2650 O2G -0.9526 O3G -0.9526 O3B -0.5322 PB 1.3852 O1B -0.8894 O2B "
        "-0.8894 O3A -0.5689 PA 1.2532 O1A -0.8799 O2A -0.8799 O5' -0.5987 C5' 0.1916 C4' "
        "0.2239 O4' -0.3548 C1' 0.2401 N9 -0.0251 C8 0.3559 N7 -0.6073 C5 0.0515 C6 0.7009 N6 "
        "-0.0789 N1 -0.7615 C2 0.6348 N3 -0.6997 C4 0.3053 C3' 0.2637 O3' -0.2165 C2' 0.1642 "
        "O2' -0.1953"),
    "ADP": (
        "O1B -0.9552 PB 1.3672 O2B -0.9552 O3B -0.9552 O3A -0.6346 PA 1.4929 O1A -0.9474 O2A "
        "-0.9474 O5' -0.6579 C5' 0.1916 C4' 0.2239 O4' -0.3548 C1' 0.2401 N9 -0.0251 C8 "
        "0.3559 N7 -0.6073 C5 0.0515 C6 0.7009 N6 -0.0789 N1 -0.7615 C2 0.6348 N3 -0.6997 C4 "
        "0.3053 C3' 0.2637 O3' -0.2165 C2' 0.1642 O2' -0.1953"),
    "GTP": (
        "O1G -0.9526 PG 1.2650 O2G -0.9526 O3G -0.9526 O3B -0.5322 PB 1.3852 O1B -0.8894 O2B "
        "-0.8894 O3A -0.5689 PA 1.2532 O1A -0.8799 O2A -0.8799 O5' -0.5987 C5' 0.1916 C4' "
        "0.2239 O4' -0.3548 C1' 0.2197 N9 0.0492 C8 0.3014 N7 -0.5709 C5 0.1744 C6 0.4770 O6 "
        "-0.5597 N1 -0.1363 C2 0.7657 N2 -0.0944 N3 -0.6323 C4 0.1222 C3' 0.2637 O3' -0.2165 "
        "C2' 0.1642 O2' -0.1953"),
    "GDP": (
        "O1B -0.9552 PB 1.3672 O2B -0.9552 O3B -0.9552 O3A -0.6346 PA 1.4929 O1A -0.9474 O2A "
        "-0.9474 O5' -0.6579 C5' 0.1916 C4' 0.2239 O4' -0.3548 C1' 0.2197 N9 0.0492 C8 0.3014 "
        "N7 -0.5709 C5 0.1744 C6 0.4770 O6 -0.5597 N1 -0.1363 C2 0.7657 N2 -0.0944 N3 -0.6323 "
        "C4 0.1222 C3' 0.2637 C2' 0.1642 O2' -0.1953 O3' -0.2165"),
    "HEM": (
        "FE 0.2500 NA -0.1800 C1A 0.0300 C2A -0.0200 CAA 0.0400 CBA -0.1000 CGA 0.3000 O1A "
        "-0.5000 O2A -0.5000 C3A 0.0200 CMA -0.0400 C4A 0.0200 CHB 0.0400 C1B 0.0300 NB "
        "-0.1800 C2B 0.0200 CMB -0.0400 C3B -0.0500 CAB 0.0200 CBB -0.1000 C4B 0.0200 CHC "
        "0.0400 C1C 0.0300 NC -0.1800 C2C 0.0200 CMC -0.0400 C3C -0.0500 CAC 0.0300 CBC "
        "-0.1000 C4C 0.0200 CHD 0.0400 C1D 0.0300 ND -0.1800 C2D 0.0200 CMD -0.0400 C3D "
        "-0.0200 C4D 0.0200 CHA 0.0400 CAD 0.0400 CBD -0.1000 CGD 0.3000 O1D -0.5000 O2D "
        "-0.5000"),
    "NAD": (
        "PA 1.3729 O1A -0.8558 O2A -0.8558 O5B -0.5661 O3 -0.5760 C5B 0.1916 C4B 0.2239 O4B "
        "-0.3548 C3B 0.2637 O3B -0.2165 C2B 0.1642 O2B -0.1953 C1B 0.2401 N9A -0.0251 C8A "
        "0.3559 N7A -0.6073 C5A 0.0515 C6A 0.7009 N6A -0.0789 N1A -0.7615 C2A 0.6348 N3A "
        "-0.6997 C4A 0.3053 PN 1.3729 O1N -0.8558 O2N -0.8558 O5D -0.5661 C5D 0.1916 C4D "
        "0.2239 O4D -0.3548 C3D 0.2637 O3D -0.2165 C2D 0.1642 O2D -0.1953 C1D 0.3995 N1N "
        "0.0831 C2N 0.1914 C6N 0.2613 C3N -0.2345 C7N 0.8217 O7N -0.5428 N7N -0.1066 C4N "
        "0.2968 C5N -0.0539"),
}
_COFACTOR = {}
for _r, _txt in _COFACTOR_TEXT.items():
    _w = _txt.split()
    for _k in range(0, len(_w), 2):
        _COFACTOR[(_r, _w[_k])] = float(_w[_k + 1])


def _atom_type(resname: str, atomname: str):
    hit = _TABLE.get((resname, atomname))
    if hit is not None:
        return hit[0]
    if resname in RESIDUES and resname != "PRO":
        return _BB_TYPES.get(atomname)
    return None


def _charge(resname: str, atomname: str) -> float:
    hit = _TABLE.get((resname, atomname))
    if hit is not None:
        return hit[1]
    return _COFACTOR.get((resname, atomname), 0.0)


def element_of(atomname: str) -> str:
    """The element the fallback reads off an atom name: leading digits
    dropped, first two characters, upper case, trailing digits dropped."""
    return atomname.lstrip("0123456789")[:2].upper().rstrip("0123456789")


def get_atom_params(resname: str, atomname: str) -> tuple[float, float, float]:
    """(charge e, R* A, epsilon kcal/mol) for one PDB atom, in the lookup
    order of python/amber_params.py:928-1001: ion residues first (so a Ca2+
    "CA" is not read as an alpha carbon), then the (residue, atom) charge
    (0 when absent) and type; atoms without a type take the element fallback
    (a two-letter element, else its first letter, else the generic default)."""
    ion = ION_PARAMS.get(resname)
    if ion is not None:
        return ion
    q = _charge(resname, atomname)
    t = _atom_type(resname, atomname)
    if t is not None:
        r, e = VDW.get(t, VDW_FALLBACK)
    else:
        el = element_of(atomname)
        r, e = ELEMENT_FALLBACK_ALL.get(el, ELEMENT_FALLBACK_ALL.get(el[:1], VDW_FALLBACK))
    return q, r, e


# Hydrogens of each heavy atom, in the residue records' order (each atom's
# hydrogens follow it), for the united-atom charge fold of
# python/amber_params.py:1004-1071.
HEAVY_CHILDREN: dict[tuple[str, str], tuple[str, ...]] = {}
for _r, _rec in RESIDUES.items():
    _heavy = None
    _kids: dict[str, list[str]] = {}
    for _a, _t, _q in _rec:
        if _a.startswith("H"):
            if _heavy is not None:
                _kids.setdefault(_heavy, []).append(_a)
        else:
            _heavy = _a
    for _h, _v in _kids.items():
        HEAVY_CHILDREN[(_r, _h)] = tuple(_v)


def missing_hydrogen_charge(resname: str, heavy_atomname: str, present) -> float:
    """Charge of the hydrogens of (resname, heavy atom) that are absent from
    `present` (the atom names of that residue instance), summed in order;
    0 when all are present or the atom has none."""
    total = 0.0
    for h in HEAVY_CHILDREN.get((resname, heavy_atomname), ()):
        if h not in present:
            total += _charge(resname, h)
    return total
