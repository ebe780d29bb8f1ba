// topology.cpp -- BondTopology of the protein_physics drop-in.
//
// Follows BondTopology::build and helpers of the reference
// (src/physics_engine.cpp:1303-1746).  The residue bond templates
// (:577-984) and rotatable-bond specifications (:1019-1108) are data the
// drop-in must share bit-for-bit (array order drives adjacency order, hence
// dihedral anchors and j-side DFS order); they are restated here as compact
// "A-B" strings parsed once.
#include "topology.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <sstream>
#include <unordered_map>

namespace alma_host {

namespace {

constexpr double kPi = 3.14159265358979323846;

// ── residue bond templates (src/physics_engine.cpp:577-984) ────────────────
const char* const BB = "N-H N-CA CA-HA CA-C C-O ";
const char* const BB_GLY = "N-H N-CA CA-HA2 CA-HA3 CA-C C-O ";
const char* const BB_PRO = "N-CA N-CD CA-HA CA-C C-O ";

struct Tmpl {
    const char* name;
    const char* bb;
    const char* body;
};

const Tmpl kBondTemplates[] = {
    {"GLY", BB_GLY, ""},
    {"ALA", BB, "CA-CB CB-HB1 CB-HB2 CB-HB3"},
    {"VAL", BB, "CA-CB CB-HB CB-CG1 CG1-HG11 CG1-HG12 CG1-HG13 CB-CG2 CG2-HG21 CG2-HG22 CG2-HG23"},
    {"LEU", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-HG CG-CD1 CD1-HD11 CD1-HD12 CD1-HD13 CG-CD2 CD2-HD21 "
     "CD2-HD22 CD2-HD23"},
    {"ILE", BB,
     "CA-CB CB-HB CB-CG1 CG1-HG12 CG1-HG13 CG1-CD1 CD1-HD11 CD1-HD12 CD1-HD13 CB-CG2 CG2-HG21 "
     "CG2-HG22 CG2-HG23"},
    {"PRO", BB_PRO, "CA-CB CB-HB2 CB-HB3 CB-CG CG-HG2 CG-HG3 CG-CD CD-HD2 CD-HD3"},
    {"PHE", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-CD1 CD1-HD1 CD1-CE1 CE1-HE1 CE1-CZ CZ-HZ CZ-CE2 CE2-HE2 "
     "CE2-CD2 CD2-HD2 CD2-CG"},
    {"TRP", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-CD1 CD1-HD1 CD1-NE1 NE1-HE1 NE1-CE2 CE2-CD2 CD2-CG CE2-CZ2 "
     "CZ2-HZ2 CZ2-CH2 CH2-HH2 CH2-CZ3 CZ3-HZ3 CZ3-CE3 CE3-HE3 CE3-CD2"},
    {"SER", BB, "CA-CB CB-HB2 CB-HB3 CB-OG OG-HG"},
    {"THR", BB, "CA-CB CB-HB CB-OG1 OG1-HG1 CB-CG2 CG2-HG21 CG2-HG22 CG2-HG23"},
    {"CYS", BB, "CA-CB CB-HB2 CB-HB3 CB-SG SG-HG"},
    {"MET", BB, "CA-CB CB-HB2 CB-HB3 CB-CG CG-HG2 CG-HG3 CG-SD SD-CE CE-HE1 CE-HE2 CE-HE3"},
    {"ASP", BB, "CA-CB CB-HB2 CB-HB3 CB-CG CG-OD1 CG-OD2"},
    {"ASN", BB, "CA-CB CB-HB2 CB-HB3 CB-CG CG-OD1 CG-ND2 ND2-HD21 ND2-HD22"},
    {"GLU", BB, "CA-CB CB-HB2 CB-HB3 CB-CG CG-HG2 CG-HG3 CG-CD CD-OE1 CD-OE2"},
    {"GLN", BB, "CA-CB CB-HB2 CB-HB3 CB-CG CG-HG2 CG-HG3 CG-CD CD-OE1 CD-NE2 NE2-HE21 NE2-HE22"},
    {"LYS", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-HG2 CG-HG3 CG-CD CD-HD2 CD-HD3 CD-CE CE-HE2 CE-HE3 CE-NZ "
     "NZ-HZ1 NZ-HZ2 NZ-HZ3"},
    {"ARG", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-HG2 CG-HG3 CG-CD CD-HD2 CD-HD3 CD-NE NE-HE NE-CZ CZ-NH1 "
     "NH1-HH11 NH1-HH12 CZ-NH2 NH2-HH21 NH2-HH22"},
    {"HID", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-ND1 ND1-HD1 ND1-CE1 CE1-HE1 CE1-NE2 NE2-CD2 CD2-HD2 CD2-CG"},
    {"HIE", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-ND1 ND1-CE1 CE1-HE1 CE1-NE2 NE2-HE2 NE2-CD2 CD2-HD2 CD2-CG"},
    {"HIP", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-ND1 ND1-HD1 ND1-CE1 CE1-HE1 CE1-NE2 NE2-HE2 NE2-CD2 CD2-HD2 "
     "CD2-CG"},
    {"TYR", BB,
     "CA-CB CB-HB2 CB-HB3 CB-CG CG-CD1 CD1-HD1 CD1-CE1 CE1-HE1 CE1-CZ CZ-OH OH-HH CZ-CE2 CE2-HE2 "
     "CE2-CD2 CD2-HD2 CD2-CG"},
    {"ATP", "",
     "N1-C2 C2-N3 N3-C4 C4-C5 C5-C6 C6-N1 C4-N9 N9-C8 C8-N7 N7-C5 C6-N6 N9-C1' C1'-C2' C2'-C3' "
     "C3'-C4' C4'-O4' O4'-C1' C2'-O2' C3'-O3' C4'-C5' C5'-O5' O5'-PA PA-O1A PA-O2A PA-O3A O3A-PB "
     "PB-O1B PB-O2B PB-O3B O3B-PG PG-O1G PG-O2G PG-O3G"},
    {"HEM", "",
     "CHA-C1A CHA-C4D CHB-C4A CHB-C1B CHC-C4B CHC-C1C CHD-C4C CHD-C1D C1A-C2A C1A-NA C2A-C3A "
     "C2A-CAA C3A-C4A C3A-CMA C4A-NA CAA-CBA CBA-CGA CGA-O1A CGA-O2A C1B-C2B C1B-NB C2B-C3B "
     "C2B-CMB C3B-C4B C3B-CAB C4B-NB CAB-CBB C1C-C2C C1C-NC C2C-C3C C2C-CMC C3C-C4C C3C-CAC "
     "C4C-NC CAC-CBC C1D-C2D C1D-ND C2D-C3D C2D-CMD C3D-C4D C3D-CAD C4D-ND CAD-CBD CBD-CGD "
     "CGD-O1D CGD-O2D FE-NA FE-NB FE-NC FE-ND"},
    {"NAD", "",
     "PA-O1A PA-O2A PA-O5B PA-O3 O5B-C5B C5B-C4B C4B-O4B C4B-C3B O4B-C1B C3B-O3B C3B-C2B "
     "C2B-O2B C2B-C1B C1B-N9A N9A-C8A N9A-C4A C8A-N7A N7A-C5A C5A-C6A C5A-C4A C6A-N6A C6A-N1A "
     "N1A-C2A C2A-N3A N3A-C4A O3-PN PN-O1N PN-O2N PN-O5D O5D-C5D C5D-C4D C4D-O4D C4D-C3D O4D-C1D "
     "C3D-O3D C3D-C2D C2D-O2D C2D-C1D C1D-N1N N1N-C2N N1N-C6N C2N-C3N C3N-C7N C3N-C4N C7N-O7N "
     "C7N-N7N C4N-C5N C5N-C6N"},
    {"FAD", "",
     "PA-O1A PA-O2A PA-O5B PA-O3P O5B-C5B C5B-C4B C4B-O4B C4B-C3B O4B-C1B C3B-O3B C3B-C2B "
     "C2B-O2B C2B-C1B C1B-N9A N9A-C8A N9A-C4A C8A-N7A N7A-C5A C5A-C6A C5A-C4A C6A-N6A C6A-N1A "
     "N1A-C2A C2A-N3A N3A-C4A N1-C2 N1-C10 C2-O2 C2-N3 N3-C4 C4-O4 C4-C4X C4X-N5 C4X-C10 N5-C5X "
     "C5X-C6 C5X-C9A C6-C7 C7-C7M C7-C8 C8-C8M C8-C9 C9-C9A C9A-N10 N10-C10 N10-C1' C1'-C2' "
     "C2'-O2' C2'-C3' C3'-O3' C3'-C4' C4'-O4' C4'-C5' C5'-O5' O5'-P P-O1P P-O2P P-O3P"},
    {"PLP", "",
     "N1-C2 N1-C6 C2-C2A C2-C3 C3-O3 C3-C4 C4-C4A C4-C5 C4A-O4A C5-C6 C5-C5A C5A-O4P O4P-P "
     "P-O1P P-O2P P-O3P"},
    {"AMP", "",
     "P-O1P P-O2P P-O3P P-O5' O5'-C5' C5'-C4' C4'-O4' C4'-C3' O4'-C1' C3'-O3' C3'-C2' C2'-O2' "
     "C2'-C1' C1'-N9 N9-C8 N9-C4 C8-N7 N7-C5 C5-C6 C5-C4 C6-N6 C6-N1 N1-C2 C2-N3 N3-C4"},
    {"ADP", "",
     "PB-O1B PB-O2B PB-O3B PB-O3A PA-O1A PA-O2A PA-O3A PA-O5' O5'-C5' C5'-C4' C4'-O4' C4'-C3' "
     "O4'-C1' C3'-O3' C3'-C2' C2'-O2' C2'-C1' C1'-N9 N9-C8 N9-C4 C8-N7 N7-C5 C5-C6 C5-C4 C6-N6 "
     "C6-N1 N1-C2 C2-N3 N3-C4"},
    {"GTP", "",
     "PG-O1G PG-O2G PG-O3G PG-O3B O3B-PB PB-O1B PB-O2B PB-O3A O3A-PA PA-O1A PA-O2A PA-O5' "
     "O5'-C5' C5'-C4' C4'-O4' C4'-C3' O4'-C1' C3'-O3' C3'-C2' C2'-O2' C2'-C1' C1'-N9 N9-C8 N9-C4 "
     "C8-N7 N7-C5 C5-C6 C5-C4 C6-O6 C6-N1 N1-C2 C2-N2 C2-N3 N3-C4"},
    {"GDP", "",
     "PB-O1B PB-O2B PB-O3B PB-O3A O3A-PA PA-O1A PA-O2A PA-O5' O5'-C5' C5'-C4' C4'-O4' C4'-C3' "
     "O4'-C1' C3'-O3' C3'-C2' C2'-O2' C2'-C1' C1'-N9 N9-C8 N9-C4 C8-N7 N7-C5 C5-C6 C5-C4 C6-O6 "
     "C6-N1 N1-C2 C2-N2 C2-N3 N3-C4"},
};

// ── rotatable bonds (src/physics_engine.cpp:1019-1108) ─────────────────────
// "A-B:k" with k = p (phi), s (psi), c (side chain)
const char* const kRotSpecs[][2] = {
    {"GLY", "N-CA:p CA-C:s"},
    {"ALA", "N-CA:p CA-C:s"},
    {"VAL", "N-CA:p CA-C:s CA-CB:c"},
    {"LEU", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
    {"ILE", "N-CA:p CA-C:s CA-CB:c CB-CG1:c"},
    {"PRO", "CA-C:s"},
    {"PHE", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
    {"TRP", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
    {"SER", "N-CA:p CA-C:s CA-CB:c CB-OG:c"},
    {"THR", "N-CA:p CA-C:s CA-CB:c"},
    {"CYS", "N-CA:p CA-C:s CA-CB:c CB-SG:c"},
    {"MET", "N-CA:p CA-C:s CA-CB:c CB-CG:c CG-SD:c SD-CE:c"},
    {"ASP", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
    {"ASN", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
    {"GLU", "N-CA:p CA-C:s CA-CB:c CB-CG:c CG-CD:c"},
    {"GLN", "N-CA:p CA-C:s CA-CB:c CB-CG:c CG-CD:c"},
    {"LYS", "N-CA:p CA-C:s CA-CB:c CB-CG:c CG-CD:c CD-CE:c CE-NZ:c"},
    {"ARG", "N-CA:p CA-C:s CA-CB:c CB-CG:c CG-CD:c CD-NE:c"},
    {"HID", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
    {"HIE", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
    {"HIP", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
    {"TYR", "N-CA:p CA-C:s CA-CB:c CB-CG:c"},
};

struct Spec {
    std::string a, b;
    BondKind kind;
};

using PairList = std::vector<std::pair<std::string, std::string>>;

const std::map<std::string, PairList>& bond_table() {
    static const std::map<std::string, PairList> t = [] {
        std::map<std::string, PairList> m;
        for (const auto& tp : kBondTemplates) {
            std::istringstream in(std::string(tp.bb) + tp.body);
            std::string tok;
            PairList v;
            while (in >> tok) {
                auto d = tok.find('-');
                v.emplace_back(tok.substr(0, d), tok.substr(d + 1));
            }
            m[tp.name] = std::move(v);
        }
        return m;
    }();
    return t;
}

const std::map<std::string, std::vector<Spec>>& rot_table() {
    static const std::map<std::string, std::vector<Spec>> t = [] {
        std::map<std::string, std::vector<Spec>> m;
        for (const auto& r : kRotSpecs) {
            std::istringstream in(r[1]);
            std::string tok;
            std::vector<Spec> v;
            while (in >> tok) {
                auto d = tok.find('-');
                auto c = tok.find(':');
                char k = tok[c + 1];
                BondKind kind = k == 'p' ? BondKind::BACKBONE_PHI
                              : k == 's' ? BondKind::BACKBONE_PSI
                                         : BondKind::SIDECHAIN;
                v.push_back({tok.substr(0, d), tok.substr(d + 1, c - d - 1), kind});
            }
            m[r[0]] = std::move(v);
        }
        return m;
    }();
    return t;
}

// get_dih_terms, src/physics_engine.cpp:508-533
std::vector<DihTerm> dih_terms(BondKind kind, const std::string& aname_j) {
    if (kind == BondKind::BACKBONE_PHI) return {{0.40, 3, kPi}, {0.29, 2, kPi}};
    if (kind == BondKind::BACKBONE_PSI) return {{0.40, 3, 0.0}, {0.29, 2, kPi}};
    char e = aname_j.empty() ? 'C' : aname_j[0];
    if (e == 'O') return {{1.18, 2, kPi}, {0.14, 3, 0.0}};
    if (e == 'S') return {{0.60, 3, 0.0}};
    if (e == 'N') return {{0.50, 3, 0.0}, {0.16, 2, 0.0}};
    return {{0.56, 3, 0.0}};
}

}  // namespace

// BondTopology::build, src/physics_engine.cpp:1341-1597
void BondTopology::build(const std::vector<std::string>& resnames,
                         const std::vector<std::string>& atomnames, const std::vector<int>& res_idx) {
    N = (int)resnames.size();
    adj.assign(N, {});
    bonds.clear();
    rot_bonds.clear();
    if (N == 0) return;
    if (atomnames.size() != resnames.size() || res_idx.size() != resnames.size())
        throw std::invalid_argument("resnames, atomnames and res_idx differ in length");

    // (residue index, atom name) -> atom; a later duplicate overwrites
    std::map<std::pair<int, std::string>, int> lookup;
    for (int k = 0; k < N; ++k) lookup[{res_idx[k], atomnames[k]}] = k;

    std::vector<int> residues;   // first-appearance order
    std::unordered_map<int, std::string> res_name;
    for (int k = 0; k < N; ++k) {
        int r = res_idx[k];
        if (res_name.find(r) == res_name.end()) {
            residues.push_back(r);
            std::string rn = resnames[k];
            if (rn == "HIS") rn = "HID";
            res_name[r] = rn;
        }
    }
    auto find = [&](int r, const std::string& a) {
        auto it = lookup.find({r, a});
        return it == lookup.end() ? -1 : it->second;
    };
    auto link = [&](int i, int j) {
        if (i == j) return;
        if (i > j) std::swap(i, j);
        adj[i].push_back(j);
        adj[j].push_back(i);
        bonds.push_back({i, j});
    };
    const auto& bt = bond_table();
    for (int r : residues) {
        auto it = bt.find(res_name[r]);
        if (it == bt.end()) continue;
        for (const auto& pr : it->second) {
            int a = find(r, pr.first), b = find(r, pr.second);
            if (a >= 0 && b >= 0) link(a, b);
        }
    }
    // peptide bonds between consecutive residue numbers (:1434-1441)
    for (size_t k = 0; k + 1 < residues.size(); ++k) {
        int r1 = residues[k], r2 = residues[k + 1];
        if (r2 != r1 + 1) continue;
        int c = find(r1, "C"), n = find(r2, "N");
        if (c >= 0 && n >= 0) link(c, n);
    }
    // terminal / tautomer patches (:1453-1466)
    static const char* const patch[][2] = {{"N", "H1"}, {"N", "H2"}, {"N", "H3"}, {"C", "OXT"},
                                           {"NE2", "HE2"}};
    for (int r : residues)
        for (const auto& pp : patch) {
            int a = find(r, pp[0]), b = find(r, pp[1]);
            if (a >= 0 && b >= 0) link(a, b);
        }
    // rotatable bonds (:1470-1481)
    const auto& rt = rot_table();
    for (int r : residues) {
        auto it = rt.find(res_name[r]);
        if (it == rt.end()) continue;
        for (const auto& sp : it->second) {
            int a = find(r, sp.a), b = find(r, sp.b);
            if (a >= 0 && b >= 0) rot_bonds.push_back({a, b, sp.kind});
        }
    }
    // exclusions: 1-2 and 1-3, sorted unique (:1489-1507)
    excl.assign(N, {});
    for (const auto& b : bonds) excl[b.first].push_back(b.second);
    for (int i = 0; i < N; ++i)
        for (int m : adj[i])
            for (int k : adj[m]) {
                if (k == i) continue;
                excl[std::min(i, k)].push_back(std::max(i, k));
            }
    for (auto& v : excl) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    // one dihedral record per rotatable bond, anchors = first other
    // neighbour in adjacency order (:1540-1553)
    dihedrals.clear();
    for (const auto& rb : rot_bonds) {
        int a = -1, d = -1;
        for (int x : adj[rb.i])
            if (x != rb.j) {
                a = x;
                break;
            }
        for (int x : adj[rb.j])
            if (x != rb.i) {
                d = x;
                break;
            }
        if (a < 0 || d < 0) continue;
        auto terms = dih_terms(rb.kind, atomnames[rb.j]);
        if (!terms.empty()) dihedrals.push_back({a, rb.i, rb.j, d, std::move(terms)});
    }
    rot_bond_sides.resize(rot_bonds.size());
    for (size_t k = 0; k < rot_bonds.size(); ++k) rot_bond_sides[k] = j_side(rot_bonds[k].i, rot_bonds[k].j);
    // lever-arm scale clamp(sqrt(10/|side|), 0.05, 1) (:1564-1570)
    rot_bond_scale.resize(rot_bonds.size());
    for (size_t k = 0; k < rot_bonds.size(); ++k) {
        double ns = std::max(1.0, (double)rot_bond_sides[k].size());
        rot_bond_scale[k] = std::min(1.0, std::max(0.05, std::sqrt(10.0 / ns)));
    }
    // crankshaft (phi, psi) pairs sharing a CA, in the iteration order of a
    // std::unordered_map<int,int> filled in rot-bond order (:1575-1595)
    {
        std::unordered_map<int, int> phi_at, psi_at;
        for (size_t k = 0; k < rot_bonds.size(); ++k) {
            if (rot_bonds[k].kind == BondKind::BACKBONE_PHI)
                phi_at[rot_bonds[k].j] = (int)k;
            else if (rot_bonds[k].kind == BondKind::BACKBONE_PSI)
                psi_at[rot_bonds[k].i] = (int)k;
        }
        for (const auto& kv : phi_at) {
            auto it = psi_at.find(kv.first);
            if (it != psi_at.end()) concerted_pairs.push_back({kv.second, it->second});
        }
    }
}

// add_disulfide, src/physics_engine.cpp:1303-1314
void BondTopology::add_disulfide(int i, int j) {
    if (i < 0 || j < 0 || i >= N || j >= N || i == j) return;
    if (i > j) std::swap(i, j);
    disulfide_pairs.push_back({i, j});
    auto& v = excl[i];
    auto p = std::lower_bound(v.begin(), v.end(), j);
    if (p == v.end() || *p != j) v.insert(p, j);
}

// identify_concerted_sidechain_pairs, src/physics_engine.cpp:1640-1673
void BondTopology::identify_concerted_sidechain_pairs(const std::vector<XYZ>& pos, double cutoff) {
    concerted_sidechain_pairs.clear();
    const size_t kMax = 4000;
    const double c2 = cutoff * cutoff;
    const int nrb = (int)rot_bonds.size();
    auto disjoint = [&](int a, int b) {
        const auto& sa = rot_bond_sides[a];
        const auto& sb = rot_bond_sides[b];
        const auto& small = sa.size() <= sb.size() ? sa : sb;
        const auto& big = sa.size() <= sb.size() ? sb : sa;
        for (int x : small)
            if (std::find(big.begin(), big.end(), x) != big.end()) return false;
        return true;
    };
    for (int a = 0; a < nrb && concerted_sidechain_pairs.size() < kMax; ++a) {
        if (rot_bonds[a].kind != BondKind::SIDECHAIN) continue;
        const XYZ& pa = pos.at(rot_bonds[a].j);
        for (int b = a + 1; b < nrb && concerted_sidechain_pairs.size() < kMax; ++b) {
            if (rot_bonds[b].kind != BondKind::SIDECHAIN) continue;
            const XYZ& pb = pos.at(rot_bonds[b].j);
            double dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
            if (dx * dx + dy * dy + dz * dz > c2) continue;
            if (!disjoint(a, b)) continue;
            concerted_sidechain_pairs.push_back({a, b});
        }
    }
}

bool BondTopology::bonded(int i, int j) const noexcept {
    if (i < 0 || i >= N) return false;
    for (int k : adj[i])
        if (k == j) return true;
    return false;
}

bool BondTopology::is_excluded(int i, int j) const noexcept {
    if (i < 0 || j < 0 || i >= N || j >= N) return false;
    if (i > j) std::swap(i, j);
    return std::binary_search(excl[i].begin(), excl[i].end(), j);
}

// j_side, src/physics_engine.cpp:1720-1737: stack DFS from bj with bi as a
// barrier; visit order is the side order
std::vector<int> BondTopology::j_side(int bi, int bj) const {
    std::vector<int> side;
    if (bi < 0 || bi >= N || bj < 0 || bj >= N) return side;
    std::vector<char> seen(N, 0);
    seen[bi] = 1;
    std::vector<int> stk{bj};
    while (!stk.empty()) {
        int cur = stk.back();
        stk.pop_back();
        if (seen[cur]) continue;
        seen[cur] = 1;
        side.push_back(cur);
        for (int nb : adj[cur]) stk.push_back(nb);
    }
    return side;
}

}  // namespace alma_host
