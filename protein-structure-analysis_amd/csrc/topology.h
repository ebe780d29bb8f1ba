// topology.h -- covalent topology of the protein_physics drop-in
// (BondTopology, src/physics_engine.cpp:1186-1746 of the reference).
// Host-side integer graph work, built once per structure.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace alma_host {

// BondKind values {BACKBONE_PHI=0, BACKBONE_PSI=1, SIDECHAIN=2, FIXED=3}
// (src/physics_engine.cpp:453-485)
enum class BondKind : uint8_t { BACKBONE_PHI = 0, BACKBONE_PSI = 1, SIDECHAIN = 2, FIXED = 3 };

struct RotBond {
    int i;
    int j;
    BondKind kind;
};

struct DihTerm {
    double V2;
    int n;
    double gamma;
};

struct DihRecord {
    int a, b, c, d;
    std::vector<DihTerm> terms;
};

struct XYZ {
    double x, y, z;
};

class BondTopology {
public:
    int N = 0;
    std::vector<std::vector<int>> adj;
    std::vector<std::pair<int, int>> bonds;
    std::vector<RotBond> rot_bonds;
    std::vector<std::vector<int>> rot_bond_sides;
    std::vector<std::vector<int>> excl;
    std::vector<DihRecord> dihedrals;
    std::vector<double> rot_bond_scale;
    std::vector<std::pair<int, int>> concerted_pairs;
    std::vector<std::pair<int, int>> concerted_sidechain_pairs;
    std::vector<std::pair<int, int>> disulfide_pairs;

    void build(const std::vector<std::string>& resnames, const std::vector<std::string>& atomnames,
               const std::vector<int>& res_idx);
    void add_disulfide(int i, int j);
    void identify_concerted_sidechain_pairs(const std::vector<XYZ>& pos, double cutoff = 6.0);
    bool bonded(int i, int j) const noexcept;
    bool is_excluded(int i, int j) const noexcept;
    std::vector<int> j_side(int bi, int bj) const;
};

}  // namespace alma_host
