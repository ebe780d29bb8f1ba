/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE.  Compiles the reference's own
 * src/physics_engine.cpp (included unmodified from where it lies) into a
 * Python module `alma_ref` that exposes, next to the reference bindings,
 * its internal statics (per-term energies, neighbour list, Born radii) and
 * a way to seed its member mt19937.  Built without OpenMP so that
 * generate_ensemble takes its serial branch, which draws from the seeded
 * member generator (src/physics_engine.cpp:2537-2541).
 * Used only to write golden vectors into tests/golden.
 */
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#define private public
#include REF_SOURCE
#undef private

namespace {

using Parts = std::vector<Particle>;

std::vector<double> ref_terms(const Parts& p, const BondTopology* topo) {
    // DH, GB, LJ, HC, SASA, DIH, SS, TOTAL using the reference's own statics
    NeighborList nl;
    nl.build(p);
    auto a = PhysicsEngine::born_radii(p, nl);
    std::vector<double> t(8, 0.0);
    t[4] = PhysicsEngine::sasa_nonpolar(p, nl);
    if (topo && !topo->dihedrals.empty()) t[5] = PhysicsEngine::dihedral_e(p, topo->dihedrals);
    if (topo && !topo->disulfide_pairs.empty()) t[6] = PhysicsEngine::ss_e(p, topo->disulfide_pairs);
    for (size_t i = 0; i < p.size(); ++i)
        for (size_t j : nl.nb[i]) {
            if (topo && topo->is_excluded((int)i, (int)j)) continue;
            if (PhysicsEngine::d2(p[i], p[j]) > PAIR_CUT2) continue;
            const Particle &pi = p[i], &pj = p[j];
            double ai = a[i], aj = a[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            double r2 = dx * dx + dy * dy + dz * dz, r = std::sqrt(r2), sig = pi.radius + pj.radius;
            double qp = pi.charge * pj.charge;
            t[0] += (COULOMB * qp) / (EPS_WATER * r) * std::exp(-KAPPA * r);
            double fgb = std::sqrt(r2 + ai * aj * std::exp(-r2 / (4.0 * ai * aj)));
            t[1] += GB_COEF * qp / fgb;
            double hc = 0.0;
            double e = PhysicsEngine::pair_e_diag(pi, pj, ai, aj, hc);
            t[3] += hc;
            t[2] += e - hc - (COULOMB * qp) / (EPS_WATER * r) * std::exp(-KAPPA * r) - GB_COEF * qp / fgb;
        }
    t[7] = t[0] + t[1] + t[2] + t[3] + t[4] + t[5] + t[6];
    return t;
}

}  // namespace

PYBIND11_MODULE(alma_ref, m) {
    pybind11_init_protein_physics(m);   // every reference binding
    m.def("terms", [](const Parts& p, pybind11::object topo) {
        const BondTopology* t = topo.is_none() ? nullptr : topo.cast<BondTopology*>();
        return ref_terms(p, t);
    }, pybind11::arg("particles"), pybind11::arg("topology") = pybind11::none());
    m.def("nlist", [](const Parts& p) {
        NeighborList nl;
        nl.build(p);
        std::vector<int> off(1, 0), js;
        for (auto& v : nl.nb) {
            for (size_t j : v) js.push_back((int)j);
            off.push_back((int)js.size());
        }
        return std::make_pair(off, js);
    });
    m.def("born", [](const Parts& p) {
        NeighborList nl;
        nl.build(p);
        return PhysicsEngine::born_radii(p, nl);
    });
    m.def("seed", [](PhysicsEngine& e, unsigned s) { e.gen.seed(s); });
    m.def("mt_words", [](unsigned s, int n) {
        std::mt19937 g(s);
        std::vector<unsigned> v(n);
        for (auto& x : v) x = (unsigned)g();
        return v;
    });
    m.def("uniform_int", [](unsigned s, int hi, int n) {
        std::mt19937 g(s);
        std::uniform_int_distribution<int> d(0, hi);
        std::vector<int> v(n);
        for (auto& x : v) x = d(g);
        return v;
    });
    m.def("uniform_real", [](unsigned s, double a, double b, int n) {
        std::mt19937 g(s);
        std::vector<double> v(n);
        for (auto& x : v) x = std::uniform_real_distribution<double>(a, b)(g);
        return v;
    });
    m.def("dihedral_angle", [](const Parts& p, int a, int b, int c, int d) {
        return PhysicsEngine::dihedral_angle(p, a, b, c, d);
    });
}
