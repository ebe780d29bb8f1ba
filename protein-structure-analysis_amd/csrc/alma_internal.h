// alma_internal.h -- shared declarations of the gfx950 kernels and the
// context that owns their device buffers.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/alma.h"

namespace alma {

// Physics constants of src/physics_engine.cpp:181-196 and :1959-1961.
constexpr double kCoulomb = 332.0636;
constexpr double kEpsWater = 78.5;
constexpr double kEpsProt = 1.0;
constexpr double kKappa = 0.1257;
constexpr double kGammaSA = 0.00542;
constexpr double kBetaSA = 0.92;
constexpr double kProbe = 1.4;
constexpr double kNlRcut2 = 196.0;        // (12 + 2)^2
constexpr double kPairCut2 = 144.0;       // 12^2
constexpr double kHalfSkin2 = 1.0;        // (2/2)^2
constexpr double kHardScale = 1.0e4;
constexpr double kHardCap = 5.0e3;
constexpr double kHardFrac = 0.6;
constexpr double kGbCoef = -0.5 * (1.0 / kEpsProt - 1.0 / kEpsWater) * kCoulomb;
constexpr double kKss = 600.0;
constexpr double kR0ss = 2.044;
constexpr uint32_t kExclBit = 0x80000000u;   // neighbour entry flag: 1-2/1-3/SS excluded
constexpr uint32_t kIdxMask = 0x7fffffffu;

// Cell-grid geometry of one neighbour-list build (FP64, reference arithmetic).
struct Grid {
    double lo[3];
    double cs;
    int nc[3];
    int ncell;
};

// Device-side topology (flat CSR), owned by the context.
struct DevTopo {
    int n_rot = 0, n_dih = 0, n_ss = 0, n_cp = 0, n_csp = 0;
    int* excl_off = nullptr;
    int* excl_idx = nullptr;
    int* rb_i = nullptr;
    int* rb_j = nullptr;
    int* rb_kind = nullptr;
    double* rb_scale = nullptr;
    int* side_off = nullptr;
    int* side_idx = nullptr;
    int2* side_run = nullptr;      // j-side as sorted index runs [lo, hi]
    int* side_run_off = nullptr;
    int* dih_a = nullptr;
    int* dih_b = nullptr;
    int* dih_c = nullptr;
    int* dih_d = nullptr;
    int* dih_toff = nullptr;
    double* dih_v2 = nullptr;
    int* dih_n = nullptr;
    double* dih_gamma = nullptr;
    int* ss_i = nullptr;
    int* ss_j = nullptr;
    int* cp_phi = nullptr;
    int* cp_psi = nullptr;
    int* csp_a = nullptr;
    int* csp_b = nullptr;
};

// Per-system device buffers.
struct DevSys {
    int n = 0;
    int cap = 0;                 // neighbour-list capacity per atom
    double* pos = nullptr;       // 4n: x, y, z, 0
    float4* prm = nullptr;       // q, R*, eps, a (Born radius)
    double* rstar = nullptr;     // FP64 R*
    double* born_sum = nullptr;  // full HCT sum per atom
    double* half_sum = nullptr;  // half-list HCT sum per atom
    double* a = nullptr;         // FP64 Born radius
    int* cell_of = nullptr;
    int* cell_cnt = nullptr;
    int* cell_start = nullptr;
    int* cell_atoms = nullptr;
    int* cell_tmp = nullptr;
    int* nl_cnt = nullptr;
    uint32_t* nl = nullptr;      // n * cap entries: j | excl bit, cell-scan order
    double* part = nullptr;      // per-atom partial sums: DH, GB, LJ, HC, SASA
    double* rec = nullptr;       // per-record partial sums: dihedral, SS
    double* terms = nullptr;     // ALMA_NTERMS
    int* flags = nullptr;        // [0] overflow, [1] max count
    Grid* grid = nullptr;
    int max_cells = 0;
    DevTopo topo;
    bool has_topo = false;
};

// ---- kernels / launchers (kernels.hip) ----
hipError_t launch_build_nl(DevSys& s, hipStream_t st);
hipError_t launch_born(DevSys& s, hipStream_t st);
hipError_t launch_energy(DevSys& s, hipStream_t st);

}  // namespace alma

struct alma_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    std::string err;
    std::string name;
    alma::DevSys sys;
    std::vector<double> h_pos;       // host copy of the current coordinates
    std::vector<double> h_q, h_r, h_e;
    bool have_eval = false;
    // MC trace of the last run
    std::vector<int> tr_kind, tr_bond, tr_acc;
    std::vector<double> tr_delta, tr_dE;
};
