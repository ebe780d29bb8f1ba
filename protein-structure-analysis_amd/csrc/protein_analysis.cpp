// protein_analysis.cpp -- the drop-in Python module `protein_analysis`.
//
// Same functions, argument names, defaults, return types and error types as
// the reference module (src/analysis_ext.cpp:364-392):
//   compute_mean_dist_matrix(coords), compute_contact_freq_matrix(coords,
//   cutoff=8.0): the O(n_snaps * n_ca^2) ensemble metrics, run by the gfx950
//   library behind include/alma.h (alma_ensemble_matrices), bit-equal to the
//   reference's serial loop;
//   kmt_reduce(poly, max_passes=200), find_crossings(poly, axis): the knot
//   pipeline's polygon reduction and projected-crossing scan, host code (each
//   vertex deletion depends on the previous one, and the reduced polygons
//   hold tens of vertices: no device work pays here).
// The device context is created on the first matrix call (device 0) and kept
// for the life of the process; the GIL is released around the device call.
// Matrix calls from several threads are serialised.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/alma.h"

namespace py = pybind11;

namespace alma_ana {

using Array = py::array_t<double, py::array::c_style | py::array::forcecast>;

// ── ensemble metrics on the device ──────────────────────────────────────────

std::mutex g_mu;
alma_ctx* g_ctx = nullptr;

alma_ctx* context() {
    if (!g_ctx) {
        alma_status s = alma_create(0, &g_ctx);
        if (s != ALMA_OK) {
            g_ctx = nullptr;
            throw std::runtime_error(s == ALMA_ENODEV ? "no GPU device available for protein_analysis"
                                                      : "failed to create the GPU context");
        }
    }
    return g_ctx;
}

// (n_snaps, n_ca, 3) layout check, src/analysis_ext.cpp:71-74
void check_coords(const py::buffer_info& b) {
    if (b.ndim != 3 || b.shape[2] != 3) throw std::invalid_argument("coords must be a 3-D array (n_snaps, n_ca, 3)");
}

py::array_t<double> matrix(Array coords, double cutoff, bool mean) {
    py::buffer_info b = coords.request();
    check_coords(b);
    const int64_t s = b.shape[0], n = b.shape[1];
    if (s > INT32_MAX || n > INT32_MAX) throw std::invalid_argument("coords too large");
    py::array_t<double> out({(py::ssize_t)n, (py::ssize_t)n});
    double* o = out.mutable_data();
    const double* x = static_cast<const double*>(b.ptr);
    alma_status st;
    std::string err;
    {
        py::gil_scoped_release rel;
        std::lock_guard<std::mutex> lk(g_mu);
        alma_ctx* c = context();
        st = alma_ensemble_matrices(c, x, (int32_t)s, (int32_t)n, cutoff, mean ? o : nullptr, mean ? nullptr : o);
        if (st != ALMA_OK) err = alma_last_error(c);
    }
    if (st == ALMA_EINVAL) throw std::invalid_argument(err);
    if (st != ALMA_OK) throw std::runtime_error(err);
    return out;
}

// ── knot pipeline: host code ───────────────────────────────────────────────

struct V3 {
    double x, y, z;
};
inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

constexpr double kEps = 1e-9;

// Does the open segment p0 -> p1 pass through triangle (a, b, c)?  The
// Moller-Trumbore ray / triangle test (Moller & Trumbore 1997) with the
// segment parameter restricted to (eps, 1 - eps) and the barycentric bounds
// widened by eps, as knot_analysis._segment_triangle_intersect decides it
// (src/analysis_ext.cpp:175-195).
bool segment_hits_triangle(const V3& p0, const V3& p1, const V3& a, const V3& b, const V3& c) {
    const V3 dir = sub(p1, p0), e1 = sub(b, a), e2 = sub(c, a);
    const V3 h = cross(dir, e2);
    const double det = dot(e1, h);
    if (std::fabs(det) < kEps) return false;
    const double f = 1.0 / det;
    const V3 s = sub(p0, a);
    const double u = dot(s, h) * f;
    if (u < -kEps || u > 1 + kEps) return false;
    const V3 q = cross(s, e1);
    const double v = dot(dir, q) * f;
    if (v < -kEps || u + v > 1 + kEps) return false;
    const double t = dot(e2, q) * f;
    return kEps < t && t < 1 - kEps;
}

// KMT reduction (src/analysis_ext.cpp:197-256): vertex i of the closed
// polygon is dropped when no segment other than the four touching the
// triangle (i-1, i, i+1) crosses that triangle.  The scan of a pass does not
// advance after a deletion (the vertex moved into slot i is tried next);
// passes repeat until one deletes nothing, the polygon has 3 vertices, or
// max_passes passes ran.
std::vector<V3> kmt(std::vector<V3> p, int max_passes) {
    for (int pass = 0; pass < max_passes && p.size() > 3; ++pass) {
        bool any = false;
        size_t i = 0;
        while (i < p.size() && p.size() > 3) {
            const int n = (int)p.size();
            const int prev = ((int)i + n - 1) % n, next = ((int)i + 1) % n;
            bool blocked = false;
            for (int j = 0; j < n && !blocked; ++j) {
                const int jn = (j + 1) % n;
                if (j == prev || j == (int)i || j == next || jn == prev) continue;
                blocked = segment_hits_triangle(p[j], p[jn], p[prev], p[i], p[next]);
            }
            if (blocked) {
                ++i;
            } else {
                p.erase(p.begin() + (ptrdiff_t)i);
                any = true;
            }
        }
        if (!any) break;
    }
    return p;
}

std::vector<V3> points(Array poly) {
    py::buffer_info b = poly.request();
    if (b.ndim != 2 || b.shape[1] != 3) throw std::invalid_argument("poly must be a (n,3) array");
    const double* d = static_cast<const double*>(b.ptr);
    std::vector<V3> p((size_t)b.shape[0]);
    for (size_t k = 0; k < p.size(); ++k) p[k] = {d[3 * k], d[3 * k + 1], d[3 * k + 2]};
    return p;
}

py::array_t<double> kmt_reduce(Array poly, int max_passes) {
    std::vector<V3> p = points(poly);
    {
        py::gil_scoped_release rel;
        p = kmt(std::move(p), max_passes);
    }
    py::array_t<double> out({(py::ssize_t)p.size(), (py::ssize_t)3});
    double* o = out.mutable_data();
    for (size_t k = 0; k < p.size(); ++k) {
        o[3 * k] = p[k].x;
        o[3 * k + 1] = p[k].y;
        o[3 * k + 2] = p[k].z;
    }
    return out;
}

// Crossings of the polygon projected along `axis` (src/analysis_ext.cpp:
// 272-361): the plane basis u = axis x e / |axis x e| (e = x unless
// |axis.x| >= 0.9, then y), v = axis x u; every pair of non-adjacent
// segments whose projections cross at interior parameters (t, s) in
// (eps, 1 - eps) gives a crossing, the segment deeper along the axis being
// over: over / under positions (segment index + parameter) and the sign of
// the projected cross product (negated when segment j is over).  A crossing
// with depths within 1e-9 or a cross product below 1e-12 makes the
// projection degenerate: (False, [], [], []).
py::tuple find_crossings(Array poly, Array axis_arr) {
    std::vector<V3> p = points(poly);
    py::buffer_info ab = axis_arr.request();
    if (ab.size != 3) throw std::invalid_argument("axis must be a length-3 array");
    const double* ad = static_cast<const double*>(ab.ptr);
    const V3 axis{ad[0], ad[1], ad[2]};
    const int n = (int)p.size();
    std::vector<double> under, over;
    std::vector<int> sgn;
    bool ok = true;
    {
        py::gil_scoped_release rel;
        const V3 e = std::fabs(axis.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
        V3 u = cross(axis, e);
        const double un = std::sqrt(dot(u, u));
        u = {u.x / un, u.y / un, u.z / un};
        const V3 v = cross(axis, u);
        std::vector<double> pu(n), pv(n), dep(n);
        for (int k = 0; k < n; ++k) {
            pu[k] = dot(p[k], u);
            pv[k] = dot(p[k], v);
            dep[k] = dot(p[k], axis);
        }
        for (int i = 0; i < n && ok; ++i) {
            const int i1 = (i + 1) % n;
            const double ax = pu[i], ay = pv[i];
            const double rx = pu[i1] - ax, ry = pv[i1] - ay;
            for (int j = i + 1; j < n; ++j) {
                const int j1 = (j + 1) % n;
                if (j1 == i || i1 == j) continue;   // segments sharing a vertex
                const double bx = pu[j], by = pv[j];
                const double sx = pu[j1] - bx, sy = pv[j1] - by;
                const double den = rx * sy - ry * sx;
                if (std::fabs(den) < kEps) continue;
                const double wx = bx - ax, wy = by - ay;
                const double t = (wx * sy - wy * sx) / den;
                const double s = (wx * ry - wy * rx) / den;
                if (!(t > kEps && t < 1 - kEps && s > kEps && s < 1 - kEps)) continue;
                const double di = dep[i] * (1 - t) + dep[i1] * t;
                const double dj = dep[j] * (1 - s) + dep[j1] * s;
                const double cz = rx * sy - ry * sx;
                if (std::fabs(di - dj) < 1e-9 || std::fabs(cz) < 1e-12) {
                    ok = false;
                    break;
                }
                if (di > dj) {
                    over.push_back(i + t);
                    under.push_back(j + s);
                    sgn.push_back(cz > 0 ? 1 : -1);
                } else {
                    over.push_back(j + s);
                    under.push_back(i + t);
                    sgn.push_back(cz < 0 ? 1 : -1);
                }
            }
        }
    }
    if (!ok) return py::make_tuple(false, py::array_t<double>(0), py::array_t<double>(0), py::array_t<int>(0));
    py::array_t<double> ua((py::ssize_t)under.size()), oa((py::ssize_t)over.size());
    py::array_t<int> sa((py::ssize_t)sgn.size());
    std::copy(under.begin(), under.end(), ua.mutable_data());
    std::copy(over.begin(), over.end(), oa.mutable_data());
    std::copy(sgn.begin(), sgn.end(), sa.mutable_data());
    return py::make_tuple(true, ua, oa, sa);
}

}  // namespace alma_ana

PYBIND11_MODULE(protein_analysis, m) {
    using namespace alma_ana;
    m.doc() = "ALMA ensemble metrics (gfx950) and knot-analysis helpers";
    m.def(
        "compute_mean_dist_matrix", [](Array coords) { return matrix(coords, 0.0, true); }, py::arg("coords"),
        "Ensemble-mean Calpha-Calpha distance matrix.\n\ncoords : (n_snaps, n_ca, 3) float64 array.\n"
        "Returns (n_ca, n_ca) float64 ndarray.");
    m.def(
        "compute_contact_freq_matrix", [](Array coords, double cutoff) { return matrix(coords, cutoff, false); },
        py::arg("coords"), py::arg("cutoff") = 8.0,
        "Fraction of snapshots each residue pair is within `cutoff` A.\n\ncoords : (n_snaps, n_ca, 3) float64 "
        "array.\nReturns (n_ca, n_ca) float64 ndarray (diagonal is 0).");
    m.def("kmt_reduce", &kmt_reduce, py::arg("poly"), py::arg("max_passes") = 200,
          "KMT triangle-elimination reduction of a closed 3D polygon.\n\npoly : (n, 3) float64 array.\n"
          "Returns (m, 3) float64 ndarray, m <= n.");
    m.def("find_crossings", &find_crossings, py::arg("poly"), py::arg("axis"),
          "2D-projected self-crossing detection.\n\npoly : (n, 3) float64 array. axis : length-3 float64 array.\n"
          "Returns (ok, under_pos, over_pos, sign); ok=False means a degenerate projection.");
}
