// alma_abi.cpp -- host side of the C ABI declared in include/alma.h:
// context lifetime, device buffers, launch sequences and error mapping.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "alma_internal.h"

using namespace alma;

namespace {

struct Fail {
    alma_status code;
    std::string msg;
};

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw Fail{ALMA_ERUNTIME, std::string(#expr) + ": " + hipGetErrorString(e_)};     \
    } while (0)

template <class T>
void dev_alloc(T** p, size_t count) {
    if (*p) {
        hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc((void**)p, sizeof(T) * count));
}

template <class T>
void dev_upload(T** p, const T* src, size_t count) {
    dev_alloc(p, count);
    if (count && src) HIP_TRY(hipMemcpy(*p, src, sizeof(T) * count, hipMemcpyHostToDevice));
}

template <class T>
void dev_free(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

void free_topo(DevTopo& t) {
    dev_free(t.excl_off);
    dev_free(t.excl_idx);
    dev_free(t.rb_i);
    dev_free(t.rb_j);
    dev_free(t.rb_kind);
    dev_free(t.rb_scale);
    dev_free(t.side_off);
    dev_free(t.side_idx);
    dev_free(t.side_run);
    dev_free(t.side_run_off);
    dev_free(t.dih_a);
    dev_free(t.dih_b);
    dev_free(t.dih_c);
    dev_free(t.dih_d);
    dev_free(t.dih_toff);
    dev_free(t.dih_v2);
    dev_free(t.dih_n);
    dev_free(t.dih_gamma);
    dev_free(t.ss_i);
    dev_free(t.ss_j);
    dev_free(t.cp_phi);
    dev_free(t.cp_psi);
    dev_free(t.csp_a);
    dev_free(t.csp_b);
    t = DevTopo{};
}

void free_sys(DevSys& s) {
    dev_free(s.pos);
    dev_free(s.prm);
    dev_free(s.rstar);
    dev_free(s.born_sum);
    dev_free(s.half_sum);
    dev_free(s.a);
    dev_free(s.cell_of);
    dev_free(s.cell_cnt);
    dev_free(s.cell_start);
    dev_free(s.cell_atoms);
    dev_free(s.cell_tmp);
    dev_free(s.nl_cnt);
    dev_free(s.nl);
    dev_free(s.part);
    dev_free(s.rec);
    dev_free(s.terms);
    dev_free(s.flags);
    dev_free(s.grid);
    free_topo(s.topo);
    s = DevSys{};
}

template <class F>
alma_status guarded(alma_ctx* ctx, F&& f) {
    if (!ctx) return ALMA_EINVAL;
    try {
        f();
        return ALMA_OK;
    } catch (const Fail& e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->err = "host allocation failed";
        return ALMA_ERUNTIME;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return ALMA_ERUNTIME;
    }
}

void validate_topo(const alma_topo* t, int n) {
    auto bad = [](const char* m) { throw Fail{ALMA_EINVAL, m}; };
    if (t->n_atoms != n) bad("topology atom count differs from the particle count");
    if (!t->excl_off || !t->excl_idx) bad("topology exclusions missing");
    if (t->n_rot < 0 || t->n_dih < 0 || t->n_ss < 0 || t->n_cp < 0 || t->n_csp < 0)
        bad("negative topology count");
    for (int i = 0; i < n; ++i)
        for (int k = t->excl_off[i]; k < t->excl_off[i + 1]; ++k)
            if (t->excl_idx[k] <= i || t->excl_idx[k] >= n) bad("exclusion index out of range");
    auto chk = [&](const int32_t* v, int cnt) {
        for (int k = 0; k < cnt; ++k)
            if (v[k] < 0 || v[k] >= n) bad("topology atom index out of range");
    };
    chk(t->rb_i, t->n_rot);
    chk(t->rb_j, t->n_rot);
    if (t->n_rot) chk(t->side_idx, t->side_off[t->n_rot]);
    chk(t->dih_a, t->n_dih);
    chk(t->dih_b, t->n_dih);
    chk(t->dih_c, t->n_dih);
    chk(t->dih_d, t->n_dih);
    chk(t->ss_i, t->n_ss);
    chk(t->ss_j, t->n_ss);
    for (int k = 0; k < t->n_cp; ++k)
        if (t->cp_phi[k] < 0 || t->cp_phi[k] >= t->n_rot || t->cp_psi[k] < 0 || t->cp_psi[k] >= t->n_rot)
            bad("concerted pair bond index out of range");
    for (int k = 0; k < t->n_csp; ++k)
        if (t->csp_a[k] < 0 || t->csp_a[k] >= t->n_rot || t->csp_b[k] < 0 || t->csp_b[k] >= t->n_rot)
            bad("side-chain pair bond index out of range");
}

// j-side atom sets as sorted runs of consecutive indices
void side_runs(const alma_topo* t, std::vector<int2>& runs, std::vector<int>& off) {
    off.assign(t->n_rot + 1, 0);
    runs.clear();
    std::vector<int> v;
    for (int b = 0; b < t->n_rot; ++b) {
        v.assign(t->side_idx + t->side_off[b], t->side_idx + t->side_off[b + 1]);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        for (size_t k = 0; k < v.size();) {
            size_t e = k;
            while (e + 1 < v.size() && v[e + 1] == v[e] + 1) ++e;
            runs.push_back(make_int2(v[k], v[e]));
            k = e + 1;
        }
        off[b + 1] = (int)runs.size();
    }
}

void upload_topo(DevTopo& d, const alma_topo* t, int n) {
    free_topo(d);
    dev_upload(&d.excl_off, t->excl_off, (size_t)n + 1);
    dev_upload(&d.excl_idx, t->excl_idx, (size_t)t->excl_off[n]);
    d.n_rot = t->n_rot;
    d.n_dih = t->n_dih;
    d.n_ss = t->n_ss;
    d.n_cp = t->n_cp;
    d.n_csp = t->n_csp;
    dev_upload(&d.rb_i, t->rb_i, t->n_rot);
    dev_upload(&d.rb_j, t->rb_j, t->n_rot);
    dev_upload(&d.rb_kind, t->rb_kind, t->n_rot);
    dev_upload(&d.rb_scale, t->rb_scale, t->n_rot);
    if (t->n_rot) {
        dev_upload(&d.side_off, t->side_off, (size_t)t->n_rot + 1);
        dev_upload(&d.side_idx, t->side_idx, (size_t)t->side_off[t->n_rot]);
        std::vector<int2> runs;
        std::vector<int> roff;
        side_runs(t, runs, roff);
        dev_upload(&d.side_run, runs.data(), runs.size());
        dev_upload(&d.side_run_off, roff.data(), roff.size());
    }
    dev_upload(&d.dih_a, t->dih_a, t->n_dih);
    dev_upload(&d.dih_b, t->dih_b, t->n_dih);
    dev_upload(&d.dih_c, t->dih_c, t->n_dih);
    dev_upload(&d.dih_d, t->dih_d, t->n_dih);
    if (t->n_dih) {
        dev_upload(&d.dih_toff, t->dih_toff, (size_t)t->n_dih + 1);
        int nt = t->dih_toff[t->n_dih];
        dev_upload(&d.dih_v2, t->dih_v2, nt);
        dev_upload(&d.dih_n, t->dih_n, nt);
        dev_upload(&d.dih_gamma, t->dih_gamma, nt);
    }
    dev_upload(&d.ss_i, t->ss_i, t->n_ss);
    dev_upload(&d.ss_j, t->ss_j, t->n_ss);
    dev_upload(&d.cp_phi, t->cp_phi, t->n_cp);
    dev_upload(&d.cp_psi, t->cp_psi, t->n_cp);
    dev_upload(&d.csp_a, t->csp_a, t->n_csp);
    dev_upload(&d.csp_b, t->csp_b, t->n_csp);
}

// number of cells the host-side bounding box implies, times a growth margin
int cells_needed(const std::vector<double>& xyz, int n) {
    double lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], xyz[3 * i + d]);
            hi[d] = std::max(hi[d], xyz[3 * i + d]);
        }
    const double cs = std::sqrt(kNlRcut2) + 1e-6;
    long c = 1;
    for (int d = 0; d < 3; ++d) c *= (long)std::ceil((hi[d] - lo[d] + 3.0 * cs) / cs) * 2 + 2;
    return (int)std::min<long>(c, 1L << 26);
}

void upload_positions(alma_ctx* ctx) {
    DevSys& s = ctx->sys;
    std::vector<double> p4(4 * (size_t)s.n, 0.0);
    for (int i = 0; i < s.n; ++i)
        for (int d = 0; d < 3; ++d) p4[4 * i + d] = ctx->h_pos[3 * i + d];
    HIP_TRY(hipMemcpy(s.pos, p4.data(), sizeof(double) * p4.size(), hipMemcpyHostToDevice));
}

// neighbour list + Born radii + terms, growing the per-atom capacity when a
// row overflows
void evaluate(alma_ctx* ctx, double* terms) {
    DevSys& s = ctx->sys;
    for (int attempt = 0; attempt < 3; ++attempt) {
        HIP_TRY(launch_build_nl(s, ctx->stream));
        int flags[2];
        HIP_TRY(hipMemcpyAsync(flags, s.flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (flags[0] & 2) throw Fail{ALMA_ERUNTIME, "cell grid exceeds its allocation"};
        if (flags[0] & 1) {
            s.cap = ((flags[1] * 5 / 4) + 63) / 64 * 64;
            dev_alloc(&s.nl, (size_t)s.n * s.cap);
            continue;
        }
        HIP_TRY(launch_born(s, ctx->stream));
        HIP_TRY(launch_energy(s, ctx->stream));
        if (terms)
            HIP_TRY(hipMemcpyAsync(terms, s.terms, sizeof(double) * ALMA_NTERMS, hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->have_eval = true;
        return;
    }
    throw Fail{ALMA_ERUNTIME, "neighbour list capacity could not be sized"};
}

}  // namespace

extern "C" {

alma_status alma_create(int device, alma_ctx** out) {
    if (!out) return ALMA_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return ALMA_ENODEV;
    if (device < 0 || device >= count) return ALMA_EINVAL;
    alma_ctx* c = new (std::nothrow) alma_ctx();
    if (!c) return ALMA_ERUNTIME;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return ALMA_ENODEV;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->name = prop.name;
        if (c->name.empty()) c->name = std::string("AMD ") + prop.gcnArchName;
        c->n_cu = prop.multiProcessorCount;
    }
    *out = c;
    return ALMA_OK;
}

void alma_destroy(alma_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    free_sys(ctx->sys);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* alma_last_error(const alma_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

alma_status alma_device_name(alma_ctx* ctx, char* buf, int32_t len) {
    return guarded(ctx, [&] {
        if (!buf || len <= 0) throw Fail{ALMA_EINVAL, "empty name buffer"};
        std::strncpy(buf, ctx->name.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    });
}

alma_status alma_set_system(alma_ctx* ctx, int32_t n, const double* xyz, const double* q,
                            const double* rstar, const double* eps, const alma_topo* topo) {
    return guarded(ctx, [&] {
        if (n < 0 || (n > 0 && (!xyz || !q || !rstar || !eps)))
            throw Fail{ALMA_EINVAL, "particle arrays missing"};
        if (topo) validate_topo(topo, n);
        for (int i = 0; i < n; ++i)
            if (!(rstar[i] > 0.0)) throw Fail{ALMA_EINVAL, "particle radius must be positive"};
        HIP_TRY(hipSetDevice(ctx->device));
        free_sys(ctx->sys);
        DevSys& s = ctx->sys;
        s.n = n;
        ctx->have_eval = false;
        ctx->h_pos.assign(xyz, xyz + 3 * (size_t)n);
        ctx->h_q.assign(q, q + n);
        ctx->h_r.assign(rstar, rstar + n);
        ctx->h_e.assign(eps, eps + n);
        if (n == 0) return;
        dev_alloc(&s.pos, 4 * (size_t)n);
        std::vector<float4> prm(n);
        for (int i = 0; i < n; ++i) prm[i] = make_float4((float)q[i], (float)rstar[i], (float)eps[i], 0.f);
        dev_upload(&s.prm, prm.data(), n);
        dev_upload(&s.rstar, rstar, n);
        dev_alloc(&s.born_sum, n);
        dev_alloc(&s.half_sum, n);
        dev_alloc(&s.a, n);
        dev_alloc(&s.cell_of, n);
        s.max_cells = cells_needed(ctx->h_pos, n);
        dev_alloc(&s.cell_cnt, (size_t)s.max_cells + 1);
        dev_alloc(&s.cell_start, (size_t)s.max_cells + 1);
        dev_alloc(&s.cell_atoms, n);
        dev_alloc(&s.cell_tmp, n);
        dev_alloc(&s.nl_cnt, n);
        s.cap = 512;
        dev_alloc(&s.nl, (size_t)n * s.cap);
        dev_alloc(&s.part, 5 * (size_t)n);
        dev_alloc(&s.terms, ALMA_NTERMS);
        dev_alloc(&s.flags, 2);
        Grid g{};
        g.cs = std::sqrt(kNlRcut2) + 1e-6;
        dev_upload(&s.grid, &g, 1);
        if (topo) {
            upload_topo(s.topo, topo, n);
            s.has_topo = true;
            dev_alloc(&s.rec, (size_t)topo->n_dih + topo->n_ss);
        }
        upload_positions(ctx);
    });
}

alma_status alma_set_positions(alma_ctx* ctx, const double* xyz) {
    return guarded(ctx, [&] {
        if (!xyz) throw Fail{ALMA_EINVAL, "positions missing"};
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->h_pos.assign(xyz, xyz + 3 * (size_t)ctx->sys.n);
        ctx->have_eval = false;
        if (ctx->sys.n) upload_positions(ctx);
    });
}

alma_status alma_energy(alma_ctx* ctx, double* terms) {
    return guarded(ctx, [&] {
        if (!terms) throw Fail{ALMA_EINVAL, "terms buffer missing"};
        HIP_TRY(hipSetDevice(ctx->device));
        if (ctx->sys.n == 0) {   // calculate_potential returns 0 for no particles
            std::fill(terms, terms + ALMA_NTERMS, 0.0);
            return;
        }
        evaluate(ctx, terms);
    });
}

alma_status alma_nlist(alma_ctx* ctx, int32_t* offs, int32_t* js, int64_t js_cap, int64_t* npairs) {
    return guarded(ctx, [&] {
        if (!offs || !npairs) throw Fail{ALMA_EINVAL, "output buffers missing"};
        HIP_TRY(hipSetDevice(ctx->device));
        DevSys& s = ctx->sys;
        if (s.n == 0) {
            offs[0] = 0;
            *npairs = 0;
            return;
        }
        if (!ctx->have_eval) evaluate(ctx, nullptr);
        std::vector<int> cnt(s.n);
        HIP_TRY(hipMemcpy(cnt.data(), s.nl_cnt, sizeof(int) * s.n, hipMemcpyDeviceToHost));
        offs[0] = 0;
        for (int i = 0; i < s.n; ++i) offs[i + 1] = offs[i] + cnt[i];
        *npairs = offs[s.n];
        if (js && js_cap >= *npairs) {
            std::vector<uint32_t> rows((size_t)s.n * s.cap);
            HIP_TRY(hipMemcpy(rows.data(), s.nl, sizeof(uint32_t) * rows.size(), hipMemcpyDeviceToHost));
            for (int i = 0; i < s.n; ++i)
                for (int k = 0; k < cnt[i]; ++k) js[offs[i] + k] = (int32_t)(rows[(size_t)i * s.cap + k] & kIdxMask);
        }
    });
}

alma_status alma_born_radii(alma_ctx* ctx, double* a) {
    return guarded(ctx, [&] {
        if (!a) throw Fail{ALMA_EINVAL, "output buffer missing"};
        HIP_TRY(hipSetDevice(ctx->device));
        if (ctx->sys.n == 0) return;
        if (!ctx->have_eval) evaluate(ctx, nullptr);
        HIP_TRY(hipMemcpy(a, ctx->sys.a, sizeof(double) * ctx->sys.n, hipMemcpyDeviceToHost));
    });
}

alma_status alma_mc_run(alma_ctx* ctx, const alma_mc_params* p, double* xyz_out, alma_mc_stats* stats,
                        double* snap_xyz, double* snap_e) {
    (void)p;
    (void)xyz_out;
    (void)stats;
    (void)snap_xyz;
    (void)snap_e;
    return guarded(ctx, [&] { throw Fail{ALMA_ERUNTIME, "device MC is not built into this library"}; });
}

alma_status alma_mc_trace(alma_ctx* ctx, int32_t* kind, int32_t* bond, double* delta, double* dE,
                          int32_t* accepted) {
    return guarded(ctx, [&] {
        size_t m = ctx->tr_kind.size();
        if (kind) std::copy(ctx->tr_kind.begin(), ctx->tr_kind.end(), kind);
        if (bond) std::copy(ctx->tr_bond.begin(), ctx->tr_bond.end(), bond);
        if (delta) std::copy(ctx->tr_delta.begin(), ctx->tr_delta.end(), delta);
        if (dE) std::copy(ctx->tr_dE.begin(), ctx->tr_dE.end(), dE);
        if (accepted) std::copy(ctx->tr_acc.begin(), ctx->tr_acc.end(), accepted);
        (void)m;
    });
}

alma_status alma_forces(alma_ctx* ctx, double* energy, double* dEdx, double* dEdtheta) {
    (void)energy;
    (void)dEdx;
    (void)dEdtheta;
    return guarded(ctx, [&] { throw Fail{ALMA_ERUNTIME, "device forces are not built into this library"}; });
}

}  // extern "C"
