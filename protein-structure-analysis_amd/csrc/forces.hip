// forces.hip -- gradient of calculate_potential on gfx950 (no reference
// counterpart: the reference has no forces).  Exact derivative of total_e
// (src/physics_engine.cpp:2205-2234) with the neighbour list and the 12 A
// cutoff held fixed:
//   * pair terms: radial derivative of DH + GB + LJ + hard-core (FP32 per
//     pair, FP64 accumulation) and dE/da of the GB term;
//   * Born chain rule: a = 1/(1/R* - S/2) for unclamped atoms, da/dS = a^2/2,
//     dS/dr from the HCT integrand (src/physics_engine.cpp:2007-2011), both
//     directions of every neighbour entry;
//   * SASA: adjoint of the ordered fold (clamped steps scale the adjoint by
//     0.15, unclamped steps pass pi*R_i*(dc - r)'s derivative);
//   * dihedral and disulfide terms in FP64;
//   * torsion gradients dE/dtheta_k = u_k . sum_{i in side_k} (x_i - o_k) x G_i,
//     from index-ordered prefix sums over each side's runs.
#include "physics_dev.h"

namespace alma {

namespace {

using namespace dev;

__device__ __forceinline__ void atomic_add3(double* g, int i, double s, const double* u) {
    atomicAdd(&g[3 * i], s * u[0]);
    atomicAdd(&g[3 * i + 1], s * u[1]);
    atomicAdd(&g[3 * i + 2], s * u[2]);
}

// d hct(r; ri, rj)/dr (FP32)
__device__ __forceinline__ float dhct_f(float r, float ri, float rj) {
    float U = r + rj;
    if (ri >= U) return 0.0f;
    float d = r - rj;
    float L = fmaxf(fabsf(d), ri);
    float dL = fabsf(d) > ri ? (d > 0.0f ? 1.0f : -1.0f) : 0.0f;
    float c = ri * ri - rj * rj;
    float ir = rcp_f(r), iri2 = rcp_f(ri * ri);
    float B = (r * r + c) * 0.25f * iri2 * ir * ir;
    float dB = -0.5f * c * iri2 * ir * ir * ir;
    float lg = log_f(L * rcp_f(U));
    float iL = rcp_f(L), iU = rcp_f(U);
    return -dL * iL * iL + iU * iU + dB * lg + B * (dL * iL - iU);
}

// pair derivatives: dE/dr and dE/dD (D = a_i a_j) of DH + GB + LJ + hard-core
__device__ __forceinline__ void dpair(double r2d, float4 pi, float4 pj, double Ri, double Rj, double ei,
                                      double ej, double* dEdr, double* dEdD) {
    float r2 = (float)r2d, r = sqrt_f(r2), ir = rcp_f(r);
    float qp = pi.x * pj.x;
    float edh = (float)(kCoulomb / kEpsWater) * qp * ir * exp_f(-(float)kKappa * r);
    float dd = edh * (-ir - (float)kKappa);
    float D = pi.w * pj.w;
    float x = r2 * rcp_f(4.0f * D);
    float g = x < 87.0f ? exp_f(-x) : 0.0f;
    float f = sqrt_f(r2 + D * g), iff = rcp_f(f);
    float egb = (float)kGbCoef * qp * iff;
    float dfdr = (2.0f * r - 0.5f * r * g) * 0.5f * iff;
    float dgb = -egb * iff * dfdr;
    float dfdD = g * (1.0f + x) * 0.5f * iff;
    double sig = Ri + Rj;
    double s2 = sig * sig / r2d, rr = sqrt(r2d);
    double s6 = s2 * s2 * s2;
    double dlj = 4.0 * sqrt(ei * ej) * (-12.0 * s6 * s6 + 6.0 * s6) / rr;
    double dhc = 0.0;
    double rc = sig * kHardFrac;
    if (r2d < rc * rc) {
        double c2 = rc * rc / r2d, c6 = c2 * c2 * c2;
        double p12 = c6 * c6, xx = p12 - 1.0;
        if (kHardScale * xx * xx < kHardCap) dhc = kHardScale * 2.0 * xx * (-12.0 * p12 / rr);
    }
    *dEdr = (double)dd + (double)dgb + dlj + dhc;
    *dEdD = (double)(-egb * iff * dfdD);
}

// pass 1: pair radial forces and dE/da
__global__ void __launch_bounds__(kBlock) k_f_pair(DevSys s, double* __restrict__ G,
                                                  double* __restrict__ dEda) {
    int i = blockIdx.x * kWavesPerBlock + (threadIdx.x / kWave);
    if (i >= s.n) return;
    int lane = lane_id();
    const double* pos = s.pos;
    const double pi[3] = {pos[4 * i], pos[4 * i + 1], pos[4 * i + 2]};
    float4 qi = s.prm[i];
    int m = min(s.nl_cnt[i], s.cap);
    const uint32_t* row = s.nl + (size_t)i * s.cap;
    double gi[3] = {0, 0, 0}, dai = 0.0, dh = 0, gb = 0, lj = 0, hc = 0;
    uint32_t e_next = lane < m ? row[lane] : 0u;   // next row entries in flight
    for (int k = lane; k < m; k += kWave) {
        const uint32_t e = e_next;
        if (k + kWave < m) e_next = row[k + kWave];
        if (s.has_topo && (e & kExclBit)) continue;
        int j = (int)(e & kIdxMask);
        double r2 = dist2_exact(pi, pos + 4 * j);
        if (r2 > kPairCut2) continue;
        float4 qj = s.prm[j];
        PairTerms pt = pair_terms(r2, qi, qj, s.rstar[i], s.rstar[j], s.eps64[i], s.eps64[j]);
        dh += pt.dh;
        gb += pt.gb;
        lj += pt.lj;
        hc += pt.hc;
        double dr, dD;
        dpair(r2, qi, qj, s.rstar[i], s.rstar[j], s.eps64[i], s.eps64[j], &dr, &dD);
        double r = sqrt(r2);
        double u[3] = {(pi[0] - pos[4 * j]) / r, (pi[1] - pos[4 * j + 1]) / r, (pi[2] - pos[4 * j + 2]) / r};
        for (int d = 0; d < 3; ++d) gi[d] += dr * u[d];
        atomic_add3(G, j, -dr, u);
        dai += dD * (double)qj.w;
        atomicAdd(&dEda[j], dD * (double)qi.w);
    }
    for (int d = 0; d < 3; ++d) gi[d] = wave_sum(gi[d]);
    dai = wave_sum(dai);
    dh = wave_sum(dh);
    gb = wave_sum(gb);
    lj = wave_sum(lj);
    hc = wave_sum(hc);
    if (lane == 0) {
        for (int d = 0; d < 3; ++d) atomicAdd(&G[3 * i + d], gi[d]);
        atomicAdd(&dEda[i], dai);
        double* part = s.part + 5 * (size_t)i;
        part[0] = dh;
        part[1] = gb;
        part[2] = lj;
        part[3] = hc;
    }
}

// pass 2: Born chain rule over every neighbour entry (both directions)
__global__ void __launch_bounds__(kBlock) k_f_born(DevSys s, const double* __restrict__ dEda,
                                                  double* __restrict__ G) {
    int i = blockIdx.x * kWavesPerBlock + (threadIdx.x / kWave);
    if (i >= s.n) return;
    int lane = lane_id();
    const double* pos = s.pos;
    // a = 1/max(1/R - S/2, 2): unclamped while a < 0.5
    auto weight = [&](int k) -> double {
        double a = s.a[k];
        return a < 0.5 ? dEda[k] * 0.5 * a * a : 0.0;
    };
    double wi = weight(i);
    const double pi[3] = {pos[4 * i], pos[4 * i + 1], pos[4 * i + 2]};
    float ri = s.prm[i].y;
    int m = min(s.nl_cnt[i], s.cap);
    const uint32_t* row = s.nl + (size_t)i * s.cap;
    double gi[3] = {0, 0, 0};
    uint32_t e_next = lane < m ? row[lane] : 0u;   // next row entries in flight
    for (int k = lane; k < m; k += kWave) {
        const uint32_t e_cur = e_next;
        if (k + kWave < m) e_next = row[k + kWave];
        int j = (int)(e_cur & kIdxMask);
        double wj = weight(j);
        if (wi == 0.0 && wj == 0.0) continue;
        double r2 = dist2_exact(pi, pos + 4 * j), r = sqrt(r2);
        float rf = (float)r, rj = s.prm[j].y;
        double d = wi * (double)dhct_f(rf, ri, rj) + wj * (double)dhct_f(rf, rj, ri);
        double u[3] = {(pi[0] - pos[4 * j]) / r, (pi[1] - pos[4 * j + 1]) / r, (pi[2] - pos[4 * j + 2]) / r};
        for (int q = 0; q < 3; ++q) gi[q] += d * u[q];
        atomic_add3(G, j, -d, u);
    }
    for (int d = 0; d < 3; ++d) gi[d] = wave_sum(gi[d]);
    if (lane == 0)
        for (int d = 0; d < 3; ++d) atomicAdd(&G[3 * i + d], gi[d]);
}

// pass 3: SASA fold adjoint, one wavefront per atom
__global__ void __launch_bounds__(kBlock) k_f_sasa(DevSys s, double* __restrict__ G) {
    int i = blockIdx.x * kWavesPerBlock + (threadIdx.x / kWave);
    if (i >= s.n) return;
    const double pi[3] = {s.pos[4 * i], s.pos[4 * i + 1], s.pos[4 * i + 2]};
    const double ri = (double)s.prm[i].y + kProbe;
    const uint32_t* row = s.nl + (size_t)i * s.cap;
    const int m = min(s.nl_cnt[i], s.cap);
    int clamped;
    double sa = sasa_fold_wave(pi, s.pos, s.prm, ri, row, m, &clamped);
    if (lane_id() == 0) s.part[5 * (size_t)i + 4] = kGammaSA * fmax(0.0, sa);
    if (!(sa > 0.0)) return;
    double gi[3] = {0, 0, 0};
    sasa_adjoint_wave(pi, s.pos, s.prm, ri, row, m, clamped, G, gi);
    for (int d = 0; d < 3; ++d) gi[d] = wave_sum(gi[d]);
    if (lane_id() == 0)
        for (int d = 0; d < 3; ++d) atomicAdd(&G[3 * i + d], gi[d]);
}

__device__ __forceinline__ void cross3(const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

// pass 4: dihedral records and disulfide pairs (FP64, reverse mode through
// the reference's atan2 form, src/physics_engine.cpp:1871-1885)
__global__ void k_f_bonded(DevSys s, double* __restrict__ G) {
    const DevTopo& t = s.topo;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    const double* pos = s.pos;
    if (k < t.n_dih) {
        int a = t.dih_a[k], b = t.dih_b[k], c = t.dih_c[k], d = t.dih_d[k];
        double b1[3], b2[3], b3[3], n1[3], n2[3], m1[3];
        for (int q = 0; q < 3; ++q) {
            b1[q] = pos[4 * a + q] - pos[4 * b + q];
            b2[q] = pos[4 * c + q] - pos[4 * b + q];
            b3[q] = pos[4 * d + q] - pos[4 * c + q];
        }
        cross3(b1, b2, n1);
        cross3(b2, b3, n2);
        cross3(n1, b2, m1);
        double X = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
        double Y = m1[0] * n2[0] + m1[1] * n2[1] + m1[2] * n2[2];
        double den = X * X + Y * Y;
        double phi = dih_angle_d(pos, a, b, c, d);
        double gphi = 0.0, e = 0.0;
        for (int m = t.dih_toff[k]; m < t.dih_toff[k + 1]; ++m) {
            double arg = (double)t.dih_n[m] * phi - t.dih_gamma[m];
            e += t.dih_v2[m] * (1.0 + cos(arg));
            gphi += -t.dih_v2[m] * t.dih_n[m] * sin(arg);
        }
        s.rec[k] = e;
        if (den == 0.0) return;
        double gX = -Y / den * gphi, gY = X / den * gphi;
        double gn1[3], gn2[3], gm1[3], gb1[3] = {0, 0, 0}, gb2[3] = {0, 0, 0}, gb3[3] = {0, 0, 0}, tt[3];
        for (int q = 0; q < 3; ++q) {
            gn1[q] = gX * n2[q];
            gn2[q] = gX * n1[q] + gY * m1[q];
            gm1[q] = gY * n2[q];
        }
        cross3(b2, gm1, tt);
        for (int q = 0; q < 3; ++q) gn1[q] += tt[q];
        cross3(gm1, n1, tt);
        for (int q = 0; q < 3; ++q) gb2[q] += tt[q];
        cross3(b3, gn2, tt);
        for (int q = 0; q < 3; ++q) gb2[q] += tt[q];
        cross3(gn2, b2, tt);
        for (int q = 0; q < 3; ++q) gb3[q] += tt[q];
        cross3(b2, gn1, tt);
        for (int q = 0; q < 3; ++q) gb1[q] += tt[q];
        cross3(gn1, b1, tt);
        for (int q = 0; q < 3; ++q) gb2[q] += tt[q];
        for (int q = 0; q < 3; ++q) {
            atomicAdd(&G[3 * a + q], gb1[q]);
            atomicAdd(&G[3 * b + q], -gb1[q] - gb2[q]);
            atomicAdd(&G[3 * c + q], gb2[q] - gb3[q]);
            atomicAdd(&G[3 * d + q], gb3[q]);
        }
    } else if (k < t.n_dih + t.n_ss) {
        int q0 = k - t.n_dih;
        int i = t.ss_i[q0], j = t.ss_j[q0];
        double dx = pos[4 * i] - pos[4 * j], dy = pos[4 * i + 1] - pos[4 * j + 1], dz = pos[4 * i + 2] - pos[4 * j + 2];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        s.rec[k] = kKss * (r - kR0ss) * (r - kR0ss);
        double sc = 2.0 * kKss * (r - kR0ss) / r;
        for (int q = 0; q < 3; ++q) {
            double v = sc * (pos[4 * i + q] - pos[4 * j + q]);
            atomicAdd(&G[3 * i + q], v);
            atomicAdd(&G[3 * j + q], -v);
        }
    }
}

// pass 5: inclusive prefix sums over atoms of x_i x G_i and G_i, one block
// per component (6 lanes of FP64): wavefront scans, then a scan of the 16
// wavefront totals; then dE/dtheta_k from each side's runs
__global__ void __launch_bounds__(1024) k_f_prefix(DevSys s, const double* __restrict__ G,
                                                   double* __restrict__ pre) {
    __shared__ double wtot[16];
    const int comp = blockIdx.x;
    const int n = s.n;
    const double* pos = s.pos;
    const int lane = lane_id(), w = threadIdx.x / kWave;
    double* out = pre + (size_t)comp * (n + 1);
    double carry = 0.0;
    for (int base = 0; base < n; base += 1024) {
        int i = base + threadIdx.x;
        double v = 0.0;
        if (i < n) {
            if (comp < 3) {
                int a1 = (comp + 1) % 3, a2 = (comp + 2) % 3;
                v = pos[4 * i + a1] * G[3 * i + a2] - pos[4 * i + a2] * G[3 * i + a1];
            } else {
                v = G[3 * i + comp - 3];
            }
        }
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            double t = __shfl_up(v, o, kWave);
            if (lane >= o) v += t;
        }
        if (lane == kWave - 1) wtot[w] = v;
        __syncthreads();
        double before = carry;
        for (int k = 0; k < w; ++k) before += wtot[k];
        if (i < n) out[i + 1] = before + v;
        double tot = 0.0;
        for (int k = 0; k < 16; ++k) tot += wtot[k];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) out[0] = 0.0;
}

__global__ void k_f_torsion(DevSys s, const double* __restrict__ pre, double* __restrict__ dEdt) {
    const DevTopo& t = s.topo;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= t.n_rot) return;
    const int n = s.n;
    const double* pos = s.pos;
    int bi = t.rb_i[k], bj = t.rb_j[k];
    double o[3], u[3];
    for (int d = 0; d < 3; ++d) {
        o[d] = pos[4 * bi + d];
        u[d] = pos[4 * bj + d] - o[d];
    }
    double l = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    if (l < 1e-10) {
        dEdt[k] = 0.0;
        return;
    }
    for (int d = 0; d < 3; ++d) u[d] /= l;
    double T[3] = {0, 0, 0}, S[3] = {0, 0, 0};
    for (int q = t.side_run_off[k]; q < t.side_run_off[k + 1]; ++q) {
        int2 r = t.side_run[q];
        for (int d = 0; d < 3; ++d) {
            T[d] += pre[(size_t)d * (n + 1) + r.y + 1] - pre[(size_t)d * (n + 1) + r.x];
            S[d] += pre[(size_t)(3 + d) * (n + 1) + r.y + 1] - pre[(size_t)(3 + d) * (n + 1) + r.x];
        }
    }
    double oxS[3];
    cross3(o, S, oxS);
    // u . sum (x_i - o) x G_i
    dEdt[k] = u[0] * (T[0] - oxS[0]) + u[1] * (T[1] - oxS[1]) + u[2] * (T[2] - oxS[2]);
}

__global__ void k_f_zero(double* __restrict__ a, int na, double* __restrict__ b, int nb) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0.0;
    else if (i < na + nb) b[i - na] = 0.0;
}

inline unsigned blocks_for(int n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// Energy terms and gradient in one sequence after the list and Born radii:
// the pair, SASA and bonded passes write the per-atom / per-record energy
// partials the energy path would, then the terms are reduced.
hipError_t launch_forces(DevSys& s, double* G, double* dEda, double* pre, double* dEdt, hipStream_t st) {
    const int n = s.n;
    k_f_zero<<<blocks_for(4 * n, kBlock), kBlock, 0, st>>>(G, 3 * n, dEda, n);
    k_f_pair<<<blocks_for(n, kWavesPerBlock), kBlock, 0, st>>>(s, G, dEda);
    k_f_born<<<blocks_for(n, kWavesPerBlock), kBlock, 0, st>>>(s, dEda, G);
    k_f_sasa<<<blocks_for(n, kWavesPerBlock), kBlock, 0, st>>>(s, G);
    int nrec = s.has_topo ? s.topo.n_dih + s.topo.n_ss : 0;
    if (nrec > 0) k_f_bonded<<<blocks_for(nrec, kBlock), kBlock, 0, st>>>(s, G);
    if (s.has_topo && s.topo.n_rot > 0 && dEdt) {
        k_f_prefix<<<6, 1024, 0, st>>>(s, G, pre);
        k_f_torsion<<<blocks_for(s.topo.n_rot, kBlock), kBlock, 0, st>>>(s, pre, dEdt);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_terms(s, st);
}

}  // namespace alma
