// Gravitational unbinding of FoF candidates (bound_check=True) on MI355X.
//
// Restates bound_fortran.f90's solver for every candidate at once:
//   packing + span       (fof_fortran.f90:263-308, bound_fortran.f90:2940-2977)
//   method selection     (host, bound_fortran.f90:1854-2082, 2901-3073)
//   direct summation     k_direct_solve: one workgroup per halo of <= 1000
//                        members, the whole pass/removal loop in LDS, FP64
//                        potentials summed in the reference's order
//   particle mesh        per grid class, all active halos batched per pass:
//                        TSC setup + self term, fixed-point deposit, FFT,
//                        x G(k), inverse FFT, interpolation (C1), short-range
//                        residual (C2), then k_pm_inner (removal loop)
//   write-back           valid = status OK and np_bound >= np_halo_min
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "bound.hpp"
#include "fft.hpp"
#include "prims.hpp"

namespace pyfof {

namespace {

constexpr int kB = 256;
constexpr int kDirectMax = 1000;  // bound_fortran.f90:1976-1980
constexpr int kDirectLargeMax = 65536;  // largest direct-summation threshold a context accepts
constexpr double kPi = 3.141592653589793238462643383279502884197;

inline uint32_t blocks(uint64_t n) { return div_up(n == 0 ? 1 : n, kB); }

struct Params {
    double app, a, b, L, Lh;
    int32_t np_min;
};

// ---------------------------------------------------------------------------
// depth kernels (bound_fortran.f90:3080-3117)

__device__ double depth_continuous(double r, double h) {
    if (r <= 0.0) return 208.0 / (70.0 * h);
    if (r >= h) return 1.0 / r;
    const double e = 2.0 * r / h;
    if (e <= 1.0) {
        const double prim = e * e * (112.0 + e * e * (-56.0 + e * (14.0 + e * (8.0 - 3.0 * e))));
        return (208.0 - prim) / (70.0 * h);
    }
    const double prim = e * e * (448.0 + e * (-280.0 + e * (56.0 + e * (14.0 + e * (-8.0 + e))))) -
                        224.0 * e - 12.0 / e;
    return (128.0 - prim) / (70.0 * h);
}

// the same kernel in f32 (mixed-precision direct summation)
__device__ float depth_continuous_f(float r, float h) {
    if (r <= 0.0f) return 208.0f / (70.0f * h);
    if (r >= h) return 1.0f / r;
    const float e = 2.0f * r / h;
    if (e <= 1.0f) {
        const float prim = e * e * (112.0f + e * e * (-56.0f + e * (14.0f + e * (8.0f - 3.0f * e))));
        return (208.0f - prim) / (70.0f * h);
    }
    const float prim = e * e * (448.0f + e * (-280.0f + e * (56.0f + e * (14.0f + e * (-8.0f + e))))) -
                       224.0f * e - 12.0f / e;
    return (128.0f - prim) / (70.0f * h);
}

// one pair term w_j (a/6pi) D(|q_j - q_i|): the reference's form in f64 (f32
// differences widened, bound_fortran.f90:2616-2628) or all in f32
template <bool F32>
__device__ __forceinline__ double pair_term(const float4& qi, const float4& qj, double Gd, float Gf,
                                            double app, float appf) {
    if (F32) {
        const float dx = qj.x - qi.x, dy = qj.y - qi.y, dz = qj.z - qi.z;
        const float r2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
        return (double)__fmul_rn(qj.w, __fmul_rn(Gf, depth_continuous_f(__fsqrt_rn(r2), appf)));
    }
    const double dx = (double)(qj.x - qi.x), dy = (double)(qj.y - qi.y), dz = (double)(qj.z - qi.z);
    const double r = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)));
    return __dmul_rn((double)qj.w, __dmul_rn(Gd, depth_continuous(r, app)));
}

// the add-back term of later passes, (w_j (a/6pi)) D(r) in the reference's
// association (bound_fortran.f90:2636-2645)
template <bool F32>
__device__ __forceinline__ double addback_term(const float4& qi, const float4& qj, double Gd, float Gf,
                                               double app, float appf) {
    if (F32) return pair_term<true>(qi, qj, Gd, Gf, app, appf);
    const double dx = (double)(qj.x - qi.x), dy = (double)(qj.y - qi.y), dz = (double)(qj.z - qi.z);
    const double r = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)));
    return __dmul_rn(__dmul_rn((double)qj.w, Gd), depth_continuous(r, app));
}

// depth_continuous(r, h) for a fixed h with the divisions by h taken out of
// the pair loop: 2r/h, x/(70h) and 12/e = 6h/r become products, and the
// caller passes 1/r once for both kernels of a C2 difference. The results
// differ from depth_continuous by f64 rounding only (relative ~1e-16); the
// pair terms are rounded to f32 afterwards.
struct DepthK {
    double h, two_over_h, inv_70h, six_h;
    __device__ __forceinline__ void init(double hh) {
        h = hh;
        two_over_h = 2.0 / hh;
        inv_70h = 1.0 / (70.0 * hh);
        six_h = 6.0 * hh;
    }
    // r >= 0; inv_r = 1/r (unused when e <= 1, so r = 0 is fine)
    __device__ __forceinline__ double eval(double r, double inv_r) const {
        if (r >= h) return inv_r;
        const double e = r * two_over_h;
        if (e <= 1.0) {
            const double prim = e * e * (112.0 + e * e * (-56.0 + e * (14.0 + e * (8.0 - 3.0 * e))));
            return (208.0 - prim) * inv_70h;
        }
        const double prim = e * e * (448.0 + e * (-280.0 + e * (56.0 + e * (14.0 + e * (-8.0 + e))))) -
                            224.0 * e - six_h * inv_r;
        return (128.0 - prim) * inv_70h;
    }
};

// C2 residual depth of a pair at distance d (in grid units: rs = d / sp),
// cutoff kernel ht = app / sp and outer kernel hs = 3.5
__device__ __forceinline__ double c2_delta_fast(double d, double inv_sp, double sp, const DepthK& Kt,
                                                const DepthK& Ks) {
    const double rs = d * inv_sp;
    if (rs >= Kt.h && rs >= Ks.h) return 0.0;
    const double inv_rs = sp / d;
    return Kt.eval(rs, inv_rs) - Ks.eval(rs, inv_rs);
}

__device__ __forceinline__ double gm_rel(double x, double xa, double L, double Lh) {
    double a = (x - xa) + Lh;
    double r = fmod(a, L);
    if (r != 0.0 && ((r < 0.0) != (L < 0.0))) r += L;
    return r - Lh;
}

// ---------------------------------------------------------------------------
// packing and span

// member slot t: float position relative to the halo's anchor (the reference's
// xv = real(dx_rel, 4), and q = xv - xv(anchor) = xv), float velocity, mass
__global__ void k_bound_pack(const uint32_t* __restrict__ members, const uint32_t* __restrict__ mhalo,
                             const uint32_t* __restrict__ anchor, const double* __restrict__ xp,
                             const double* __restrict__ vp, const double* __restrict__ mp,
                             uint32_t nmem, double L, double Lh, float4* __restrict__ qw,
                             float4* __restrict__ vv) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const size_t i = members[t];
    const size_t a = anchor[mhalo[t]];
    float4 q;
    q.x = (float)gm_rel(xp[3 * i], xp[3 * a], L, Lh);
    q.y = (float)gm_rel(xp[3 * i + 1], xp[3 * a + 1], L, Lh);
    q.z = (float)gm_rel(xp[3 * i + 2], xp[3 * a + 2], L, Lh);
    q.w = mp ? (float)mp[i] : 1.0f;
    qw[t] = q;
    vv[t] = vp ? make_float4((float)vp[3 * i], (float)vp[3 * i + 1], (float)vp[3 * i + 2], 0.f)
               : make_float4(0.f, 0.f, 0.f, 0.f);
}

// per halo: lo/hi start at 0 (the anchor) as bound_classify's float min/max
__global__ void __launch_bounds__(kB) k_bound_span(const float4* __restrict__ qw,
                                                   const uint32_t* __restrict__ hoff,
                                                   float* __restrict__ lohi) {
    __shared__ float red[6][kB / kWave];
    const uint32_t h = blockIdx.x;
    const uint32_t s = h ? hoff[h - 1] : 0u, e = hoff[h];
    float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
    for (uint32_t t = s + threadIdx.x; t < e; t += kB) {
        const float4 q = qw[t];
        lo[0] = fminf(lo[0], q.x); hi[0] = fmaxf(hi[0], q.x);
        lo[1] = fminf(lo[1], q.y); hi[1] = fmaxf(hi[1], q.y);
        lo[2] = fminf(lo[2], q.z); hi[2] = fmaxf(hi[2], q.z);
    }
    for (int d = 0; d < 3; ++d)
        for (int o = 32; o > 0; o >>= 1) {
            lo[d] = fminf(lo[d], __shfl_down(lo[d], o, kWave));
            hi[d] = fmaxf(hi[d], __shfl_down(hi[d], o, kWave));
        }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int d = 0; d < 3; ++d) { red[d][w] = lo[d]; red[3 + d][w] = hi[d]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        float v = red[d][0];
        for (int k = 1; k < kB / kWave; ++k) v = d < 3 ? fminf(v, red[d][k]) : fmaxf(v, red[d][k]);
        lohi[6 * (size_t)h + d] = v;
    }
}

// ---------------------------------------------------------------------------
// direct summation halos: one workgroup each, everything in LDS

struct DirectLds {
    float4 q[kDirectMax];     // position, mass
    float3 v[kDirectMax];
    double phid[kDirectMax];
    float phi[kDirectMax];
    float en[kDirectMax];
    int16_t just[kDirectMax];
    int16_t ub[kDirectMax];
    uint8_t keep[kDirectMax];
    double red[8];
    int32_t cnt[4];
};

__device__ __forceinline__ bool e_greater(float eb, int b, float ea, int a) {
    // bound_fortran.f90:2808-2844: larger energy, ties -> smaller local index
    return eb > ea || (eb == ea && b < a);
}

// F32: pair terms and potential sums in f32 (the mixed-precision form);
// otherwise FP64 exactly as the reference. phi_rec (optional) receives the
// first pass's potential.
template <bool F32>
__global__ void __launch_bounds__(kB) k_direct_solve(const int32_t* __restrict__ dlist,
                                                     const uint32_t* __restrict__ hoff,
                                                     const float4* __restrict__ qw,
                                                     const float4* __restrict__ vv,
                                                     uint8_t* __restrict__ keep_out,
                                                     int32_t* __restrict__ status,
                                                     int32_t* __restrict__ npb_out,
                                                     int32_t* __restrict__ err,
                                                     float* __restrict__ phi_rec, Params P) {
    __shared__ DirectLds S;
    const int32_t h = dlist[blockIdx.x];
    const uint32_t s0 = h ? hoff[h - 1] : 0u;
    const int n = (int)(hoff[h] - s0);
    const int tid = threadIdx.x;
    for (int i = tid; i < n; i += kB) {
        S.q[i] = qw[s0 + i];
        const float4 v = vv[s0 + i];
        S.v[i] = make_float3(v.x, v.y, v.z);
        S.keep[i] = 1;
    }
    if (tid == 0) { S.cnt[0] = 0; S.cnt[1] = 0; }
    __syncthreads();
    const double Gd = P.a / (6.0 * kPi);
    const float Gf = (float)Gd, appf = (float)P.app;
    const float b32 = (float)P.b;
    int np_bound = n, n_just = 0, st = 5;
    for (int pass = 1; pass <= 20; ++pass) {
        if (pass == 1) {
            // phi_direct(k) = sum over partners j (ascending) of -w_j (a/6pi) D(r)
            for (int k = tid; k < n; k += kB) {
                const float4 qk = S.q[k];
                double acc = 0.0;
                float accf = 0.0f;
                for (int j = 0; j < n; ++j) {
                    if (j == k) continue;
                    const float4 qj = S.q[j];
                    const double t = pair_term<F32>(qk, qj, Gd, Gf, P.app, appf);
                    if (F32) {
                        accf = __fsub_rn(accf, (float)t);
                        if ((j & 255) == 255) {  // f32 within 256 partners, f64 across
                            acc = __dadd_rn(acc, (double)accf);
                            accf = 0.0f;
                        }
                    } else {
                        acc = __dsub_rn(acc, t);
                    }
                }
                S.phid[k] = F32 ? __dadd_rn(acc, (double)accf) : acc;
            }
        } else {
            // add back the sources deleted since the last refresh, in deletion order
            for (int i = tid; i < n; i += kB) {
                if (!S.keep[i]) continue;
                const float4 qi = S.q[i];
                double acc = S.phid[i];
                for (int k = 0; k < n_just; ++k) {
                    const int j = S.just[k];
                    const float4 qj = S.q[j];
                    if (F32) {  // f32 term, f64 sum
                        acc = __dadd_rn(acc, (double)(float)pair_term<true>(qi, qj, Gd, Gf, P.app, appf));
                    } else {
                        const double dx = (double)(qj.x - qi.x), dy = (double)(qj.y - qi.y),
                                     dz = (double)(qj.z - qi.z);
                        const double r = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)));
                        acc = __dadd_rn(acc, __dmul_rn(__dmul_rn((double)qj.w, Gd), depth_continuous(r, P.app)));
                    }
                }
                S.phid[i] = acc;
            }
        }
        n_just = 0;
        __syncthreads();
        bool bad = false;
        for (int i = tid; i < n; i += kB) {
            if (!S.keep[i]) continue;
            float p = (float)S.phid[i];
            if (np_bound == 1) p = 0.0f;
            if (!isfinite(p)) bad = true;
            S.phi[i] = p;
            if (pass == 1 && phi_rec) phi_rec[s0 + i] = p;
        }
        if (__syncthreads_or(bad)) { st = -4; break; }
        int deleted = 0;
        bool refresh = false;
        while (!refresh) {
            if (tid == 0) {
                // f64 centre and bulk velocity over kept members, in index order
                double ws = 0.0, qc[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
                for (int i = 0; i < n; ++i) {
                    if (!S.keep[i]) continue;
                    const double wi = (double)S.q[i].w;
                    ws = __dadd_rn(ws, wi);
                    qc[0] = __dadd_rn(qc[0], __dmul_rn(wi, (double)S.q[i].x));
                    qc[1] = __dadd_rn(qc[1], __dmul_rn(wi, (double)S.q[i].y));
                    qc[2] = __dadd_rn(qc[2], __dmul_rn(wi, (double)S.q[i].z));
                    vc[0] = __dadd_rn(vc[0], __dmul_rn(wi, (double)S.v[i].x));
                    vc[1] = __dadd_rn(vc[1], __dmul_rn(wi, (double)S.v[i].y));
                    vc[2] = __dadd_rn(vc[2], __dmul_rn(wi, (double)S.v[i].z));
                }
                for (int d = 0; d < 3; ++d) { S.red[d] = qc[d] / ws; S.red[3 + d] = vc[d] / ws; }
                S.cnt[0] = 0;
            }
            __syncthreads();
            const float qc32[3] = {(float)S.red[0], (float)S.red[1], (float)S.red[2]};
            const float vc32[3] = {(float)S.red[3], (float)S.red[4], (float)S.red[5]};
            bool badE = false;
            for (int i = tid; i < n; i += kB) {
                if (!S.keep[i]) continue;
                const float4 q = S.q[i];
                const float3 v = S.v[i];
                const float d0 = __fadd_rn(__fsub_rn(v.x, vc32[0]), __fmul_rn(b32, __fsub_rn(q.x, qc32[0])));
                const float d1 = __fadd_rn(__fsub_rn(v.y, vc32[1]), __fmul_rn(b32, __fsub_rn(q.y, qc32[1])));
                const float d2 = __fadd_rn(__fsub_rn(v.z, vc32[2]), __fmul_rn(b32, __fsub_rn(q.z, qc32[2])));
                float s = __fmul_rn(d0, d0);
                s = __fadd_rn(s, __fmul_rn(d1, d1));
                s = __fadd_rn(s, __fmul_rn(d2, d2));
                const float E = __fadd_rn(s / 2.0f, S.phi[i]);
                S.en[i] = E;
                if (!isfinite(E)) badE = true;
                if (E >= 0.0f) S.ub[atomicAdd(&S.cnt[0], 1)] = (int16_t)i;
            }
            if (__syncthreads_or(badE)) { st = -1; break; }
            const int nu = S.cnt[0];
            if (nu == 0) {
                if (deleted == 0) { st = 0; break; }
                refresh = true;
                __syncthreads();
                continue;
            }
            int nb = (int)ceil(0.02 * (double)np_bound);
            nb = nb < 1 ? 1 : nb;
            nb = nb > nu ? nu : nb;
            // rank of each unbound member in decreasing "greatness"; the nb first go
            for (int u = tid; u < nu; u += kB) {
                const int a = S.ub[u];
                const float ea = S.en[a];
                int rank = 0;
                for (int t = 0; t < nu; ++t) {
                    const int b = S.ub[t];
                    if (e_greater(S.en[b], b, ea, a)) ++rank;
                }
                if (rank < nb) S.just[n_just + rank] = (int16_t)a;
            }
            __syncthreads();
            for (int k = tid; k < nb; k += kB) S.keep[S.just[n_just + k]] = 0;
            n_just += nb;
            np_bound -= nb;
            deleted += nb;
            __syncthreads();
            if (np_bound < P.np_min) { st = 1; break; }
            if ((double)deleted >= 0.2 * (double)np_bound) refresh = true;
        }
        if (st != 5) break;
    }
    __syncthreads();
    if (st < 0) {
        if (tid == 0) atomicCAS(err, 0, st == -4 ? 4 : 1);
    }
    for (int i = tid; i < n; i += kB) keep_out[s0 + i] = S.keep[i];
    if (tid == 0) {
        status[h] = st < 0 ? 5 : st;
        npb_out[h] = np_bound;
    }
}

// ---------------------------------------------------------------------------
// direct summation of halos above the LDS size (up to the context's direct
// threshold): potentials in global memory, one workgroup per 256-member tile
// of a halo, partners staged in LDS in ascending order; the removal loop is
// k_pm_inner's. Pass 1 sums every partner; later passes add back the
// members deleted since the previous potential, as the reference's direct
// path does (bound_fortran.f90:2613-2647).

template <bool F32>
__global__ void __launch_bounds__(kB) k_dl_full(const int2* __restrict__ tiles,
                                                const uint32_t* __restrict__ hoff,
                                                const float4* __restrict__ qw,
                                                double* __restrict__ phid, Params P) {
    __shared__ float4 sq[kB];
    const int2 tl = tiles[blockIdx.x];
    const uint32_t s0 = tl.x ? hoff[tl.x - 1] : 0u;
    const int n = (int)(hoff[tl.x] - s0);
    const int i = tl.y + (int)threadIdx.x;
    const bool valid = i < n;
    const float4 qi = valid ? qw[s0 + i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const double Gd = P.a / (6.0 * kPi);
    const float Gf = (float)Gd, appf = (float)P.app;
    double acc = 0.0;
    float accf;
    for (int jb = 0; jb < n; jb += kB) {
        __syncthreads();
        if (jb + (int)threadIdx.x < n) sq[threadIdx.x] = qw[s0 + jb + threadIdx.x];
        __syncthreads();
        const int jn = min(kB, n - jb);
        if (!valid) continue;
        accf = 0.0f;
        for (int jj = 0; jj < jn; ++jj) {
            if (jb + jj == i) continue;
            const double t = pair_term<F32>(qi, sq[jj], Gd, Gf, P.app, appf);
            if (F32) accf = __fsub_rn(accf, (float)t);
            else acc = __dsub_rn(acc, t);
        }
        if (F32) acc = __dadd_rn(acc, (double)accf);  // f32 within a tile, f64 across tiles
    }
    if (valid) phid[s0 + i] = acc;
}

// add back the deleted partners dj[dstart[h] .. dstart[h+1]) to kept members
template <bool F32>
__global__ void __launch_bounds__(kB) k_dl_update(const int2* __restrict__ tiles,
                                                  const uint32_t* __restrict__ hoff,
                                                  const float4* __restrict__ qw,
                                                  const uint8_t* __restrict__ keep,
                                                  const uint32_t* __restrict__ dstart,
                                                  const uint32_t* __restrict__ dj,
                                                  double* __restrict__ phid, Params P) {
    __shared__ float4 sq[kB];
    const int2 tl = tiles[blockIdx.x];
    const uint32_t s0 = tl.x ? hoff[tl.x - 1] : 0u;
    const int n = (int)(hoff[tl.x] - s0);
    const int i = tl.y + (int)threadIdx.x;
    const bool valid = i < n && keep[s0 + i];
    const float4 qi = valid ? qw[s0 + i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const double Gd = P.a / (6.0 * kPi);
    const float Gf = (float)Gd, appf = (float)P.app;
    double acc = valid ? phid[s0 + i] : 0.0;
    const uint32_t d0 = dstart[tl.x], d1 = dstart[tl.x + 1];
    for (uint32_t jb = d0; jb < d1; jb += kB) {
        __syncthreads();
        if (jb + threadIdx.x < d1) sq[threadIdx.x] = qw[dj[jb + threadIdx.x]];
        __syncthreads();
        const int jn = (int)min((uint32_t)kB, d1 - jb);
        if (!valid) continue;
        float accf = 0.0f;
        for (int jj = 0; jj < jn; ++jj) {
            const double t = addback_term<F32>(qi, sq[jj], Gd, Gf, P.app, appf);
            if (F32) accf = __fadd_rn(accf, (float)t);
            else acc = __dadd_rn(acc, t);
        }
        if (F32) acc = __dadd_rn(acc, (double)accf);
    }
    if (valid) phid[s0 + i] = acc;
}

// members of active large-direct halos deleted since the last potential
__global__ void k_dl_del_flags(const uint32_t* __restrict__ mhalo, const int8_t* __restrict__ kind,
                               const uint8_t* __restrict__ active, const uint8_t* __restrict__ keep,
                               const uint8_t* __restrict__ keep_prev, uint32_t nmem,
                               uint32_t* __restrict__ flag) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const uint32_t h = mhalo[t];
    flag[t] = (kind[h] == 1 && active[h] && keep_prev[t] && !keep[t]) ? 1u : 0u;
}

__global__ void k_dl_del_compact(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ idx,
                                 const uint32_t* __restrict__ mhalo, uint32_t nmem,
                                 uint32_t* __restrict__ dkeys, uint32_t* __restrict__ dvals) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem || !flag[t]) return;
    dkeys[idx[t]] = mhalo[t];
    dvals[idx[t]] = t;
}

// phi of kept members of active large-direct halos (0 for a single member)
__global__ void k_dl_phi(const uint32_t* __restrict__ mhalo, const int8_t* __restrict__ kind,
                         const uint8_t* __restrict__ keep, const uint8_t* __restrict__ active,
                         const int32_t* __restrict__ npb, const double* __restrict__ phid, uint32_t nmem,
                         float* __restrict__ phi, int32_t* __restrict__ err) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem || !keep[t]) return;
    const uint32_t h = mhalo[t];
    if (kind[h] != 1 || !active[h]) return;
    float p = (float)phid[t];
    if (npb[h] == 1) p = 0.0f;
    if (!isfinite(p)) atomicCAS(err, 0, 4);
    phi[t] = p;
}

// first-pass potential of the members of active halos
__global__ void k_record_phi(const uint32_t* __restrict__ mhalo, const uint8_t* __restrict__ active,
                             const float* __restrict__ phi, uint32_t nmem, float* __restrict__ rec) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t < nmem && active[mhalo[t]]) rec[t] = phi[t];
}

// ---------------------------------------------------------------------------
// Green function (bound_fortran.f90:2300-2382), full spectrum, f64 alias sum

__device__ __forceinline__ double wave_k(int idx, int n) {
    int mode = idx < n / 2 ? idx : idx - n;
    if (idx == n / 2) mode = -n / 2;
    return 2.0 * kPi * (double)mode / (double)n;
}

// range 0: the outer grids' kernel (truncated at R = n/2); range > 0: the
// band kernel of the inner two-level grids, not truncated
// (bound_fortran.f90:2318-2347)
__global__ void k_green(float* __restrict__ G, int32_t n, double range, int32_t truncated) {
    const uint64_t id = (uint64_t)blockIdx.x * kB + threadIdx.x;
    const uint64_t tot = (uint64_t)n * n * n;
    if (id >= tot) return;
    const int x = (int)(id % n), y = (int)((id / n) % n), z = (int)(id / ((uint64_t)n * n));
    const double kv[3] = {wave_k(x, n), wave_k(y, n), wave_k(z, n)};
    const double radius = 0.5 * (double)n, soft = 3.5;
    // per-axis sinc^6 of the aliases
    double w6[3][5];
    for (int ax = 0; ax < 3; ++ax)
        for (int a = 0; a < 5; ++a) {
            const double k = kv[ax] + 2.0 * kPi * (double)(a - 2);
            const double hx = 0.5 * k;
            double s = 1.0;
            if (hx != 0.0) s = sin(hx) / hx;
            const double s2 = s * s;
            w6[ax][a] = s2 * s2 * s2;
        }
    double asum = 0.0, bsum = 0.0;
    for (int ka = 0; ka < 5; ++ka)
        for (int ja = 0; ja < 5; ++ja)
            for (int ia = 0; ia < 5; ++ia) {
                const double k0 = kv[0] + 2.0 * kPi * (double)(ia - 2);
                const double k1 = kv[1] + 2.0 * kPi * (double)(ja - 2);
                const double k2 = kv[2] + 2.0 * kPi * (double)(ka - 2);
                const double kmag = sqrt(k0 * k0 + k1 * k1 + k2 * k2);
                const double uk2 = w6[0][ia] * w6[1][ja] * w6[2][ka];
                asum += uk2;
                double target;
                if (kmag == 0.0) {
                    if (range == 0.0) target = -4.0 * kPi * radius * radius / 6.0;
                    else if (truncated) target = 0.0;
                    else target = -2.0 * kPi * (range * range - soft * soft) / 15.0;
                } else {
                    const double yv = soft * kmag / 2.0;
                    const double y4 = yv * yv * yv * yv;
                    const double ss = 12.0 * (1.0 / y4) * (2.0 - 2.0 * cos(yv) - yv * sin(yv));
                    double sr = 0.0;
                    if (range != 0.0) {
                        const double yr = range * kmag / 2.0;
                        const double yr4 = yr * yr * yr * yr;
                        sr = 12.0 * (1.0 / yr4) * (2.0 - 2.0 * cos(yr) - yr * sin(yr));
                    }
                    const double kr = kmag * radius;
                    const double cut = truncated ? 1.0 - sin(kr) / kr : 1.0;
                    target = -4.0 * kPi * (ss * ss - sr * sr) * cut / (kmag * kmag);
                }
                bsum += uk2 * target;
            }
    G[id] = (float)(bsum / (asum * asum));
}

__global__ void k_green_to_complex(const float* __restrict__ G, float2* __restrict__ g, uint64_t tot) {
    const uint64_t id = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (id < tot) g[id] = make_float2(G[id], 0.0f);
}

__global__ void k_self_extract(const float2* __restrict__ g, int32_t n, float* __restrict__ self) {
    const int t = threadIdx.x;
    if (t >= 125) return;
    const int ix = t % 5 - 2, iy = (t / 5) % 5 - 2, iz = t / 25 - 2;
    const int x = (ix + n) % n, y = (iy + n) % n, z = (iz + n) % n;
    const float norm = (float)((double)n * n * n);
    self[t] = g[(size_t)x + (size_t)n * ((size_t)y + (size_t)n * z)].x / norm;
}

// ---------------------------------------------------------------------------
// particle-mesh halos

struct PmHalo {       // per PM halo
    float4 origin;    // (x, y, z, spacing)
    int32_t slot;     // grid slot within its class
    int32_t cls;      // grid class 0..3
    int32_t n;        // grid size
    int32_t c2off;    // first C2 cell of the halo (global)
    int32_t nc[3];    // C2 cells per axis
    int32_t tl;       // 1: two-level (per-member short-range cutoff, inner grids)
    double umin[3];   // C2 cell grid origin in u = q / spacing
};

// TSC setup (bound_fortran.f90:2560-2602): nearest node, weights, self term
__global__ void k_tsc_setup(const float4* __restrict__ qw, const uint32_t* __restrict__ mhalo,
                            const int32_t* __restrict__ pm_of_halo, const PmHalo* __restrict__ pm,
                            const float* __restrict__ selfs, uint32_t nmem,
                            int4* __restrict__ first, float* __restrict__ wt,
                            float* __restrict__ self, int32_t* __restrict__ err) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const int32_t k = pm_of_halo[mhalo[t]];
    if (k < 0) return;
    const PmHalo H = pm[k];
    const float4 q = qw[t];
    const float sp = H.origin.w;
    const float qa[3] = {q.x, q.y, q.z}, oa[3] = {H.origin.x, H.origin.y, H.origin.z};
    float w9[9];
    int f[3];
    for (int ax = 0; ax < 3; ++ax) {
        const float tt = __fdiv_rn(__fsub_rn(qa[ax], oa[ax]), sp);
        if (!(tt >= 0.0f) || !(tt < (float)(H.n / 2 - 1))) {
            atomicCAS(err, 0, 1);  // bound_fortran.f90:2577-2580
            return;
        }
        const int nearest = (int)roundf(tt);  // Fortran NINT: half away from zero
        const float d = __fsub_rn(tt, (float)nearest);
        f[ax] = nearest;
        const float a0 = __fsub_rn(0.5f, d), a2 = __fadd_rn(0.5f, d);
        w9[3 * ax + 0] = __fmul_rn(0.5f, __fmul_rn(a0, a0));
        w9[3 * ax + 1] = __fsub_rn(0.75f, __fmul_rn(d, d));
        w9[3 * ax + 2] = __fmul_rn(0.5f, __fmul_rn(a2, a2));
    }
    const float* sk = selfs + 125 * H.cls;
    double acc = 0.0;
    for (int sz = 0; sz < 3; ++sz)
        for (int sy = 0; sy < 3; ++sy)
            for (int sx = 0; sx < 3; ++sx) {
                const float sw = __fmul_rn(__fmul_rn(w9[sx], w9[3 + sy]), w9[6 + sz]);
                for (int tz = 0; tz < 3; ++tz)
                    for (int ty = 0; ty < 3; ++ty)
                        for (int tx = 0; tx < 3; ++tx) {
                            const float tw = __fmul_rn(__fmul_rn(w9[tx], w9[3 + ty]), w9[6 + tz]);
                            acc = __dadd_rn(acc, __dmul_rn((double)__fmul_rn(sw, tw),
                                                           (double)sk[(tx - sx + 2) + 5 * ((ty - sy + 2) + 5 * (tz - sz + 2))]));
                        }
            }
    first[t] = make_int4(f[0], f[1], f[2], 0);
    for (int i = 0; i < 9; ++i) wt[9 * (size_t)t + i] = w9[i];
    self[t] = (float)acc;
}

// deposit of kept members into int64 fixed-point grids (order independent)
// TSC deposit of the kept members of active halos of one grid class
// (bound_fortran.f90:2660-2668) as int64 fixed point, so the sums do not
// depend on the order of the additions. Members come in C2-cell order (order
// = the class's range of the C2 member list), so the 256 members of a
// workgroup are neighbours: when they belong to one halo and their stencils
// fit in a box of BOX nodes, they are summed into that box in LDS first
// and the box is added to the grid once per node; otherwise each member adds
// its 27 nodes to the grid directly.
// workgroup size and LDS box (nodes) of the two deposits: the two-level
// sources are sparser per node block, so a 256-entry run spans more nodes
// than a 4096-node box; runs of 64 fit a 1024-node box far more often
// (measured: C3 unbinding 113.5 -> 109.9 ms per call)
constexpr int kDepTB = 256, kDepBox = 4096;
constexpr int kTlDepTB = 64, kTlDepBox = 1024;

template <int TB, int BOX>
__global__ void __launch_bounds__(TB) k_deposit(const float4* __restrict__ qw, const uint32_t* __restrict__ mhalo,
                                                const int32_t* __restrict__ pm_of_halo, const PmHalo* __restrict__ pm,
                                                const uint8_t* __restrict__ keep, const int4* __restrict__ first,
                                                const float* __restrict__ wt, const uint8_t* __restrict__ active,
                                                const double* __restrict__ fscale, int32_t cls,
                                                const uint32_t* __restrict__ order, uint32_t count,
                                                unsigned long long* __restrict__ acc) {
    __shared__ unsigned long long box[BOX];
    __shared__ int32_t blo[3], bhi[3], bslot, bn, mixed;
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; ++d) { blo[d] = 0x7fffffff; bhi[d] = -0x7fffffff; }
        bslot = -1;
        bn = 0;
        mixed = 0;
    }
    __syncthreads();
    const uint32_t ps = blockIdx.x * TB + threadIdx.x;
    bool dep = false;
    uint32_t t = 0, h = 0;
    int32_t slot = -1, n = 0;
    int4 f = make_int4(0, 0, 0, 0);
    if (ps < count) {
        t = order[ps];
        if (keep[t]) {
            h = mhalo[t];
            const int32_t k = pm_of_halo[h];
            if (k >= 0 && pm[k].cls == cls && active[h]) {
                dep = true;
                slot = pm[k].slot;
                n = pm[k].n;
                f = first[t];
            }
        }
    }
    if (dep) {
        atomicMin(&blo[0], f.x); atomicMin(&blo[1], f.y); atomicMin(&blo[2], f.z);
        atomicMax(&bhi[0], f.x + 2); atomicMax(&bhi[1], f.y + 2); atomicMax(&bhi[2], f.z + 2);
        if (atomicCAS(&bslot, -1, slot) == -1) bn = n;
    }
    __syncthreads();
    if (dep && slot != bslot) mixed = 1;
    __syncthreads();
    // stencil nodes nearest..nearest+2 per axis; k_tsc_setup's guard t in
    // [0, n/2 - 1) keeps them inside [0, n/2 + 1]
    const float w = dep ? qw[t].w : 0.0f;
    const float* w9 = wt + 9 * (size_t)t;
    const double fs = dep ? fscale[h] : 0.0;
    const int ex = bhi[0] - blo[0] + 1, ey = bhi[1] - blo[1] + 1, ez = bhi[2] - blo[2] + 1;
    const bool in_lds = bslot >= 0 && !mixed && (int64_t)ex * ey * ez <= BOX;
    if (!in_lds) {
        if (!dep) return;
        unsigned long long* g = acc + (size_t)slot * n * n * n;
        for (int sz = 0; sz < 3; ++sz)
            for (int sy = 0; sy < 3; ++sy)
                for (int sx = 0; sx < 3; ++sx) {
                    const float v = __fmul_rn(__fmul_rn(__fmul_rn(w, w9[sx]), w9[3 + sy]), w9[6 + sz]);
                    const long long fx = llrint((double)v * fs);
                    const size_t node = (size_t)(f.x + sx) + (size_t)n * ((size_t)(f.y + sy) + (size_t)n * (f.z + sz));
                    atomicAdd(&g[node], (unsigned long long)fx);
                }
        return;
    }
    const int vol = ex * ey * ez;
    for (int i = threadIdx.x; i < vol; i += TB) box[i] = 0ull;
    __syncthreads();
    if (dep) {
        for (int sz = 0; sz < 3; ++sz)
            for (int sy = 0; sy < 3; ++sy)
                for (int sx = 0; sx < 3; ++sx) {
                    const float v = __fmul_rn(__fmul_rn(__fmul_rn(w, w9[sx]), w9[3 + sy]), w9[6 + sz]);
                    const long long fx = llrint((double)v * fs);
                    const int i = (f.x + sx - blo[0]) + ex * ((f.y + sy - blo[1]) + ey * (f.z + sz - blo[2]));
                    atomicAdd(&box[i], (unsigned long long)fx);
                }
    }
    __syncthreads();
    const int nn = bn;
    unsigned long long* g = acc + (size_t)bslot * nn * nn * nn;
    for (int i = threadIdx.x; i < vol; i += TB) {
        const unsigned long long v = box[i];
        if (v == 0ull) continue;  // no member reached this node (all terms are >= 0)
        const int x = blo[0] + i % ex, y = blo[1] + (i / ex) % ey, z = blo[2] + i / (ex * ey);
        atomicAdd(&g[(size_t)x + (size_t)nn * ((size_t)y + (size_t)nn * z)], v);
    }
}

// zero the deposit nodes [0, mc)^3 of each listed slot
__global__ void k_zero_region(unsigned long long* __restrict__ acc, const int32_t* __restrict__ slots,
                              int32_t n, uint32_t mc) {
    const uint64_t id = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (id >= (uint64_t)mc * mc * mc) return;
    const uint32_t x = (uint32_t)(id % mc), y = (uint32_t)((id / mc) % mc), z = (uint32_t)(id / ((uint64_t)mc * mc));
    const uint64_t grid = (uint64_t)n * n * n;
    acc[(size_t)slots[blockIdx.y] * grid + x + (uint64_t)n * (y + (uint64_t)n * z)] = 0ull;
}

// C1 potential (bound_fortran.f90:2680-2703)
__global__ void k_interp(const float4* __restrict__ qw, const uint32_t* __restrict__ mhalo,
                         const int32_t* __restrict__ pm_of_halo, const PmHalo* __restrict__ pm,
                         const uint8_t* __restrict__ keep, const int4* __restrict__ first,
                         const float* __restrict__ wt, const float* __restrict__ self,
                         const uint8_t* __restrict__ active, const double* __restrict__ mkeep,
                         const float2* __restrict__ g, int32_t cls, uint32_t nmem, double a,
                         float* __restrict__ phi_c1) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem || !keep[t]) return;
    const uint32_t h = mhalo[t];
    const int32_t k = pm_of_halo[h];
    if (k < 0 || pm[k].cls != cls || !active[h]) return;
    const PmHalo H = pm[k];
    const int n = H.n;
    const float norm = (float)((double)n * n * n);
    const float2* gg = g + (size_t)H.slot * n * n * n;
    const int4 f = first[t];
    const float* w9 = wt + 9 * (size_t)t;
    float raw = 0.0f;
    for (int sz = 0; sz < 3; ++sz) {
        const int iz = f.z + sz;
        for (int sy = 0; sy < 3; ++sy) {
            const int iy = f.y + sy;
            for (int sx = 0; sx < 3; ++sx) {
                const int ix = f.x + sx;
                const float rho = __fdiv_rn(gg[(size_t)ix + (size_t)n * ((size_t)iy + (size_t)n * iz)].x, norm);
                raw = __fadd_rn(raw, __fmul_rn(__fmul_rn(__fmul_rn(rho, w9[sx]), w9[3 + sy]), w9[6 + sz]));
            }
        }
    }
    const float w = qw[t].w;
    const float radius = (float)n / 2.0f;
    const float dz0 = __fdiv_rn((float)fmax(0.0, mkeep[h] - (double)w), radius);
    const float pref = (float)(a / (6.0 * kPi * (double)H.origin.w));
    phi_c1[t] = __fmul_rn(pref, __fsub_rn(__fsub_rn(raw, __fmul_rn(w, self[t])), dz0));
}

// ---------------------------------------------------------------------------
// two-level refinement of 256^3 halos (bound_fortran.f90:403-1517)

constexpr int kTlRegions = 512;   // 8^3 blocks of 16^3 coarse cells
constexpr int kTlN12 = 3926, kTlN24 = 14688;
constexpr int kTlN2 = 128, kTlN4 = 216;

struct TlSlot {        // one refined block (inner grid)
    float lo[3], hi[3];  // block bounds in coarse-cell units
    float bo;            // band half-width 3.5 + 3 / r
    float u0[3];         // inner grid origin in coarse-cell units
    int32_t r, n_in, halo, cls_slot;  // ratio, inner grid, halo, slot within its ratio class
    float sp;            // the halo's spacing
};

// coarse-cell coordinate of a member, as bound_fortran.f90:567 computes it
__device__ __forceinline__ void tl_coord(const float4& q, const PmHalo& H, float t[3]) {
    const double sp = (double)H.origin.w;
    t[0] = (float)(((double)q.x - (double)H.origin.x) / sp);
    t[1] = (float)(((double)q.y - (double)H.origin.y) / sp);
    t[2] = (float)(((double)q.z - (double)H.origin.z) / sp);
}

__device__ __forceinline__ int tl_block(float t) {
    const int cell = (int)floor((double)t);
    const int v = cell / 16;  // truncation toward zero, as Fortran
    return v < 0 ? 0 : (v > 7 ? 7 : v);
}

// block of every member of a two-level candidate, and member counts per block
__global__ void k_tl_count(const float4* __restrict__ qw, const uint32_t* __restrict__ mhalo,
                           const int32_t* __restrict__ tl_of_halo, const int32_t* __restrict__ pm_of_halo,
                           const PmHalo* __restrict__ pm, uint32_t nmem, int16_t* __restrict__ treg,
                           uint32_t* __restrict__ counts) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const uint32_t h = mhalo[t];
    const int32_t k = tl_of_halo[h];
    if (k < 0) { treg[t] = -1; return; }
    float c[3];
    tl_coord(qw[t], pm[pm_of_halo[h]], c);
    const int b = tl_block(c[0]) + 8 * tl_block(c[1]) + 64 * tl_block(c[2]);
    treg[t] = (int16_t)b;
    atomicAdd(&counts[(size_t)k * kTlRegions + b], 1u);
}

// per-member ratio (1 outside two-level candidates and unrefined blocks)
__global__ void k_tl_ratio(const uint32_t* __restrict__ mhalo, const int32_t* __restrict__ tl_of_halo,
                           const int16_t* __restrict__ treg, const int8_t* __restrict__ region_r,
                           uint32_t nmem, uint8_t* __restrict__ pr) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const int32_t k = tl_of_halo[mhalo[t]];
    pr[t] = (uint8_t)(k < 0 ? 1 : region_r[(size_t)k * kTlRegions + treg[t]]);
}

__device__ __forceinline__ bool tl_in_band(const float c[3], const TlSlot& S) {
    return c[0] >= S.lo[0] - S.bo && c[0] <= S.hi[0] + S.bo && c[1] >= S.lo[1] - S.bo &&
           c[1] <= S.hi[1] + S.bo && c[2] >= S.lo[2] - S.bo && c[2] <= S.hi[2] + S.bo;
}

// inner-grid TSC of a coarse coordinate; false outside nodes [1, n_in/2]
__device__ __forceinline__ bool tl_tsc(const float c[3], const TlSlot& S, int f[3], float w9[9]) {
    bool ok = true;
    for (int ax = 0; ax < 3; ++ax) {
        const float coord = __fmul_rn(__fsub_rn(c[ax], S.u0[ax]), (float)S.r);
        const int nearest = (int)roundf(coord);
        const float d = __fsub_rn(coord, (float)nearest);
        f[ax] = nearest;
        const float a0 = __fsub_rn(0.5f, d), a2 = __fadd_rn(0.5f, d);
        w9[3 * ax + 0] = __fmul_rn(0.5f, __fmul_rn(a0, a0));
        w9[3 * ax + 1] = __fsub_rn(0.75f, __fmul_rn(d, d));
        w9[3 * ax + 2] = __fmul_rn(0.5f, __fmul_rn(a2, a2));
        if (nearest < 1 || nearest + 2 > S.n_in / 2) ok = false;
    }
    return ok;
}

// sources: every member of a candidate within the band of a refined block
// (its own or a neighbouring one). FILL = false counts, true writes entries.
template <bool FILL>
__global__ void k_tl_sources(const float4* __restrict__ qw, const uint32_t* __restrict__ mhalo,
                             const int32_t* __restrict__ tl_of_halo, const int32_t* __restrict__ pm_of_halo,
                             const PmHalo* __restrict__ pm, const int16_t* __restrict__ treg,
                             const int32_t* __restrict__ reg_slot, const TlSlot* __restrict__ slots,
                             uint32_t nmem, uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                             int32_t* __restrict__ src_slot, uint32_t* __restrict__ src_member,
                             int4* __restrict__ src_first, float* __restrict__ src_wt,
                             int32_t* __restrict__ err) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const uint32_t h = mhalo[t];
    const int32_t k = tl_of_halo[h];
    if (k < 0) {
        if (!FILL) cnt[t] = 0;
        return;
    }
    float c[3];
    tl_coord(qw[t], pm[pm_of_halo[h]], c);
    const int b = treg[t];
    const int bx = b % 8, by = (b / 8) % 8, bz = b / 64;
    uint32_t m = 0, o = FILL ? off[t] : 0u;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = bx + dx, ny = by + dy, nz = bz + dz;
                if (nx < 0 || nx > 7 || ny < 0 || ny > 7 || nz < 0 || nz > 7) continue;
                const int32_t sl = reg_slot[(size_t)k * kTlRegions + nx + 8 * ny + 64 * nz];
                if (sl < 0) continue;
                const TlSlot S = slots[sl];
                if (!tl_in_band(c, S)) continue;
                if (FILL) {
                    int f[3];
                    float w9[9];
                    if (!tl_tsc(c, S, f, w9)) atomicCAS(err, 0, 1);  // bound_fortran.f90:736-739
                    src_slot[o + m] = sl;
                    src_member[o + m] = t;
                    src_first[o + m] = make_int4(f[0], f[1], f[2], 0);
                    for (int i = 0; i < 9; ++i) src_wt[9 * (size_t)(o + m) + i] = w9[i];
                }
                ++m;
            }
    if (!FILL) cnt[t] = m;
}

// targets: members of refined blocks, their inner stencil and self term
__global__ void k_tl_targets(const float4* __restrict__ qw, const uint32_t* __restrict__ mhalo,
                             const int32_t* __restrict__ tl_of_halo, const int32_t* __restrict__ pm_of_halo,
                             const PmHalo* __restrict__ pm, const int16_t* __restrict__ treg,
                             const int32_t* __restrict__ reg_slot, const TlSlot* __restrict__ slots,
                             const float* __restrict__ self_r2, const float* __restrict__ self_r4,
                             uint32_t nmem, int32_t* __restrict__ tgt_slot, int4* __restrict__ tgt_first,
                             float* __restrict__ tgt_wt, float* __restrict__ self_in,
                             int32_t* __restrict__ err) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const uint32_t h = mhalo[t];
    const int32_t k = tl_of_halo[h];
    const int32_t sl = k < 0 ? -1 : reg_slot[(size_t)k * kTlRegions + treg[t]];
    tgt_slot[t] = sl;
    if (sl < 0) return;
    const TlSlot S = slots[sl];
    float c[3];
    tl_coord(qw[t], pm[pm_of_halo[h]], c);
    int f[3];
    float w9[9];
    if (!tl_tsc(c, S, f, w9)) {  // bound_fortran.f90:766-769
        atomicCAS(err, 0, 1);
        return;
    }
    tgt_first[t] = make_int4(f[0], f[1], f[2], 0);
    for (int i = 0; i < 9; ++i) tgt_wt[9 * (size_t)t + i] = w9[i];
    const float* sk = S.r == 2 ? self_r2 : self_r4;
    double acc = 0.0;
    for (int sz = 0; sz < 3; ++sz)
        for (int sy = 0; sy < 3; ++sy)
            for (int sx = 0; sx < 3; ++sx) {
                const float sw = __fmul_rn(__fmul_rn(w9[sx], w9[3 + sy]), w9[6 + sz]);
                for (int tz = 0; tz < 3; ++tz)
                    for (int ty = 0; ty < 3; ++ty)
                        for (int tx = 0; tx < 3; ++tx) {
                            const float tw = __fmul_rn(__fmul_rn(w9[tx], w9[3 + ty]), w9[6 + tz]);
                            acc = __dadd_rn(acc, __dmul_rn((double)__fmul_rn(sw, tw),
                                                           (double)sk[(tx - sx + 2) + 5 * ((ty - sy + 2) + 5 * (tz - sz + 2))]));
                        }
            }
    self_in[t] = (float)acc;
}

// fixed-point deposit of kept sources of the slots [s_lo, s_hi) of one ratio
// class into their inner grids (local slot = slot's class index - s_lo)
// sort keys of the source entries: slot rank (ratio-2 slots first, then
// ratio 4, each by slot), then the 4^3-node block of the stencil's first
// node
__global__ void k_tl_src_keys(const int32_t* __restrict__ src_slot, const int4* __restrict__ src_first,
                              const TlSlot* __restrict__ slots, uint32_t nsrc, int32_t nslot2,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t e = blockIdx.x * kB + threadIdx.x;
    if (e >= nsrc) return;
    const TlSlot S = slots[src_slot[e]];
    const uint32_t rank = (uint32_t)(S.r == 2 ? S.cls_slot : nslot2 + S.cls_slot);
    const int4 f = src_first[e];
    keys[e] = (rank << 18) | ((uint32_t)(f.z >> 2) & 63u) << 12 | ((uint32_t)(f.y >> 2) & 63u) << 6 |
              ((uint32_t)(f.x >> 2) & 63u);
    vals[e] = e;
}

// end of each slot rank's run in the sorted source keys (rank = key >> 18)
__global__ void k_rank_ends(const uint32_t* __restrict__ skeys, uint32_t n, uint32_t* __restrict__ ends) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = skeys[i] >> 18;
    if (i + 1 == n || (skeys[i + 1] >> 18) != r) ends[r] = i + 1;
}

// deposit of the band sources of one chunk of inner grids (bound_fortran.f90:
// 895-904), as int64 fixed point. With order (the entries sorted by slot and
// position) the launch covers the chunk's range only and a workgroup's
// entries are neighbours in one inner grid: their stencils are summed in an
// LDS box when it fits (kTlDepBox nodes), as in k_deposit.
template <int TB, int BOX>
__global__ void __launch_bounds__(TB) k_tl_deposit(const int32_t* __restrict__ src_slot,
                                                   const uint32_t* __restrict__ src_member,
                                                   const int4* __restrict__ src_first,
                                                   const float* __restrict__ src_wt,
                                                   const uint32_t* __restrict__ order, uint32_t begin,
                                                   uint32_t count, const TlSlot* __restrict__ slots,
                                                   int32_t ratio, int32_t s_lo, int32_t s_hi,
                                                   const uint8_t* __restrict__ keep,
                                                   const uint8_t* __restrict__ active,
                                                   const float4* __restrict__ qw,
                                                   const double* __restrict__ fscale,
                                                   unsigned long long* __restrict__ acc) {
    __shared__ unsigned long long box[BOX];
    __shared__ int32_t blo[3], bhi[3], bslot, bn, mixed;
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; ++d) { blo[d] = 0x7fffffff; bhi[d] = -0x7fffffff; }
        bslot = -1;
        bn = 0;
        mixed = 0;
    }
    __syncthreads();
    const uint32_t ps = blockIdx.x * TB + threadIdx.x;
    bool dep = false;
    uint32_t e = 0, t = 0;
    int32_t slot = -1, n = 0, halo = 0;
    int4 f = make_int4(0, 0, 0, 0);
    if (ps < count) {
        e = order ? order[begin + ps] : begin + ps;
        const TlSlot S = slots[src_slot[e]];
        if (S.r == ratio && S.cls_slot >= s_lo && S.cls_slot < s_hi && active[S.halo]) {
            t = src_member[e];
            if (keep[t]) {
                dep = true;
                slot = S.cls_slot - s_lo;
                n = S.n_in;
                halo = S.halo;
                f = src_first[e];
            }
        }
    }
    if (dep) {
        atomicMin(&blo[0], f.x); atomicMin(&blo[1], f.y); atomicMin(&blo[2], f.z);
        atomicMax(&bhi[0], f.x + 2); atomicMax(&bhi[1], f.y + 2); atomicMax(&bhi[2], f.z + 2);
        if (atomicCAS(&bslot, -1, slot) == -1) bn = n;
    }
    __syncthreads();
    if (dep && slot != bslot) mixed = 1;
    __syncthreads();
    const float w = dep ? qw[t].w : 0.0f;
    const float* w9 = src_wt + 9 * (size_t)e;
    const double fs = dep ? fscale[halo] : 0.0;
    const int ex = bhi[0] - blo[0] + 1, ey = bhi[1] - blo[1] + 1, ez = bhi[2] - blo[2] + 1;
    const bool in_lds = bslot >= 0 && !mixed && (int64_t)ex * ey * ez <= BOX;
    if (!in_lds) {
        if (!dep) return;
        unsigned long long* g = acc + (size_t)slot * n * n * n;
        for (int sz = 0; sz < 3; ++sz)
            for (int sy = 0; sy < 3; ++sy)
                for (int sx = 0; sx < 3; ++sx) {
                    const float v = __fmul_rn(__fmul_rn(__fmul_rn(w, w9[sx]), w9[3 + sy]), w9[6 + sz]);
                    const long long fx = llrint((double)v * fs);
                    atomicAdd(&g[(size_t)(f.x + sx) + (size_t)n * ((size_t)(f.y + sy) + (size_t)n * (f.z + sz))],
                              (unsigned long long)fx);
                }
        return;
    }
    const int vol = ex * ey * ez;
    for (int i = threadIdx.x; i < vol; i += TB) box[i] = 0ull;
    __syncthreads();
    if (dep) {
        for (int sz = 0; sz < 3; ++sz)
            for (int sy = 0; sy < 3; ++sy)
                for (int sx = 0; sx < 3; ++sx) {
                    const float v = __fmul_rn(__fmul_rn(__fmul_rn(w, w9[sx]), w9[3 + sy]), w9[6 + sz]);
                    const long long fx = llrint((double)v * fs);
                    atomicAdd(&box[(f.x + sx - blo[0]) + ex * ((f.y + sy - blo[1]) + ey * (f.z + sz - blo[2]))],
                              (unsigned long long)fx);
                }
    }
    __syncthreads();
    const int nn = bn;
    unsigned long long* g = acc + (size_t)bslot * nn * nn * nn;
    for (int i = threadIdx.x; i < vol; i += TB) {
        const unsigned long long v = box[i];
        if (v == 0ull) continue;
        const int x = blo[0] + i % ex, y = blo[1] + (i / ex) % ey, z = blo[2] + i / (ex * ey);
        atomicAdd(&g[(size_t)x + (size_t)nn * ((size_t)y + (size_t)nn * z)], v);
    }
}

// band potential of kept targets (bound_fortran.f90:906-938)
__global__ void k_tl_interp(const int32_t* __restrict__ tgt_slot, const int4* __restrict__ tgt_first,
                            const float* __restrict__ tgt_wt, const float* __restrict__ self_in,
                            const TlSlot* __restrict__ slots, int32_t ratio, int32_t s_lo, int32_t s_hi,
                            const uint8_t* __restrict__ keep, const uint8_t* __restrict__ active,
                            const float4* __restrict__ qw, const float2* __restrict__ grid, uint32_t nmem,
                            double a, float* __restrict__ phi_in) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem || !keep[t]) return;
    const int32_t sl = tgt_slot[t];
    if (sl < 0) return;
    const TlSlot S = slots[sl];
    if (S.r != ratio || S.cls_slot < s_lo || S.cls_slot >= s_hi || !active[S.halo]) return;
    const int n = S.n_in;
    const float norm = (float)((double)n * n * n);
    const float2* gg = grid + (size_t)(S.cls_slot - s_lo) * n * n * n;
    const int4 f = tgt_first[t];
    const float* w9 = tgt_wt + 9 * (size_t)t;
    float raw = 0.0f;
    for (int sz = 0; sz < 3; ++sz)
        for (int sy = 0; sy < 3; ++sy)
            for (int sx = 0; sx < 3; ++sx) {
                const float rho = __fdiv_rn(gg[(size_t)(f.x + sx) + (size_t)n * ((size_t)(f.y + sy) + (size_t)n * (f.z + sz))].x, norm);
                raw = __fadd_rn(raw, __fmul_rn(__fmul_rn(__fmul_rn(rho, w9[sx]), w9[3 + sy]), w9[6 + sz]));
            }
    const float pref = (float)(a * (double)S.r / (6.0 * kPi * (double)S.sp));
    phi_in[t] = __fmul_rn(pref, __fsub_rn(raw, __fmul_rn(qw[t].w, self_in[t])));
}

// short-range term of a two-level candidate for target i with cutoff a_i
// (bound_fortran.f90:1160-1175): f64 differences of the f32 positions, depth
// kernels in cube units, w_j * dphi in f32
__device__ __forceinline__ float tl_pair(const float4& qi, const float4& qj, const DepthK& Ke,
                                         const DepthK& Ka, double factor) {
    const double dx = (double)qj.x - (double)qi.x, dy = (double)qj.y - (double)qi.y,
                 dz = (double)qj.z - (double)qi.z;
    const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    if (!(r2 < Ka.h * Ka.h)) return 0.0f;  // Ka.h = a_i, the target's cutoff
    const double d = sqrt(r2);
    const double inv_d = 1.0 / d;
    const float dphi = -(float)(factor * (Ke.eval(d, inv_d) - Ka.eval(d, inv_d)));
    return __fmul_rn(qj.w, dphi);
}

// C2 cells (bound_fortran.f90:3125-3223): width 3.5 in u = q / spacing
__global__ void __launch_bounds__(kB) k_c2_extent(const float4* __restrict__ qw,
                                                  const uint32_t* __restrict__ hoff,
                                                  const int32_t* __restrict__ pm_list,
                                                  const int32_t* __restrict__ pm_halo,
                                                  const PmHalo* __restrict__ pm,
                                                  double* __restrict__ uext) {
    __shared__ double red[6][kB / kWave];
    const int32_t k = pm_list[blockIdx.x];
    const int32_t h = pm_halo[k];
    const uint32_t s = h ? hoff[h - 1] : 0u, e = hoff[h];
    const double sp = (double)pm[k].origin.w;
    double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
    for (uint32_t t = s + threadIdx.x; t < e; t += kB) {
        const float4 q = qw[t];
        const double u[3] = {(double)q.x / sp, (double)q.y / sp, (double)q.z / sp};
        for (int d = 0; d < 3; ++d) { lo[d] = fmin(lo[d], u[d]); hi[d] = fmax(hi[d], u[d]); }
    }
    for (int d = 0; d < 3; ++d)
        for (int o = 32; o > 0; o >>= 1) {
            lo[d] = fmin(lo[d], __shfl_down(lo[d], o, kWave));
            hi[d] = fmax(hi[d], __shfl_down(hi[d], o, kWave));
        }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int d = 0; d < 3; ++d) { red[d][w] = lo[d]; red[3 + d][w] = hi[d]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        double v = red[d][0];
        for (int q = 1; q < kB / kWave; ++q) v = d < 3 ? fmin(v, red[d][q]) : fmax(v, red[d][q]);
        uext[6 * (size_t)k + d] = v;
    }
}

__device__ __forceinline__ int c2_axis_cell(double u, double umin, int nc) {
    int v = (int)floor((u - umin) / 3.5) + 1;
    v = v < 1 ? 1 : v;
    return (v > nc ? nc : v) - 1;  // 0-based
}

// C2 cell of every member (the sentinel `ncells` for members of direct halos)
__global__ void k_c2_keys(const float4* __restrict__ qw, const uint32_t* __restrict__ mhalo,
                          const int32_t* __restrict__ pm_of_halo, const PmHalo* __restrict__ pm,
                          uint32_t nmem, uint32_t ncells, uint32_t* __restrict__ keys,
                          uint32_t* __restrict__ vals) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    vals[t] = t;
    const int32_t k = pm_of_halo[mhalo[t]];
    if (k < 0) {
        keys[t] = ncells;
        return;
    }
    const PmHalo H = pm[k];
    const float4 q = qw[t];
    const double sp = (double)H.origin.w;
    const int cx = c2_axis_cell((double)q.x / sp, H.umin[0], H.nc[0]);
    const int cy = c2_axis_cell((double)q.y / sp, H.umin[1], H.nc[1]);
    const int cz = c2_axis_cell((double)q.z / sp, H.umin[2], H.nc[2]);
    keys[t] = (uint32_t)H.c2off + (uint32_t)cx + (uint32_t)H.nc[0] * ((uint32_t)cy + (uint32_t)H.nc[1] * (uint32_t)cz);
}

// cell of each member, and run ends of the sorted cell keys
__global__ void k_c2_cells(const uint32_t* __restrict__ skeys, const uint32_t* __restrict__ svals,
                           uint32_t n, uint32_t ncells, uint32_t* __restrict__ cell_of,
                           uint32_t* __restrict__ ends) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = skeys[i];
    cell_of[svals[i]] = k;
    if (k < ncells && (i + 1 == n || skeys[i + 1] != k)) ends[k] = i + 1;
}

__global__ void k_fill32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = v;
}

// total mass of each PM halo (the kept mass of the first pass)
// and the fixed-point scale of its deposit: 2^40 units per unit mass, fewer
// when the whole halo mass would not stay below 2^61 units
__global__ void __launch_bounds__(kB) k_halo_mass(const float4* __restrict__ qw,
                                                  const uint32_t* __restrict__ hoff,
                                                  const int32_t* __restrict__ pm_halo,
                                                  double* __restrict__ mkeep,
                                                  double* __restrict__ fscale) {
    __shared__ double red[kB / kWave];
    const int32_t h = pm_halo[blockIdx.x];
    const uint32_t s = h ? hoff[h - 1] : 0u, e = hoff[h];
    double m = 0.0;
    for (uint32_t t = s + threadIdx.x; t < e; t += kB) m += (double)qw[t].w;
    for (int o = 32; o > 0; o >>= 1) m += __shfl_down(m, o, kWave);
    if (__lane_id() == 0) red[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm = 0.0;
        for (int j = 0; j < kB / kWave; ++j) sm += red[j];
        mkeep[h] = sm;
        int e = 60 - ilogb(fmax(sm, 1.0));
        fscale[h] = ldexp(1.0, e < 40 ? e : 40);
    }
}

// positions of the members in C2-cell order (qs[p] = qw[clist[p]]), mass 0
// for members no longer kept
__global__ void k_c2_sorted_pos(const float4* __restrict__ qw, const uint8_t* __restrict__ keep,
                                const uint32_t* __restrict__ clist, uint32_t n, float4* __restrict__ qs) {
    const uint32_t p = blockIdx.x * kB + threadIdx.x;
    if (p >= n) return;
    const uint32_t t = clist[p];
    float4 q = qw[t];
    if (!keep[t]) q.w = 0.0f;
    qs[p] = q;
}

// Full C2 residual over kept members (bound_fortran.f90:3403-3483), as a
// gather: member i sums over the kept members j of its 27 neighbour cells.
// Threads take members in cell order (clist), so the lanes of a wavefront
// share neighbour cells: equal trip counts and broadcast loads. Partners are
// read from qs, the positions in the same cell order: one contiguous load
// each; a partner of mass 0 (not kept) adds nothing and is skipped, which
// leaves the sum bit for bit as skipping the removed members does.
__global__ void k_c2_full(const float4* __restrict__ qw, const float4* __restrict__ qs,
                          const uint32_t* __restrict__ mhalo,
                          const int32_t* __restrict__ pm_of_halo, const PmHalo* __restrict__ pm,
                          const uint8_t* __restrict__ keep, const uint8_t* __restrict__ active,
                          const uint32_t* __restrict__ cell_of, const uint32_t* __restrict__ cstart,
                          const uint32_t* __restrict__ clist, uint32_t nmem, double a, double app,
                          const uint8_t* __restrict__ pr, float* __restrict__ phi_c2) {
    const uint32_t ps = blockIdx.x * kB + threadIdx.x;
    if (ps >= nmem) return;
    const uint32_t t = clist[ps];
    if (!keep[t]) return;
    const uint32_t h = mhalo[t];
    const int32_t k = pm_of_halo[h];
    if (k < 0 || !active[h]) return;
    const PmHalo H = pm[k];
    const float sp = H.origin.w;
    const double ht = (double)(float)app / (double)sp, hs = 3.5;
    const double r2cut = 12.25 * (double)sp * (double)sp;
    const float pref = (float)(a / (6.0 * kPi * (double)sp));
    const double ai = 3.5 * (double)sp / (double)(pr ? pr[t] : 1), eps = (double)(float)app,
                 factor = a / (6.0 * kPi);
    DepthK Kt, Ks, Ke, Ka;  // C2: ht and hs in grid units; two-level: eps and a_i
    Kt.init(ht);
    Ks.init(hs);
    Ke.init(eps);
    Ka.init(ai);
    const double spd = (double)sp, inv_sp = 1.0 / spd;
    // f32 screen: the f32 sum of squares is within ~4e-7 of r^2, so pairs
    // above cut^2 (1 + 1e-6) are outside the cutoff in the f64 test as well
    // (and add nothing); the rest take the f64 test unchanged
    const float cut32 = nextafterf((float)((H.tl ? ai * ai : r2cut) * (1.0 + 1e-6)), INFINITY);
    const uint32_t lc = cell_of[t] - (uint32_t)H.c2off;
    const int cx = (int)(lc % (uint32_t)H.nc[0]);
    const int cy = (int)((lc / (uint32_t)H.nc[0]) % (uint32_t)H.nc[1]);
    const int cz = (int)(lc / (uint32_t)(H.nc[0] * H.nc[1]));
    const float4 qi = qw[t];
    float acc = 0.0f;
    for (int z = cz - 1; z <= cz + 1; ++z) {
        if (z < 0 || z >= H.nc[2]) continue;
        for (int y = cy - 1; y <= cy + 1; ++y) {
            if (y < 0 || y >= H.nc[1]) continue;
            for (int x = cx - 1; x <= cx + 1; ++x) {
                if (x < 0 || x >= H.nc[0]) continue;
                const uint32_t c = (uint32_t)H.c2off + (uint32_t)x + (uint32_t)H.nc[0] * ((uint32_t)y + (uint32_t)H.nc[1] * (uint32_t)z);
                for (uint32_t p = cstart[c]; p < cstart[c + 1]; ++p) {
                    if (p == ps) continue;
                    const float4 qj = qs[p];
                    if (qj.w == 0.0f) continue;
                    {
                        const float ex = qj.x - qi.x, ey = qj.y - qi.y, ez = qj.z - qi.z;
                        if (ex * ex + ey * ey + ez * ez > cut32) continue;
                    }
                    if (H.tl) {  // two-level: the target's own cutoff 3.5 sp / r
                        acc = __fadd_rn(acc, tl_pair(qi, qj, Ke, Ka, factor));
                        continue;
                    }
                    const double dx = (double)(qj.x - qi.x), dy = (double)(qj.y - qi.y), dz = (double)(qj.z - qi.z);
                    const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
                    if (r2 < r2cut) {
                        const float dphi = __fmul_rn(-pref, (float)c2_delta_fast(sqrt(r2), inv_sp, spd, Kt, Ks));
                        acc = __fadd_rn(acc, __fmul_rn(qj.w, dphi));
                    }
                }
            }
        }
    }
    phi_c2[t] = acc;
}

// Incremental C2 (bound_fortran.f90:3490-3552): members deleted since the
// last refresh stop contributing; their residual is taken off every kept
// neighbour. Deleted members are listed per C2 cell and each kept member
// gathers from its 27 neighbour cells (fixed order, no atomics).
__global__ void k_c2_del_flags(const uint32_t* __restrict__ mhalo, const int32_t* __restrict__ pm_of_halo,
                               const uint8_t* __restrict__ active, const uint8_t* __restrict__ keep,
                               const uint8_t* __restrict__ keep_c2, uint32_t nmem,
                               uint32_t* __restrict__ flag) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const uint32_t h = mhalo[t];
    flag[t] = (pm_of_halo[h] >= 0 && active[h] && keep_c2[t] && !keep[t]) ? 1u : 0u;
}

__global__ void k_c2_del_compact(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ idx,
                                 const uint32_t* __restrict__ cell_of, uint32_t nmem,
                                 uint32_t* __restrict__ dkeys, uint32_t* __restrict__ dvals) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem || !flag[t]) return;
    dkeys[idx[t]] = cell_of[t];
    dvals[idx[t]] = t;
}

__global__ void k_run_ends(const uint32_t* __restrict__ skeys, uint32_t n, uint32_t* __restrict__ ends) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = skeys[i];
    if (i + 1 == n || skeys[i + 1] != k) ends[k] = i + 1;
}

__global__ void k_c2_update(const float4* __restrict__ qw, const uint32_t* __restrict__ mhalo,
                            const int32_t* __restrict__ pm_of_halo, const PmHalo* __restrict__ pm,
                            const uint8_t* __restrict__ keep, const uint8_t* __restrict__ active,
                            const uint32_t* __restrict__ cell_of, const uint32_t* __restrict__ dstart,
                            const uint32_t* __restrict__ dlist, const uint32_t* __restrict__ clist,
                            uint32_t nmem, double a, double app, const uint8_t* __restrict__ pr,
                            float* __restrict__ phi_c2) {
    const uint32_t ps = blockIdx.x * kB + threadIdx.x;
    if (ps >= nmem) return;
    const uint32_t t = clist[ps];
    if (!keep[t]) return;
    const uint32_t h = mhalo[t];
    const int32_t k = pm_of_halo[h];
    if (k < 0 || !active[h]) return;
    const PmHalo H = pm[k];
    const float sp = H.origin.w;
    const double ht = (double)(float)app / (double)sp, hs = 3.5;
    const double r2cut = 12.25 * (double)sp * (double)sp;
    const float pref = (float)(a / (6.0 * kPi * (double)sp));
    const double ai = 3.5 * (double)sp / (double)(pr ? pr[t] : 1), eps = (double)(float)app,
                 factor = a / (6.0 * kPi);
    DepthK Kt, Ks, Ke, Ka;  // C2: ht and hs in grid units; two-level: eps and a_i
    Kt.init(ht);
    Ks.init(hs);
    Ke.init(eps);
    Ka.init(ai);
    const double spd = (double)sp, inv_sp = 1.0 / spd;
    // f32 screen: the f32 sum of squares is within ~4e-7 of r^2, so pairs
    // above cut^2 (1 + 1e-6) are outside the cutoff in the f64 test as well
    // (and add nothing); the rest take the f64 test unchanged
    const float cut32 = nextafterf((float)((H.tl ? ai * ai : r2cut) * (1.0 + 1e-6)), INFINITY);
    const uint32_t lc = cell_of[t] - (uint32_t)H.c2off;
    const int cx = (int)(lc % (uint32_t)H.nc[0]);
    const int cy = (int)((lc / (uint32_t)H.nc[0]) % (uint32_t)H.nc[1]);
    const int cz = (int)(lc / (uint32_t)(H.nc[0] * H.nc[1]));
    const float4 qi = qw[t];
    float acc = phi_c2[t];
    for (int z = cz - 1; z <= cz + 1; ++z) {
        if (z < 0 || z >= H.nc[2]) continue;
        for (int y = cy - 1; y <= cy + 1; ++y) {
            if (y < 0 || y >= H.nc[1]) continue;
            for (int x = cx - 1; x <= cx + 1; ++x) {
                if (x < 0 || x >= H.nc[0]) continue;
                const uint32_t c = (uint32_t)H.c2off + (uint32_t)x + (uint32_t)H.nc[0] * ((uint32_t)y + (uint32_t)H.nc[1] * (uint32_t)z);
                for (uint32_t p = dstart[c]; p < dstart[c + 1]; ++p) {
                    const float4 qj = qw[dlist[p]];
                    {
                        const float ex = qj.x - qi.x, ey = qj.y - qi.y, ez = qj.z - qi.z;
                        if (ex * ex + ey * ey + ez * ez > cut32) continue;
                    }
                    if (H.tl) {
                        acc = __fsub_rn(acc, tl_pair(qi, qj, Ke, Ka, factor));
                        continue;
                    }
                    const double dx = (double)(qi.x - qj.x), dy = (double)(qi.y - qj.y), dz = (double)(qi.z - qj.z);
                    const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
                    if (r2 < r2cut) {
                        const float dphi = __fmul_rn(-pref, (float)c2_delta_fast(sqrt(r2), inv_sp, spd, Kt, Ks));
                        acc = __fsub_rn(acc, __fmul_rn(qj.w, dphi));
                    }
                }
            }
        }
    }
    phi_c2[t] = acc;
}

__global__ void k_copy_u8(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// phi = C1 + C2 (bound_fortran.f90:2739-2754); zero for a single member
__global__ void k_phi_combine(const uint32_t* __restrict__ mhalo, const int32_t* __restrict__ pm_of_halo,
                              const PmHalo* __restrict__ pm,
                              const uint8_t* __restrict__ keep, const uint8_t* __restrict__ active,
                              const int32_t* __restrict__ npb, const float* __restrict__ c1,
                              const float* __restrict__ c2, const float* __restrict__ pin, uint32_t nmem,
                              float* __restrict__ phi, int32_t* __restrict__ err) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem || !keep[t]) return;
    const uint32_t h = mhalo[t];
    const int32_t k = pm_of_halo[h];
    if (k < 0 || !active[h]) return;
    // two-level: C1 + band + C2 (bound_fortran.f90:2742-2744)
    float p = pm[k].tl ? __fadd_rn(__fadd_rn(c1[t], pin[t]), c2[t]) : __fadd_rn(c1[t], c2[t]);
    if (npb[h] == 1) p = 0.0f;
    if (!isfinite(p)) atomicCAS(err, 0, 4);
    phi[t] = p;
}

// Removal loop of one PM halo per workgroup (bound_fortran.f90:2756-2856):
// repeated 2 % batches by energy until no unbound member is left or 20 % of
// the kept members have gone since the potential was computed.
// Per batch: one pass for the f64 centre of mass and bulk velocity, one pass
// for the f32 energies that also lists the unbound members in index order
// (ul, with their energy keys uk); the nb-th largest key is then found by a
// radix select over that short list, and deletion walks it in index order
// (ties at the threshold: smaller index first).
constexpr int kPB = 1024;  // threads of k_pm_inner
constexpr int kPW = kPB / kWave;

__device__ __forceinline__ uint32_t ekey(float e) { return e > 0.0f ? __float_as_uint(e) : 0u; }

// exclusive rank of `flag` among the workgroup's threads in thread order,
// and the workgroup total; wcnt holds kPW counters in LDS
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t* wcnt, uint32_t& total) {
    const int w = threadIdx.x / kWave;
    const uint64_t bal = __ballot(flag);
    const uint32_t in_wave = __popcll(bal & lanemask_lt());
    if (__lane_id() == 0) wcnt[w] = __popcll(bal);
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int k = 0; k < kPW; ++k) {
        const uint32_t c = wcnt[k];
        before += k < w ? c : 0u;
        total += c;
    }
    __syncthreads();
    return before + in_wave;
}

// Large candidates (kBigMin members or more) run on several workgroups when
// the context may use cooperative launches (k_pm_big) and on one otherwise
// (k_pm_inner). Both form their f64 sums the same way, so both give the same
// result: the members split into big_blocks() contiguous chunks; each chunk
// is summed as a 256-thread workgroup sums it (thread-strided, then a wave
// shuffle tree, then the 4 waves in order); the chunk sums are added in
// chunk order.
constexpr int kBigMin = 32768;    // members: at least this many use k_pm_big
constexpr int kBigChunk = 16384;  // members per workgroup
constexpr int kBigMaxB = 64;      // workgroups per candidate at most
constexpr int kBW = kB / kWave;

__host__ __device__ __forceinline__ uint32_t big_blocks(uint32_t n) {
    const uint32_t b = (n + kBigChunk - 1) / kBigChunk;
    return b < (uint32_t)kBigMaxB ? b : (uint32_t)kBigMaxB;
}

// the chunked f64 sums of a large candidate inside one workgroup of kPB
// threads: NV values per member (7: mass, mass x position, mass x velocity;
// 1: mass), kept members only; out[k] for k < NV, valid after the call
template <int NV>
__device__ void chunked_sums(uint32_t s0, uint32_t e0, const uint8_t* __restrict__ keep,
                             const float4* __restrict__ qw, const float4* __restrict__ vv,
                             double (*part)[8], double (*wred)[kPB / kWave], double* out) {
    const uint32_t n = e0 - s0, nb = big_blocks(n), chunk = (n + nb - 1) / nb;
    const int tid = threadIdx.x, sub = tid / kB, lt = tid % kB, w = tid / kWave;
    constexpr int kSub = kPB / kB;
    for (uint32_t r = 0; r < (nb + kSub - 1) / kSub; ++r) {
        const uint32_t b = r * kSub + (uint32_t)sub;
        double v[NV];
        for (int k = 0; k < NV; ++k) v[k] = 0.0;
        if (b < nb) {
            const uint32_t c0 = s0 + min(n, b * chunk), c1 = s0 + min(n, (b + 1) * chunk);
            for (uint32_t t = c0 + lt; t < c1; t += kB) {
                if (!keep[t]) continue;
                const float4 q = qw[t];
                const double wi = (double)q.w;
                v[0] += wi;
                if (NV == 7) {
                    const float4 u = vv[t];
                    v[1] += wi * (double)q.x; v[2] += wi * (double)q.y; v[3] += wi * (double)q.z;
                    v[4] += wi * (double)u.x; v[5] += wi * (double)u.y; v[6] += wi * (double)u.z;
                }
            }
        }
        for (int k = 0; k < NV; ++k)
            for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o, kWave);
        if (__lane_id() == 0) for (int k = 0; k < NV; ++k) wred[k][w] = v[k];
        __syncthreads();
        if (b < nb && lt < NV) {
            double sacc = 0.0;
            for (int j = 0; j < kBW; ++j) sacc += wred[lt][sub * kBW + j];
            part[b][lt] = sacc;
        }
        __syncthreads();
    }
    if (tid < NV) {
        double sacc = 0.0;
        for (uint32_t j = 0; j < nb; ++j) sacc += part[j][tid];
        out[tid] = sacc;
    }
    __syncthreads();
}

// Radix-select step over a 256-bin digit histogram h (LDS), by the first 256
// threads of the workgroup (all threads must call it): the digit d holding
// the rem-th largest key counted from the top digit, and the rank left within
// it. This is the first d, from 255 down, whose count from the top reaches
// rem, found with a reversed prefix sum instead of a serial loop; as in that
// loop, d = 0 and rem minus the total if rem exceeds the total.
__device__ void select_digit256(const uint32_t* h, uint32_t rem, uint32_t* sel, uint32_t* wsum) {
    const int t = threadIdx.x;
    uint32_t c = 0, incl = 0;
    if (t < 256) {
        c = h[255 - t];
        incl = c;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, kWave);
            if ((int)__lane_id() >= o) incl += y;
        }
        if (__lane_id() == kWave - 1) wsum[t / kWave] = incl;
    }
    __syncthreads();
    if (t < 256) {
        uint32_t off = 0;
        for (int w = 0; w < t / kWave; ++w) off += wsum[w];
        incl += off;
        const uint32_t excl = incl - c;
        if (excl < rem && rem <= incl) { sel[0] = (uint32_t)(255 - t); sel[1] = rem - excl; }
        if (t == 255 && rem > incl) { sel[0] = 0u; sel[1] = rem - incl; }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kPB) k_pm_inner(const int32_t* __restrict__ alist,
                                                  const uint32_t* __restrict__ hoff,
                                                  const float4* __restrict__ qw,
                                                  const float4* __restrict__ vv,
                                                  const float* __restrict__ phi,
                                                  uint32_t* __restrict__ ul, uint32_t* __restrict__ uk,
                                                  uint8_t* __restrict__ keep,
                                                  int32_t* __restrict__ status,
                                                  int32_t* __restrict__ npb,
                                                  uint8_t* __restrict__ active,
                                                  double* __restrict__ mkeep, int32_t last_pass,
                                                  int32_t* __restrict__ err, Params P) {
    __shared__ double red[7][kPW];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t wcnt[kPW];
    __shared__ uint32_t sh[4];
    __shared__ double bc[7];
    __shared__ int bad_any;
    __shared__ double part[kBigMaxB][8];
    const int32_t h = alist[blockIdx.x];
    const uint32_t s0 = h ? hoff[h - 1] : 0u, e0 = hoff[h];
    const int tid = threadIdx.x, w = tid / kWave;
    const float b32 = (float)P.b;
    const bool large = e0 - s0 >= (uint32_t)kBigMin;
    int np_bound = npb[h];
    int deleted = 0, st = -2;  // -2: still iterating
    for (;;) {
        if (large) {
            chunked_sums<7>(s0, e0, keep, qw, vv, part, red, bc);
            if (tid == 0) bad_any = 0;
            __syncthreads();
        } else {
        // centre of mass and bulk velocity, f64, over kept members
        double v[7] = {0, 0, 0, 0, 0, 0, 0};
        for (uint32_t t = s0 + tid; t < e0; t += kPB) {
            if (!keep[t]) continue;
            const float4 q = qw[t];
            const float4 u = vv[t];
            const double wi = (double)q.w;
            v[0] += wi;
            v[1] += wi * (double)q.x; v[2] += wi * (double)q.y; v[3] += wi * (double)q.z;
            v[4] += wi * (double)u.x; v[5] += wi * (double)u.y; v[6] += wi * (double)u.z;
        }
        for (int k = 0; k < 7; ++k)
            for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o, kWave);
        if (__lane_id() == 0) for (int k = 0; k < 7; ++k) red[k][w] = v[k];
        if (tid == 0) bad_any = 0;
        __syncthreads();
        if (tid < 7) {
            double sacc = 0.0;
            for (int j = 0; j < kPW; ++j) sacc += red[tid][j];
            bc[tid] = sacc;
        }
        __syncthreads();
        }
        const double ws = bc[0];
        const float qc32[3] = {(float)(bc[1] / ws), (float)(bc[2] / ws), (float)(bc[3] / ws)};
        const float vc32[3] = {(float)(bc[4] / ws), (float)(bc[5] / ws), (float)(bc[6] / ws)};
        // energies; unbound members listed in index order
        uint32_t nu = 0;
        for (uint32_t base = s0; base < e0; base += kPB) {
            const uint32_t t = base + tid;
            bool unb = false;
            uint32_t key = 0;
            if (t < e0 && keep[t]) {
                const float4 q = qw[t];
                const float4 u = vv[t];
                const float d0 = __fadd_rn(__fsub_rn(u.x, vc32[0]), __fmul_rn(b32, __fsub_rn(q.x, qc32[0])));
                const float d1 = __fadd_rn(__fsub_rn(u.y, vc32[1]), __fmul_rn(b32, __fsub_rn(q.y, qc32[1])));
                const float d2 = __fadd_rn(__fsub_rn(u.z, vc32[2]), __fmul_rn(b32, __fsub_rn(q.z, qc32[2])));
                float sq = __fmul_rn(d0, d0);
                sq = __fadd_rn(sq, __fmul_rn(d1, d1));
                sq = __fadd_rn(sq, __fmul_rn(d2, d2));
                const float E = __fadd_rn(sq / 2.0f, phi[t]);
                if (!isfinite(E)) bad_any = 1;
                else if (E >= 0.0f) { unb = true; key = ekey(E); }
            }
            uint32_t tot;
            const uint32_t r = block_rank(unb, wcnt, tot);
            if (unb) { ul[s0 + nu + r] = t; uk[s0 + nu + r] = key; }
            nu += tot;
        }
        __syncthreads();
        if (bad_any) { st = -1; break; }
        if (nu == 0) {
            if (deleted == 0) st = 0;
            break;  // status OK, or refresh the potential
        }
        int nb = (int)ceil(0.02 * (double)np_bound);
        nb = nb < 1 ? 1 : nb;
        nb = nb > (int)nu ? (int)nu : nb;
        // radix select of the nb-th largest key in uk[s0, s0 + nu)
        uint32_t prefix = 0, mask = 0, need = (uint32_t)nb;
        for (int byte = 3; byte >= 0; --byte) {
            for (int i = tid; i < 256; i += kPB) hist[i] = 0;
            __syncthreads();
            for (uint32_t k = tid; k < nu; k += kPB) {
                const uint32_t key = uk[s0 + k];
                if ((key & mask) == prefix) atomicAdd(&hist[(key >> (8 * byte)) & 255u], 1u);
            }
            __syncthreads();
            select_digit256(hist, need, &sh[1], wcnt);
            prefix |= sh[1] << (8 * byte);
            mask |= 255u << (8 * byte);
            need = sh[2];
            __syncthreads();
        }
        // delete keys > T, and the `need` first (smallest index) keys == T
        const uint32_t T = prefix;
        uint32_t taken = 0;
        for (uint32_t base = 0; base < nu; base += kPB) {
            const uint32_t k = base + tid;
            bool del = false, tie = false;
            uint32_t t = 0;
            if (k < nu) {
                t = ul[s0 + k];
                const uint32_t key = uk[s0 + k];
                if (key > T) del = true;
                else if (key == T) tie = true;
            }
            uint32_t tot;
            const uint32_t r = block_rank(tie, wcnt, tot);
            if (tie && taken + r < need) del = true;
            if (del) keep[t] = 0;
            taken += tot;
        }
        __syncthreads();
        np_bound -= nb;
        deleted += nb;
        if (np_bound < P.np_min) { st = 1; break; }
        if ((double)deleted >= 0.2 * (double)np_bound) break;
    }
    __syncthreads();
    if (tid == 0) {
        npb[h] = np_bound;
        if (st == -1) {
            atomicCAS(err, 0, 1);
            status[h] = 5;
            active[h] = 0;
        } else if (st >= 0) {
            status[h] = st;
            active[h] = 0;
        } else if (last_pass) {
            status[h] = 5;
            active[h] = 0;
        }
    }
    // kept mass for the next potential (zero-mode term)
    if (large) {
        chunked_sums<1>(s0, e0, keep, qw, vv, part, red, bc);
        if (tid == 0) mkeep[h] = bc[0];
        return;
    }
    double m = 0.0;
    for (uint32_t t = s0 + tid; t < e0; t += kPB)
        if (keep[t]) m += (double)qw[t].w;
    for (int o = 32; o > 0; o >>= 1) m += __shfl_down(m, o, kWave);
    if (__lane_id() == 0) red[0][w] = m;
    __syncthreads();
    if (tid == 0) {
        double sm = 0.0;
        for (int j = 0; j < kPW; ++j) sm += red[0][j];
        mkeep[h] = sm;
    }
}

// ---------------------------------------------------------------------------
// The same removal loop for a large candidate on several workgroups (one
// cooperative launch, so that all its workgroups are resident together).
// Workgroup b of a candidate owns a contiguous chunk of its members and the
// matching stretch of the unbound list; the candidate's workgroups meet at a
// barrier in global memory (bounded spin) after each reduction. Values that
// other workgroups wrote are read with device-scope atomic loads, and sums
// over workgroups are taken in workgroup order, so every workgroup derives
// the same COM, threshold and decisions.


struct BigScratch {
    double part[kBigMaxB][8];
    uint32_t ucnt[kBigMaxB], tcnt[kBigMaxB];
    uint32_t hist[4][256];
    uint32_t bar_count, bar_gen, bad, timeout;
};

__device__ __forceinline__ uint32_t ld_dev(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_dev(const double* p) {
    return __longlong_as_double(
        __hip_atomic_load(reinterpret_cast<const long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// barrier of the nb workgroups of one candidate; false if the wait exceeded
// its bound (then the timeout flag is set and every workgroup stops)
__device__ bool big_barrier(BigScratch* S, uint32_t nb, int* ok_s) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __threadfence();
        const uint32_t gen = ld_dev(&S->bar_gen);
        if (atomicAdd(&S->bar_count, 1u) == nb - 1) {
            __hip_atomic_store(&S->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            atomicAdd(&S->bar_gen, 1u);
        } else {
            uint32_t spins = 0;
            while (ld_dev(&S->bar_gen) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > (1u << 24) || ld_dev(&S->timeout)) { ok = 0; break; }
            }
        }
        if (!ok) __hip_atomic_store(&S->timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        *ok_s = ok;
    }
    __syncthreads();
    return *ok_s != 0;
}

// rank of flagged threads in the workgroup (kB threads), and their total
__device__ __forceinline__ uint32_t wg_rank(bool flag, uint32_t* wcnt, uint32_t& total) {
    const int w = threadIdx.x / kWave;
    const uint64_t bal = __ballot(flag);
    const uint32_t in_wave = __popcll(bal & lanemask_lt());
    if (__lane_id() == 0) wcnt[w] = __popcll(bal);
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int k = 0; k < kBW; ++k) {
        const uint32_t c = wcnt[k];
        before += k < w ? c : 0u;
        total += c;
    }
    __syncthreads();
    return before + in_wave;
}

__global__ void __launch_bounds__(kB) k_pm_big(const int2* __restrict__ bmap, const int32_t* __restrict__ big_h,
                                               const int32_t* __restrict__ big_nb, BigScratch* __restrict__ scr,
                                               const uint32_t* __restrict__ hoff, const float4* __restrict__ qw,
                                               const float4* __restrict__ vv, const float* __restrict__ phi,
                                               uint32_t* __restrict__ ul, uint32_t* __restrict__ uk,
                                               uint8_t* __restrict__ keep, int32_t* __restrict__ status,
                                               int32_t* __restrict__ npb, uint8_t* __restrict__ active,
                                               double* __restrict__ mkeep, int32_t last_pass,
                                               int32_t* __restrict__ err, Params P) {
    __shared__ double red[8][kBW];
    __shared__ uint32_t lhist[256];
    __shared__ uint32_t wcnt[kBW];
    __shared__ double bc[8];
    __shared__ uint32_t sel[2];
    __shared__ int ok_s, bad_s;
    const int2 bm = bmap[blockIdx.x];
    const int32_t kb = bm.x;
    const uint32_t b = (uint32_t)bm.y;
    const int32_t h = big_h[kb];
    const uint32_t nb = (uint32_t)big_nb[kb];
    BigScratch* S = scr + kb;
    const uint32_t s0 = h ? hoff[h - 1] : 0u, e0 = hoff[h];
    const uint32_t n = e0 - s0, chunk = (n + nb - 1) / nb;
    const uint32_t c0 = s0 + min(n, b * chunk), c1 = s0 + min(n, (b + 1) * chunk);
    const int tid = threadIdx.x, w = tid / kWave;
    const float b32 = (float)P.b;
    int np_bound = npb[h];
    int deleted = 0, st = -2;
    bool ok = true;
    for (;;) {
        // the select histograms start from zero each iteration; every
        // workgroup finished reading the previous ones before the last barrier
        if (b == 0)
            for (int i = tid; i < 4 * 256; i += kB) (&S->hist[0][0])[i] = 0u;
        // centre of mass and bulk velocity, f64: partial sums per workgroup
        double v[7] = {0, 0, 0, 0, 0, 0, 0};
        for (uint32_t t = c0 + tid; t < c1; t += kB) {
            if (!keep[t]) continue;
            const float4 q = qw[t];
            const float4 u = vv[t];
            const double wi = (double)q.w;
            v[0] += wi;
            v[1] += wi * (double)q.x; v[2] += wi * (double)q.y; v[3] += wi * (double)q.z;
            v[4] += wi * (double)u.x; v[5] += wi * (double)u.y; v[6] += wi * (double)u.z;
        }
        for (int k = 0; k < 7; ++k)
            for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o, kWave);
        if (__lane_id() == 0) for (int k = 0; k < 7; ++k) red[k][w] = v[k];
        __syncthreads();
        if (tid < 7) {
            double sacc = 0.0;
            for (int j = 0; j < kBW; ++j) sacc += red[tid][j];
            S->part[b][tid] = sacc;
        }
        if (!(ok = big_barrier(S, nb, &ok_s))) break;
        if (tid < 7) {
            double sacc = 0.0;
            for (uint32_t j = 0; j < nb; ++j) sacc += ld_dev(&S->part[j][tid]);
            bc[tid] = sacc;
        }
        __syncthreads();
        const double ws = bc[0];
        const float qc32[3] = {(float)(bc[1] / ws), (float)(bc[2] / ws), (float)(bc[3] / ws)};
        const float vc32[3] = {(float)(bc[4] / ws), (float)(bc[5] / ws), (float)(bc[6] / ws)};
        // energies; this workgroup's unbound members listed in index order
        // from c0 on
        if (tid == 0) bad_s = 0;
        __syncthreads();
        uint32_t nu_b = 0;
        for (uint32_t base = c0; base < c1; base += kB) {
            const uint32_t t = base + tid;
            bool unb = false;
            uint32_t key = 0;
            if (t < c1 && keep[t]) {
                const float4 q = qw[t];
                const float4 u = vv[t];
                const float d0 = __fadd_rn(__fsub_rn(u.x, vc32[0]), __fmul_rn(b32, __fsub_rn(q.x, qc32[0])));
                const float d1 = __fadd_rn(__fsub_rn(u.y, vc32[1]), __fmul_rn(b32, __fsub_rn(q.y, qc32[1])));
                const float d2 = __fadd_rn(__fsub_rn(u.z, vc32[2]), __fmul_rn(b32, __fsub_rn(q.z, qc32[2])));
                float sq = __fmul_rn(d0, d0);
                sq = __fadd_rn(sq, __fmul_rn(d1, d1));
                sq = __fadd_rn(sq, __fmul_rn(d2, d2));
                const float E = __fadd_rn(sq / 2.0f, phi[t]);
                if (!isfinite(E)) bad_s = 1;
                else if (E >= 0.0f) { unb = true; key = ekey(E); }
            }
            uint32_t tot;
            const uint32_t r = wg_rank(unb, wcnt, tot);
            if (unb) { ul[c0 + nu_b + r] = t; uk[c0 + nu_b + r] = key; }
            nu_b += tot;
        }
        if (tid == 0) {
            S->ucnt[b] = nu_b;
            if (bad_s) atomicOr(&S->bad, 1u);
        }
        if (!(ok = big_barrier(S, nb, &ok_s))) break;
        uint32_t nu = 0;
        for (uint32_t j = 0; j < nb; ++j) nu += ld_dev(&S->ucnt[j]);
        if (ld_dev(&S->bad)) { st = -1; break; }
        if (nu == 0) {
            if (deleted == 0) st = 0;
            break;  // status OK, or refresh the potential
        }
        int nbd = (int)ceil(0.02 * (double)np_bound);
        nbd = nbd < 1 ? 1 : nbd;
        nbd = nbd > (int)nu ? (int)nu : nbd;
        // radix select of the nbd-th largest key over all workgroups' lists
        uint32_t prefix = 0, mask = 0, need = (uint32_t)nbd;
        for (int byte = 3; byte >= 0; --byte) {
            for (int i = tid; i < 256; i += kB) lhist[i] = 0;
            __syncthreads();
            for (uint32_t k = tid; k < nu_b; k += kB) {
                const uint32_t key = uk[c0 + k];
                if ((key & mask) == prefix) atomicAdd(&lhist[(key >> (8 * byte)) & 255u], 1u);
            }
            __syncthreads();
            for (int i = tid; i < 256; i += kB)
                if (lhist[i]) atomicAdd(&S->hist[byte][i], lhist[i]);
            if (!(ok = big_barrier(S, nb, &ok_s))) break;
            for (int i = tid; i < 256; i += kB) lhist[i] = ld_dev(&S->hist[byte][i]);
            __syncthreads();
            select_digit256(lhist, need, sel, wcnt);
            prefix |= sel[0] << (8 * byte);
            mask |= 255u << (8 * byte);
            need = sel[1];
            __syncthreads();
        }
        if (!ok) break;
        // delete keys > T; keys == T (ties) are deleted in index order, the
        // `need` first over all workgroups
        const uint32_t T = prefix;
        uint32_t ties = 0;
        for (uint32_t base = 0; base < nu_b; base += kB) {
            const uint32_t k = base + tid;
            bool tie = false;
            if (k < nu_b) {
                const uint32_t key = uk[c0 + k];
                if (key > T) keep[ul[c0 + k]] = 0;
                else if (key == T) tie = true;
            }
            uint32_t tot;
            wg_rank(tie, wcnt, tot);
            ties += tot;
        }
        if (tid == 0) S->tcnt[b] = ties;
        if (!(ok = big_barrier(S, nb, &ok_s))) break;
        uint32_t taken = 0;
        for (uint32_t j = 0; j < b; ++j) taken += ld_dev(&S->tcnt[j]);
        if (taken < need) {
            for (uint32_t base = 0; base < nu_b; base += kB) {
                const uint32_t k = base + tid;
                const bool tie = k < nu_b && uk[c0 + k] == T;
                uint32_t tot;
                const uint32_t r = wg_rank(tie, wcnt, tot);
                if (tie && taken + r < need) keep[ul[c0 + k]] = 0;
                taken += tot;
            }
        }
        np_bound -= nbd;
        deleted += nbd;
        if (np_bound < P.np_min) { st = 1; break; }
        if ((double)deleted >= 0.2 * (double)np_bound) break;
    }
    if (!ok) {
        if (b == 0 && tid == 0) {
            atomicCAS(err, 0, IERR_HIP);
            status[h] = 5;
            active[h] = 0;
        }
        return;
    }
    // kept mass for the next potential (zero-mode term)
    double m = 0.0;
    for (uint32_t t = c0 + tid; t < c1; t += kB)
        if (keep[t]) m += (double)qw[t].w;
    for (int o = 32; o > 0; o >>= 1) m += __shfl_down(m, o, kWave);
    if (__lane_id() == 0) red[7][w] = m;
    __syncthreads();
    if (tid == 0) {
        double sm = 0.0;
        for (int j = 0; j < kBW; ++j) sm += red[7][j];
        S->part[b][7] = sm;
    }
    if (!big_barrier(S, nb, &ok_s)) {
        if (b == 0 && tid == 0) atomicCAS(err, 0, IERR_HIP);
        return;
    }
    if (b == 0 && tid == 0) {
        double sm = 0.0;
        for (uint32_t j = 0; j < nb; ++j) sm += ld_dev(&S->part[j][7]);
        mkeep[h] = sm;
        npb[h] = np_bound;
        if (st == -1) {
            atomicCAS(err, 0, 1);
            status[h] = 5;
            active[h] = 0;
        } else if (st >= 0) {
            status[h] = st;
            active[h] = 0;
        } else if (last_pass) {
            status[h] = 5;
            active[h] = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// write-back (fof_fortran.f90:313-351)

__global__ void k_valid_flags(const int32_t* __restrict__ status, const int32_t* __restrict__ npb,
                              uint32_t nh, int32_t np_min, uint32_t* __restrict__ flag) {
    const uint32_t h = blockIdx.x * kB + threadIdx.x;
    if (h < nh) flag[h] = (status[h] == 0 && npb[h] >= np_min) ? 1u : 0u;
}

__global__ void k_bound_write(const uint32_t* __restrict__ members, const uint32_t* __restrict__ mhalo,
                              const uint8_t* __restrict__ keep, const uint32_t* __restrict__ vflag,
                              const uint32_t* __restrict__ vidx, uint32_t nmem,
                              int32_t* __restrict__ in_halo, int32_t* __restrict__ halo_id,
                              uint8_t* __restrict__ keep_out) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= nmem) return;
    const uint32_t h = mhalo[t];
    const bool ok = vflag[h] && keep[t];
    const uint32_t i = members[t];
    in_halo[i] = ok ? 1 : 0;
    halo_id[i] = ok ? (int32_t)vidx[h] + 1 : 0;
    keep_out[t] = ok ? 1 : 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// host

namespace {

struct SelCtx {
    int ngrid, bn[4], n_max, m_max;
    double s15, spacing_max;
    int fgl_nlev[4];
    double fgl_sp[4][32], fgl_cov[4][32];
    int ntot, tab_gid[64], tab_lev[64];
    double tab_cov[64];
};

// bound_set_context + level tables (bound_fortran.f90:291-330, 1854-1955)
void sel_init(SelCtx& c, double s15, int64_t ng) {
    static const int cand[4] = {48, 64, 128, 256};
    memset(&c, 0, sizeof(c));
    const int ngi = (int)ng;
    c.ngrid = 1 + (ngi >= 64) + (ngi >= 128) + (ngi >= 256);
    for (int g = 0; g < 4; ++g) c.bn[g] = std::min(cand[g], ngi);
    c.n_max = std::min(256, ngi);
    c.m_max = c.n_max / 2 - 3;
    c.s15 = s15;
    c.spacing_max = s15 / (double)c.m_max;
    for (int g = 1; g <= c.ngrid; ++g) {
        const int m = c.bn[g - 1] / 2 - 3;
        int nl = 0;
        double sp = 1.0 / 32.0;
        while ((double)m * sp < s15 && nl < 32) { c.fgl_sp[g - 1][nl++] = sp; sp = 2.0 * sp; }
        if (g == c.ngrid && nl < 32) c.fgl_sp[g - 1][nl++] = c.spacing_max;
        c.fgl_nlev[g - 1] = nl;
        for (int l = 0; l < nl; ++l) c.fgl_cov[g - 1][l] = (double)m * c.fgl_sp[g - 1][l];
    }
    int nt = 0;
    for (int g = 1; g <= c.ngrid; ++g)
        for (int l = 0; l < c.fgl_nlev[g - 1]; ++l) {
            c.tab_gid[nt] = g; c.tab_lev[nt] = l + 1; c.tab_cov[nt] = c.fgl_cov[g - 1][l]; ++nt;
        }
    for (int j = 0; j < nt - 1; ++j) {  // selection sort by coverage (bound_build_ladder)
        int k = j;
        for (int i = j + 1; i < nt; ++i) if (c.tab_cov[i] < c.tab_cov[k]) k = i;
        if (k != j) {
            std::swap(c.tab_gid[j], c.tab_gid[k]);
            std::swap(c.tab_lev[j], c.tab_lev[k]);
            std::swap(c.tab_cov[j], c.tab_cov[k]);
        }
    }
    c.ntot = nt;
}

// bound_select_grid_and_level (bound_fortran.f90:1963-2019); 0 ok, 2 none
int sel_grid(const SelCtx& c, int64_t np, double span, int& gid, double& sp_out) {
    static const int ncand[3] = {11, 10, 8};
    static const int cg[3][12] = {{1, 2, 2, 3, 3, 3, 4, 4, 4, 4, 4},
                                  {3, 3, 4, 4, 4, 4, 4, 4, 4, 4},
                                  {4, 4, 4, 4, 4, 4, 4, 4}};
    static const double cs[3][12] = {
        {0.125, 0.125, 0.25, 0.125, 0.25, 0.5, 0.0625, 0.125, 0.25, 0.5, -1.0},
        {0.0625, 0.125, 0.0625, 0.125, 0.25, 0.5, 1.0, 2.0, 4.0, -1.0},
        {0.0625, 0.125, 0.25, 0.5, 1.0, 2.0, 4.0, -1.0}};
    gid = 0;
    sp_out = 0.0;
    if (np <= 1000) return 0;
    const int band = np <= 5000 ? 0 : (np <= 20000 ? 1 : 2);
    for (int s = 0; s < ncand[band]; ++s) {
        const int g = cg[band][s];
        if (g < 1 || g > c.ngrid) continue;
        double sp = cs[band][s];
        if (sp == -1.0) sp = c.spacing_max;
        int l = 0;
        for (int k = 0; k < c.fgl_nlev[g - 1]; ++k)
            if (fabs(c.fgl_sp[g - 1][k] - sp) <= 1e-12 * std::max(1.0, fabs(sp))) { l = k + 1; break; }
        if (l == 0) continue;
        if (c.fgl_cov[g - 1][l - 1] < span) continue;
        gid = g;
        sp_out = c.fgl_sp[g - 1][l - 1];
        return 0;
    }
    if (span <= c.s15) {
        for (int i = 0; i < c.ntot; ++i)
            if (c.tab_cov[i] >= span) {
                gid = c.tab_gid[i];
                sp_out = c.fgl_sp[gid - 1][c.tab_lev[i] - 1];
                return 0;
            }
    }
    return 2;
}

// range 0: outer kernel of grid n; range > 0: inner band kernel
GreenCache& green_for(Ctx& C, int n, hipStream_t s, double range = 0.0) {
    GreenCache*& gc = C.greens[n + (range > 0.0 ? 100000 * (int)range : 0)];
    if (!gc) gc = new GreenCache();
    if (gc->ready) return *gc;
    gc->n = n;
    const uint64_t tot = (uint64_t)n * n * n;
    float* G = gc->G.as<float>(tot);
    float2* tw = gc->tw.as<float2>(n);
    fft_make_plan(gc->plan, n, tw, s);
    k_green<<<blocks(tot), kB, 0, s>>>(G, n, range, range > 0.0 ? 0 : 1);
    PYFOF_LAUNCH_CHECK();
    // self kernel: inverse transform of G, / n^3, at the 5^3 corner
    DevBuf tmp;
    float2* g = tmp.as<float2>(tot);
    int32_t* slot0 = C.buf("green_slot").as<int32_t>(1);
    PYFOF_HIP(hipMemsetAsync(slot0, 0, sizeof(int32_t), s));
    k_green_to_complex<<<blocks(tot), kB, 0, s>>>(G, g, tot);
    PYFOF_LAUNCH_CHECK();
    fft3d(gc->plan, g, slot0, 1, true, s);
    float* self = gc->self.as<float>(125);
    k_self_extract<<<1, 128, 0, s>>>(g, n, self);
    PYFOF_LAUNCH_CHECK();
    PYFOF_HIP(hipStreamSynchronize(s));
    tmp.release();
    gc->ready = true;
    return *gc;
}

}  // namespace

// bound_fortran.f90:279-289: the first error of a call is kept with its
// message, and printed at verbose >= 1
static int32_t bound_fail(const BoundArgs& A, BoundResult& res, int32_t code, const char* msg) {
    res.err_msg = msg;
    if (A.verbose >= 1) {
        printf(" [BOUND_ERROR] code=%12d  msg= %s\n", code, msg);
        fflush(stdout);
    }
    return code;
}

// bound_fortran.f90:2062-2080, printed at verbose >= 1 when a candidate is
// rejected for its span; box and h0 are the fixed 300 Mpc/h and 0.67 the
// reference's unbinding module uses for this report
static void span_report(const BoundArgs& A, const SelCtx& sc, uint32_t ident, double span) {
    const double box = 300.0, h0 = 0.67;
    const double l_ng = box / (h0 * A.ng_global);
    printf(" ===============================================================\n");
    printf(" bound grid selection: candidate rejected, no valid result written\n");
    printf("   subroutine           = bound_classify\n");
    printf("   reason               = candidate span exceeds D_max = 15 Mpc\n");
    printf("   candidate identifier = %12u\n", ident);
    printf("   actual span          = %25.17g  standard grid cells\n", span);
    printf("   actual span          = %25.17g  Mpc (comoving)\n", span * l_ng);
    printf("   upper limit D_max    = 15 Mpc = %25.17g  standard grid cells\n", A.s15);
    printf("   box [Mpc/h], h0      = %25.17g %25.17g\n", box, h0);
    printf("   ng, ng_global        = %12lld %25.17g\n", (long long)A.ng, A.ng_global);
    printf("   n_max, m_max         = %12d %12d\n", sc.n_max, sc.m_max);
    printf("   spacing_max          = %25.17g\n", sc.spacing_max);
    printf("   nlev ladder size     = %12d  ngrid = %12d\n", sc.ntot, sc.ngrid);
    printf("   L_ng [Mpc]           = %25.17g\n", l_ng);
    printf(" ===============================================================\n");
    fflush(stdout);
}

static int32_t run_bound_impl(Ctx& C, const BoundArgs& A, uint8_t* keep_out, BoundResult& res) {
    hipStream_t s = C.stream;
    res = BoundResult();
    const uint32_t nh = A.nh, M = A.nmem;
    if (nh == 0 || M == 0) return IERR_OK;
    const auto t_start = std::chrono::steady_clock::now();
    auto fail = [&](int32_t code, const char* msg) { return bound_fail(A, res, code, msg); };
    SelCtx sc;
    sel_init(sc, A.s15, A.ng);
    // bound_fortran.f90:1559-1563, 1586-1590
    if (sc.ngrid != 4) return fail(IERR_INVALID, "bound_check_batch: fixed groups require ng >= 256");
    if (!(A.a > 0.0) || !std::isfinite(A.a) || !std::isfinite(A.b))
        return fail(IERR_INVALID, "bound_check_batch: invalid physical context");
    if (A.direct_max < kDirectMax || A.direct_max > kDirectLargeMax)
        return fail(IERR_INVALID, "direct-summation threshold out of range");
    Params P{A.app, A.a, A.b, A.L, A.Lh, A.np_min};

    float4* qw = C.buf("qw").as<float4>(M);
    float4* vv = C.buf("vv").as<float4>(M);
    k_bound_pack<<<blocks(M), kB, 0, s>>>(A.members, A.mhalo, A.anchor, A.xp, A.vp, A.mp, M, A.L, A.Lh, qw, vv);
    PYFOF_LAUNCH_CHECK();
    float* d_lohi = C.buf("lohi").as<float>(6 * (size_t)nh);
    k_bound_span<<<nh, kB, 0, s>>>(qw, A.hoff, d_lohi);
    PYFOF_LAUNCH_CHECK();
    std::vector<float> lohi(6 * (size_t)nh);
    std::vector<uint32_t> hoff(nh);
    PYFOF_HIP(hipMemcpyAsync(lohi.data(), d_lohi, 4 * lohi.size(), hipMemcpyDeviceToHost, s));
    PYFOF_HIP(hipMemcpyAsync(hoff.data(), A.hoff, 4 * (size_t)nh, hipMemcpyDeviceToHost, s));
    PYFOF_HIP(hipStreamSynchronize(s));
    auto hsize = [&](uint32_t h) { return (int64_t)hoff[h] - (h ? (int64_t)hoff[h - 1] : 0); };

    // classification in candidate order (bound_fortran.f90:2940-3055).
    // kind: 0 direct in LDS (<= 1000), 1 direct in global memory (up to the
    // context's threshold; 1000 by default, as the reference), 2 mesh
    std::vector<int8_t> kind(nh, 0);
    std::vector<int32_t> pm_of_halo(nh, -1), dsmall, dlarge;
    std::vector<PmHalo> pm;
    std::vector<int32_t> pm_halo;
    for (uint32_t h = 0; h < nh; ++h) {
        const int64_t n = hsize(h);
        const float* lh = &lohi[6 * (size_t)h];
        float span = lh[3] - lh[0];
        span = std::max(span, lh[4] - lh[1]);
        span = std::max(span, lh[5] - lh[2]);
        if (!std::isfinite(span)) {
            printf(" bound_classify: non-finite span\n");  // bound_fortran.f90:2966-2970
            fflush(stdout);
            return fail(IERR_INVALID, "bound_classify: non-finite span");
        }
        if (n > 1000 && span > (float)A.s15) {
            const int32_t e = fail(IERR_SPAN, "candidate span exceeds D_max = 15 Mpc");
            if (A.verbose >= 1) span_report(A, sc, h ? hoff[h - 1] + 1 : 1u, (double)span);
            return e;
        }
        if (n <= A.direct_max) {
            if (n <= kDirectMax) {
                dsmall.push_back((int32_t)h);
            } else {
                kind[h] = 1;
                dlarge.push_back((int32_t)h);
            }
            continue;
        }
        int gid;
        double sp;
        if (sel_grid(sc, n, (double)span, gid, sp) != 0 || gid == 0)  // n > 1000 always selects a grid
            return fail(IERR_INVALID, "no enabled grid class or direct combination can cover candidate");
        const float sp32 = (float)sp;
        if ((double)(sc.bn[gid - 1] / 2 - 3) * (double)sp32 < (double)span)
            return fail(IERR_INVALID, "actual single precision spacing does not cover span");
        PmHalo H;
        memset(&H, 0, sizeof(H));
        H.origin = make_float4(lh[0] - sp32, lh[1] - sp32, lh[2] - sp32, sp32);
        H.cls = gid - 1;
        H.n = sc.bn[gid - 1];
        kind[h] = 2;
        pm_of_halo[h] = (int32_t)pm.size();
        pm.push_back(H);
        pm_halo.push_back((int32_t)h);
    }
    res.n_direct = (uint32_t)(dsmall.size() + dlarge.size());
    // bound_fortran.f90:1679-1681. The batch is one phase on the GPU: every
    // candidate is processed concurrently, by workgroups of kB threads.
    if (A.verbose >= 1) {
        printf("[SCHEDULE] phase=1 owners=%u team_threads=%d\n", nh, kB);
        fflush(stdout);
    }

    int32_t* status = C.buf("bstatus").as<int32_t>(nh);
    int32_t* npb = C.buf("bnpb").as<int32_t>(nh);
    uint8_t* keep = C.buf("bkeep").as<uint8_t>(M);
    int32_t* err = C.buf("berr").as<int32_t>(4);
    PYFOF_HIP(hipMemsetAsync(err, 0, 16, s));
    PYFOF_HIP(hipMemsetAsync(keep, 1, M, s));
    {
        std::vector<int32_t> npb0(nh), st0(nh, 5);
        for (uint32_t h = 0; h < nh; ++h) npb0[h] = (int32_t)hsize(h);
        PYFOF_HIP(hipMemcpyAsync(npb, npb0.data(), 4 * (size_t)nh, hipMemcpyHostToDevice, s));
        PYFOF_HIP(hipMemcpyAsync(status, st0.data(), 4 * (size_t)nh, hipMemcpyHostToDevice, s));
    }
    int8_t* d_kind = C.buf("bkind").as<int8_t>(nh);
    PYFOF_HIP(hipMemcpyAsync(d_kind, kind.data(), nh, hipMemcpyHostToDevice, s));
    float* phi_rec = A.phi_rec;

    auto direct_small = [&](const int32_t* list_dev, uint32_t count) {
        if (A.direct_f32)
            k_direct_solve<true><<<count, kB, 0, s>>>(list_dev, A.hoff, qw, vv, keep, status, npb, err, phi_rec, P);
        else
            k_direct_solve<false><<<count, kB, 0, s>>>(list_dev, A.hoff, qw, vv, keep, status, npb, err, phi_rec, P);
        PYFOF_LAUNCH_CHECK();
    };

    // ---- state of the iterative halos (mesh and large direct)
    const uint32_t npm = (uint32_t)pm.size();
    const uint32_t nit = npm + (uint32_t)dlarge.size();
    int32_t passes = 0;
    float* phi = C.buf("phi").as<float>(M);
    uint32_t* ulist = C.buf("ulist").as<uint32_t>(M);
    uint32_t* ukey = C.buf("ukey").as<uint32_t>(M);
    uint8_t* active = C.buf("active").as<uint8_t>(nh);
    double* mkeep = C.buf("mkeep").as<double>(nh);
    uint8_t* keep_prev = C.buf("keepc2").as<uint8_t>(M);
    uint32_t* dflag = C.buf("c2dflag").as<uint32_t>(M);
    uint32_t* didx = C.buf("c2didx").as<uint32_t>(M);
    uint32_t* dtot = C.buf("c2dtot").as<uint32_t>(1);
    uint32_t* dk = C.buf("c2dk").as<uint32_t>(M);
    uint32_t* dv = C.buf("c2dv").as<uint32_t>(M);
    uint32_t* dk2 = C.buf("c2dk2").as<uint32_t>(M);
    uint32_t* dv2 = C.buf("c2dv2").as<uint32_t>(M);
    int32_t* d_alist = C.buf("alist").as<int32_t>(std::max<uint32_t>(nit, 1));
    PYFOF_HIP(hipMemsetAsync(active, 0, nh, s));

    // mesh set-up: Green functions, C2 cells, TSC stencils, grids
    int nclass[4] = {0, 0, 0, 0};
    const float* Gfull[4] = {nullptr, nullptr, nullptr, nullptr};
    const FftPlan* plans[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned long long* acc[4] = {nullptr, nullptr, nullptr, nullptr};
    float2* grid[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t* d_slot_halo[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t *d_pmof = nullptr, *d_slots = nullptr;
    PmHalo* d_pm = nullptr;
    uint64_t c2cells = 0;
    const uint32_t *clist = nullptr;
    uint32_t cls_first[4] = {0, 0, 0, 0}, cls_count[4] = {0, 0, 0, 0}, cls_end_prev = 0;
    uint32_t *cell_of = nullptr, *cstart = nullptr, *dstart = nullptr;
    int4* first = nullptr;
    float *wt = nullptr, *selfv = nullptr, *phi_c1 = nullptr, *phi_c2 = nullptr;
    double* fscale = nullptr;
    void* tmp = C.buf("c2tmp").get(std::max({radix_temp_bytes(M), scan_temp_bytes(M), scan_temp_bytes(nh + 1)}) + 256);
    // two-level state (256^3 candidates with refined blocks)
    bool tl_on = false;
    std::vector<TlSlot> tl_slots;
    int tl_nslot[2] = {0, 0};  // slots of ratio 2 and 4
    TlSlot* d_tlslots = nullptr;
    int32_t *src_slot = nullptr, *tgt_slot = nullptr, *d_tl_ls = nullptr, *d_tl_sh = nullptr;
    uint32_t* src_member = nullptr;
    int4 *src_first = nullptr, *tgt_first = nullptr;
    float *src_wt = nullptr, *tgt_wt = nullptr, *self_in = nullptr, *phi_in = nullptr;
    uint8_t* pr = nullptr;
    uint32_t nsrc = 0;
    const uint32_t* src_order = nullptr;  // entries sorted by slot rank and position (or null)
    std::vector<uint32_t> rank_start;     // first sorted entry of each slot rank
    const float* Gin[2] = {nullptr, nullptr};
    const FftPlan* plan_in[2] = {nullptr, nullptr};
    unsigned long long* acc_in[2] = {nullptr, nullptr};
    float2* grid_in[2] = {nullptr, nullptr};
    const int tl_chunk[2] = {96, 24};  // inner grids per launch: 3.2 GB (128^3), 3.6 GB (216^3)
    if (npm > 0) {
        for (auto& H : pm) H.slot = nclass[H.cls]++;
        for (int g = 0; g < 4; ++g) res.n_pm[g] = (uint32_t)nclass[g];
        float* selfs = C.buf("selfs").as<float>(4 * 125);
        for (int g = 0; g < 4; ++g) {
            if (!nclass[g]) continue;
            GreenCache& gc = green_for(C, sc.bn[g], s);
            Gfull[g] = gc.G.as<float>(1);
            plans[g] = &gc.plan;
            PYFOF_HIP(hipMemcpyAsync(selfs + 125 * g, gc.self.as<float>(125), 125 * 4, hipMemcpyDeviceToDevice, s));
        }
        // C2 cell grids per PM halo (bound_fortran.f90:3125-3145)
        std::vector<int32_t> pm_list(npm);
        for (uint32_t k = 0; k < npm; ++k) pm_list[k] = (int32_t)k;
        int32_t* d_pmlist = C.buf("pmlist").as<int32_t>(npm);
        int32_t* d_pmhalo = C.buf("pmhalo").as<int32_t>(npm);
        d_pm = C.buf("pm").as<PmHalo>(npm);
        d_pmof = C.buf("pmof").as<int32_t>(nh);
        PYFOF_HIP(hipMemcpyAsync(d_pmlist, pm_list.data(), 4 * (size_t)npm, hipMemcpyHostToDevice, s));
        PYFOF_HIP(hipMemcpyAsync(d_pmhalo, pm_halo.data(), 4 * (size_t)npm, hipMemcpyHostToDevice, s));
        PYFOF_HIP(hipMemcpyAsync(d_pm, pm.data(), sizeof(PmHalo) * npm, hipMemcpyHostToDevice, s));
        PYFOF_HIP(hipMemcpyAsync(d_pmof, pm_of_halo.data(), 4 * (size_t)nh, hipMemcpyHostToDevice, s));
        double* uext = C.buf("uext").as<double>(6 * (size_t)npm);
        k_c2_extent<<<npm, kB, 0, s>>>(qw, A.hoff, d_pmlist, d_pmhalo, d_pm, uext);
        PYFOF_LAUNCH_CHECK();
        std::vector<double> ue(6 * (size_t)npm);
        PYFOF_HIP(hipMemcpyAsync(ue.data(), uext, 8 * ue.size(), hipMemcpyDeviceToHost, s));
        PYFOF_HIP(hipStreamSynchronize(s));
        for (uint32_t k = 0; k < npm; ++k) {
            PmHalo& H = pm[k];
            for (int d = 0; d < 3; ++d) {
                const int nc = (int)floor((ue[6 * k + 3 + d] - ue[6 * k + d]) / 3.5) + 1;
                H.nc[d] = nc < 1 ? 1 : nc;
                H.umin[d] = ue[6 * k + d];
            }
        }
        // C2 cells numbered class by class, so that after the sort by cell the
        // members of each grid class form one range of the C2 member list
        // (the deposit walks that range); the order inside a cell is the
        // member order either way
        for (int g = 0; g < 4; ++g) {
            cls_first[g] = cls_end_prev;
            for (uint32_t k = 0; k < npm; ++k) {
                PmHalo& H = pm[k];
                if (H.cls != g) continue;
                H.c2off = (int32_t)c2cells;
                c2cells += (uint64_t)H.nc[0] * H.nc[1] * H.nc[2];
                cls_end_prev += (uint32_t)hsize((uint32_t)pm_halo[k]);
            }
            cls_count[g] = cls_end_prev - cls_first[g];
        }
        if (c2cells >= 0x7fffffffull) return fail(IERR_TOO_LARGE, "residual cell table exceeds 32-bit indices");
        PYFOF_HIP(hipMemcpyAsync(d_pm, pm.data(), sizeof(PmHalo) * npm, hipMemcpyHostToDevice, s));
        // member -> C2 cell; members sorted by (cell, member index)
        uint32_t* ck = C.buf("c2k").as<uint32_t>(M);
        uint32_t* cv = C.buf("c2v").as<uint32_t>(M);
        uint32_t* ck2 = C.buf("c2k2").as<uint32_t>(M);
        uint32_t* cv2 = C.buf("c2v2").as<uint32_t>(M);
        k_c2_keys<<<blocks(M), kB, 0, s>>>(qw, A.mhalo, d_pmof, d_pm, M, (uint32_t)c2cells, ck, cv);
        PYFOF_LAUNCH_CHECK();
        tmp = C.buf("c2tmp").get(std::max({radix_temp_bytes(M), scan_temp_bytes(c2cells + 1), scan_temp_bytes(M),
                                           scan_temp_bytes(nh + 1)}) + 256);
        const bool alt = radix_sort_pairs(ck, cv, ck2, cv2, M, bits_for(c2cells), tmp, s);
        const uint32_t* skeys = alt ? ck2 : ck;
        clist = alt ? cv2 : cv;
        cell_of = C.buf("c2cellof").as<uint32_t>(M);
        cstart = C.buf("c2start").as<uint32_t>(c2cells + 1);
        k_fill32<<<blocks(c2cells + 1), kB, 0, s>>>(cstart, c2cells + 1, 0u);
        PYFOF_LAUNCH_CHECK();
        k_c2_cells<<<blocks(M), kB, 0, s>>>(skeys, clist, M, (uint32_t)c2cells, cell_of, cstart + 1);
        PYFOF_LAUNCH_CHECK();
        inclusive_max_scan_u32(cstart + 1, cstart + 1, c2cells, tmp, s);

        // TSC setup
        first = C.buf("first").as<int4>(M);
        wt = C.buf("wt").as<float>(9 * (size_t)M);
        selfv = C.buf("selfv").as<float>(M);
        k_tsc_setup<<<blocks(M), kB, 0, s>>>(qw, A.mhalo, d_pmof, d_pm, selfs, M, first, wt, selfv, err);
        PYFOF_LAUNCH_CHECK();
        {
            // a member outside the source domain leaves its stencil unset:
            // stop here (ierr 1, bound_fortran.f90:2571-2575) before any deposit
            int32_t herr = 0;
            PYFOF_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
            PYFOF_HIP(hipStreamSynchronize(s));
            if (herr) return herr;
        }
        for (int g = 0; g < 4; ++g) {
            if (!nclass[g]) continue;
            const uint64_t gsz = (uint64_t)sc.bn[g] * sc.bn[g] * sc.bn[g] * (uint64_t)nclass[g];
            acc[g] = C.buf("acc" + std::to_string(g)).as<unsigned long long>(gsz);
            grid[g] = C.buf("grid" + std::to_string(g)).as<float2>(gsz);
        }
        std::vector<std::vector<int32_t>> slot_halo(4);
        for (int g = 0; g < 4; ++g) slot_halo[g].assign(nclass[g], 0);
        for (uint32_t k = 0; k < npm; ++k) slot_halo[pm[k].cls][pm[k].slot] = pm_halo[k];
        for (int g = 0; g < 4; ++g) {
            if (!nclass[g]) continue;
            d_slot_halo[g] = C.buf("slothalo" + std::to_string(g)).as<int32_t>(nclass[g]);
            PYFOF_HIP(hipMemcpyAsync(d_slot_halo[g], slot_halo[g].data(), 4 * (size_t)nclass[g], hipMemcpyHostToDevice, s));
        }
        // ---- two-level refinement (bound_fortran.f90:551-823, 2603-2606)
        std::vector<int32_t> tl_of_halo(nh, -1), tl_halos;
        for (uint32_t k = 0; k < npm; ++k) {
            // a 256^3 candidate; no block survives the demotion when 3.5 sp / 2 < app
            if (pm[k].cls != 3 || 3.5 * (double)pm[k].origin.w / 2.0 < (double)(float)A.app) continue;
            tl_of_halo[pm_halo[k]] = (int32_t)tl_halos.size();
            tl_halos.push_back(pm_halo[k]);
        }
        if (!tl_halos.empty()) {
            const size_t ntl = tl_halos.size();
            int16_t* treg = C.buf("tlreg").as<int16_t>(M);
            uint32_t* tcnt = C.buf("tlcnt").as<uint32_t>(ntl * kTlRegions);
            int32_t* d_tlof = C.buf("tlof").as<int32_t>(nh);
            PYFOF_HIP(hipMemcpyAsync(d_tlof, tl_of_halo.data(), 4 * (size_t)nh, hipMemcpyHostToDevice, s));
            k_fill32<<<blocks(ntl * kTlRegions), kB, 0, s>>>(tcnt, ntl * kTlRegions, 0u);
            PYFOF_LAUNCH_CHECK();
            k_tl_count<<<blocks(M), kB, 0, s>>>(qw, A.mhalo, d_tlof, d_pmof, d_pm, M, treg, tcnt);
            PYFOF_LAUNCH_CHECK();
            std::vector<uint32_t> cnt(ntl * kTlRegions);
            PYFOF_HIP(hipMemcpyAsync(cnt.data(), tcnt, 4 * cnt.size(), hipMemcpyDeviceToHost, s));
            PYFOF_HIP(hipStreamSynchronize(s));
            // ratios with demotion (bound_fortran.f90:569-586), refined blocks -> slots
            std::vector<int8_t> region_r(ntl * kTlRegions, 1);
            std::vector<int32_t> reg_slot(ntl * kTlRegions, -1);
            for (size_t k = 0; k < ntl; ++k) {
                PmHalo& H = pm[pm_of_halo[tl_halos[k]]];
                const double sp = (double)H.origin.w, app32 = (double)(float)A.app;
                for (int b = 0; b < kTlRegions; ++b) {
                    const uint32_t c = cnt[k * kTlRegions + b];
                    int r = c < (uint32_t)kTlN12 ? 1 : (c < (uint32_t)kTlN24 ? 2 : 4);
                    if (r == 4 && 3.5 * sp / 4.0 < app32) r = 2;
                    if (r == 2 && 3.5 * sp / 2.0 < app32) r = 1;
                    region_r[k * kTlRegions + b] = (int8_t)r;
                    if (c == 0 || r == 1) continue;
                    TlSlot S;
                    const int bb[3] = {b % 8, (b / 8) % 8, b / 64};
                    S.bo = 3.5f + 3.0f / (float)r;
                    for (int d = 0; d < 3; ++d) {
                        S.lo[d] = (float)(bb[d] * 16);
                        S.hi[d] = fminf(S.lo[d] + 16.0f, 127.0f);
                        S.u0[d] = S.lo[d] - S.bo - (1.0f / (float)r) + (0.0f / (float)r);
                    }
                    S.r = r;
                    S.n_in = r == 2 ? kTlN2 : kTlN4;
                    S.halo = tl_halos[k];
                    S.cls_slot = tl_nslot[r == 2 ? 0 : 1]++;
                    S.sp = H.origin.w;
                    reg_slot[k * kTlRegions + b] = (int32_t)tl_slots.size();
                    tl_slots.push_back(S);
                    H.tl = 1;
                }
            }
            tl_on = !tl_slots.empty();
            if (tl_on) {
                PYFOF_HIP(hipMemcpyAsync(d_pm, pm.data(), sizeof(PmHalo) * npm, hipMemcpyHostToDevice, s));
                int8_t* d_rr = C.buf("tlrr").as<int8_t>(region_r.size());
                int32_t* d_rs = C.buf("tlrs").as<int32_t>(reg_slot.size());
                d_tlslots = C.buf("tlslots").as<TlSlot>(tl_slots.size());
                PYFOF_HIP(hipMemcpyAsync(d_rr, region_r.data(), region_r.size(), hipMemcpyHostToDevice, s));
                PYFOF_HIP(hipMemcpyAsync(d_rs, reg_slot.data(), 4 * reg_slot.size(), hipMemcpyHostToDevice, s));
                PYFOF_HIP(hipMemcpyAsync(d_tlslots, tl_slots.data(), sizeof(TlSlot) * tl_slots.size(),
                                         hipMemcpyHostToDevice, s));
                pr = C.buf("tlpr").as<uint8_t>(M);
                k_tl_ratio<<<blocks(M), kB, 0, s>>>(A.mhalo, d_tlof, treg, d_rr, M, pr);
                PYFOF_LAUNCH_CHECK();
                // sources: count, scan, fill
                uint32_t* scnt = C.buf("tlscnt").as<uint32_t>(M);
                uint32_t* soff = C.buf("tlsoff").as<uint32_t>(M);
                uint32_t* stot = C.buf("tlstot").as<uint32_t>(1);
                k_tl_sources<false><<<blocks(M), kB, 0, s>>>(qw, A.mhalo, d_tlof, d_pmof, d_pm, treg, d_rs,
                                                             d_tlslots, M, scnt, nullptr, nullptr, nullptr,
                                                             nullptr, nullptr, err);
                PYFOF_LAUNCH_CHECK();
                exclusive_scan_u32(scnt, soff, M, stot, tmp, s);
                PYFOF_HIP(hipMemcpyAsync(&nsrc, stot, 4, hipMemcpyDeviceToHost, s));
                PYFOF_HIP(hipStreamSynchronize(s));
                src_slot = C.buf("tlsslot").as<int32_t>(std::max<uint32_t>(nsrc, 1));
                src_member = C.buf("tlsmem").as<uint32_t>(std::max<uint32_t>(nsrc, 1));
                src_first = C.buf("tlsfirst").as<int4>(std::max<uint32_t>(nsrc, 1));
                src_wt = C.buf("tlswt").as<float>(9 * (size_t)std::max<uint32_t>(nsrc, 1));
                k_tl_sources<true><<<blocks(M), kB, 0, s>>>(qw, A.mhalo, d_tlof, d_pmof, d_pm, treg, d_rs,
                                                            d_tlslots, M, nullptr, soff, src_slot, src_member,
                                                            src_first, src_wt, err);
                PYFOF_LAUNCH_CHECK();
                // entries sorted by slot and position: each chunk of inner
                // grids is one range, and neighbouring entries share LDS boxes
                const uint32_t nrank = (uint32_t)(tl_nslot[0] + tl_nslot[1]);
                if (nsrc > 0 && nrank < (1u << 14)) {
                    uint32_t* sk = C.buf("tlsk").as<uint32_t>(nsrc);
                    uint32_t* sv = C.buf("tlsv").as<uint32_t>(nsrc);
                    uint32_t* sk2 = C.buf("tlsk2").as<uint32_t>(nsrc);
                    uint32_t* sv2 = C.buf("tlsv2").as<uint32_t>(nsrc);
                    uint32_t* rend = C.buf("tlrend").as<uint32_t>(nrank);
                    void* stmp = C.buf("tlstmp").get(std::max(radix_temp_bytes(nsrc), scan_temp_bytes(nrank)) + 256);
                    k_tl_src_keys<<<blocks(nsrc), kB, 0, s>>>(src_slot, src_first, d_tlslots, nsrc, tl_nslot[0], sk,
                                                              sv);
                    PYFOF_LAUNCH_CHECK();
                    const bool salt = radix_sort_pairs(sk, sv, sk2, sv2, nsrc, bits_for(((uint64_t)nrank << 18) - 1),
                                                       stmp, s);
                    src_order = salt ? sv2 : sv;
                    // rank r's entries end at rend[r] (empty ranks: the end of
                    // the previous one, by a max-scan)
                    k_fill32<<<blocks(nrank), kB, 0, s>>>(rend, nrank, 0u);
                    PYFOF_LAUNCH_CHECK();
                    k_rank_ends<<<blocks(nsrc), kB, 0, s>>>(salt ? sk2 : sk, nsrc, rend);
                    PYFOF_LAUNCH_CHECK();
                    inclusive_max_scan_u32(rend, rend, nrank, stmp, s);
                    std::vector<uint32_t> re(nrank);
                    PYFOF_HIP(hipMemcpyAsync(re.data(), rend, 4 * (size_t)nrank, hipMemcpyDeviceToHost, s));
                    PYFOF_HIP(hipStreamSynchronize(s));
                    rank_start.assign(nrank + 1, 0u);
                    for (uint32_t r = 0; r < nrank; ++r) rank_start[r + 1] = re[r];
                }
                // inner kernels (bound_init_inner_gk) and targets
                for (int c = 0; c < 2; ++c) {
                    if (!tl_nslot[c]) continue;
                    const int n_in = c == 0 ? kTlN2 : kTlN4;
                    GreenCache& gc = green_for(C, n_in, s, c == 0 ? 7.0 : 14.0);
                    Gin[c] = gc.G.as<float>(1);
                    plan_in[c] = &gc.plan;
                    const uint64_t cap = (uint64_t)std::min(tl_nslot[c], tl_chunk[c]) * n_in * n_in * n_in;
                    acc_in[c] = C.buf(c == 0 ? "tlacc2" : "tlacc4").as<unsigned long long>(cap);
                    grid_in[c] = C.buf(c == 0 ? "tlgrid2" : "tlgrid4").as<float2>(cap);
                }
                const float* self2 = tl_nslot[0] ? green_for(C, kTlN2, s, 7.0).self.as<float>(125) : nullptr;
                const float* self4 = tl_nslot[1] ? green_for(C, kTlN4, s, 14.0).self.as<float>(125) : nullptr;
                tgt_slot = C.buf("tltslot").as<int32_t>(M);
                tgt_first = C.buf("tltfirst").as<int4>(M);
                tgt_wt = C.buf("tltwt").as<float>(9 * (size_t)M);
                self_in = C.buf("tlself").as<float>(M);
                k_tl_targets<<<blocks(M), kB, 0, s>>>(qw, A.mhalo, d_tlof, d_pmof, d_pm, treg, d_rs, d_tlslots,
                                                      self2, self4, M, tgt_slot, tgt_first, tgt_wt, self_in, err);
                PYFOF_LAUNCH_CHECK();
                phi_in = C.buf("tlphiin").as<float>(M);
                PYFOF_HIP(hipMemsetAsync(phi_in, 0, 4 * (size_t)M, s));
                d_tl_ls = C.buf("tlls").as<int32_t>(std::max(tl_chunk[0], tl_chunk[1]));
                d_tl_sh = C.buf("tlsh").as<int32_t>(std::max(tl_chunk[0], tl_chunk[1]));
                int32_t herr = 0;
                PYFOF_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
                PYFOF_HIP(hipStreamSynchronize(s));
                if (herr) return herr;  // a stencil left its inner grid (ierr 1)
            }
        }
        fscale = C.buf("fscale").as<double>(nh);
        phi_c1 = C.buf("phic1").as<float>(M);
        phi_c2 = C.buf("phic2").as<float>(M);
        dstart = C.buf("c2dstart").as<uint32_t>(c2cells + 1);
        d_slots = C.buf("slots").as<int32_t>(npm);
        k_halo_mass<<<npm, kB, 0, s>>>(qw, A.hoff, d_pmhalo, mkeep, fscale);
        PYFOF_LAUNCH_CHECK();
    }
    double* phid = dlarge.empty() ? nullptr : C.buf("phid").as<double>(M);
    uint32_t* hstart = dlarge.empty() ? nullptr : C.buf("dlstart").as<uint32_t>(nh + 1);
    int2* d_tiles = nullptr;

    // the pass loop over a set of mesh / large-direct halos
    // (bound_fortran.f90:2613-2864): potential, then the removal loop
    // workgroups of k_pm_big that can be resident at once (0: large
    // candidates stay on k_pm_inner)
    int64_t big_cap = 0;
    {
        int dev = 0, coop = 0, ncu = 0, per_cu = 0;
        PYFOF_HIP(hipGetDevice(&dev));
        PYFOF_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
        PYFOF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        PYFOF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_pm_big), kB, 0));
        if (A.multi_wg && coop && per_cu > 0) big_cap = (int64_t)per_cu * ncu;
    }
    const uint32_t* hoff_p = A.hoff;

    auto iterate = [&](std::vector<int32_t> act_list) -> int32_t {
        if (act_list.empty()) return IERR_OK;
        std::vector<uint8_t> on(nh, 0);
        for (int32_t h : act_list) on[h] = 1;
        PYFOF_HIP(hipMemcpyAsync(active, on.data(), nh, hipMemcpyHostToDevice, s));
        bool any_pm = false, any_dl = false;
        for (int32_t h : act_list) (kind[h] == 2 ? any_pm : any_dl) = true;
        for (int pass = 1; pass <= 20 && !act_list.empty(); ++pass) {
            passes = std::max(passes, pass);
            // ---- mesh halos: C1 by class, then C2
            if (any_pm) {
                for (int g = 0; g < 4; ++g) {
                    std::vector<int32_t> sl;
                    for (int32_t h : act_list) {
                        if (kind[h] != 2) continue;
                        const PmHalo& H = pm[pm_of_halo[h]];
                        if (H.cls == g) sl.push_back(H.slot);
                    }
                    if (sl.empty()) continue;
                    const int n = sc.bn[g];
                    PYFOF_HIP(hipMemcpyAsync(d_slots, sl.data(), 4 * sl.size(), hipMemcpyHostToDevice, s));
                    // deposit nodes [0, m]^3, m = n/2 + 1 (k_tsc_setup's domain guard)
                    const uint32_t mc = (uint32_t)(n / 2 + 2);
                    const dim3 gz((uint32_t)blocks((uint64_t)mc * mc * mc), (uint32_t)sl.size());
                    k_zero_region<<<gz, kB, 0, s>>>(acc[g], d_slots, n, mc);
                    PYFOF_LAUNCH_CHECK();
                    k_deposit<kDepTB, kDepBox><<<(uint32_t)((cls_count[g] + kDepTB - 1) / kDepTB), kDepTB, 0, s>>>(qw, A.mhalo, d_pmof, d_pm, keep, first, wt,
                                                                  active, fscale, g, clist + cls_first[g],
                                                                  cls_count[g], acc[g]);
                    PYFOF_LAUNCH_CHECK();
                    C.fft_begin();
                    fft_conv3d(*plans[g], C.fft_lines_for(n), grid[g], acc[g], d_slots, (int32_t)sl.size(),
                               d_slot_halo[g], fscale, Gfull[g], s);
                    C.fft_end(n, (int)sl.size());
                    k_interp<<<blocks(M), kB, 0, s>>>(qw, A.mhalo, d_pmof, d_pm, keep, first, wt, selfv, active,
                                                      mkeep, grid[g], g, M, A.a, phi_c1);
                    PYFOF_LAUNCH_CHECK();
                }
                // two-level band potentials on the inner grids (bound_fortran.f90:825-942)
                for (int c = 0; tl_on && c < 2; ++c) {
                    const int ratio = c == 0 ? 2 : 4;
                    const int n_in = c == 0 ? kTlN2 : kTlN4;
                    for (int cs = 0; cs < tl_nslot[c]; cs += tl_chunk[c]) {
                        std::vector<int32_t> ls, sh(tl_chunk[c], 0);
                        for (const TlSlot& S : tl_slots) {
                            if (S.r != ratio || S.cls_slot < cs || S.cls_slot >= cs + tl_chunk[c] || !on[S.halo])
                                continue;
                            ls.push_back(S.cls_slot - cs);
                            sh[S.cls_slot - cs] = S.halo;
                        }
                        if (ls.empty()) continue;
                        PYFOF_HIP(hipMemcpyAsync(d_tl_ls, ls.data(), 4 * ls.size(), hipMemcpyHostToDevice, s));
                        PYFOF_HIP(hipMemcpyAsync(d_tl_sh, sh.data(), 4 * sh.size(), hipMemcpyHostToDevice, s));
                        const uint32_t mc = (uint32_t)(n_in / 2 + 2);
                        const dim3 gz((uint32_t)blocks((uint64_t)mc * mc * mc), (uint32_t)ls.size());
                        k_zero_region<<<gz, kB, 0, s>>>(acc_in[c], d_tl_ls, n_in, mc);
                        PYFOF_LAUNCH_CHECK();
                        uint32_t sbeg = 0, scount = nsrc;
                        if (src_order) {
                            const uint32_t base = c == 0 ? 0u : (uint32_t)tl_nslot[0];
                            const uint32_t r_lo = base + (uint32_t)cs;
                            const uint32_t r_hi = base + (uint32_t)std::min(cs + tl_chunk[c], tl_nslot[c]);
                            sbeg = rank_start[r_lo];
                            scount = rank_start[r_hi] - sbeg;
                        }
                        if (scount > 0) {
                            k_tl_deposit<kTlDepTB, kTlDepBox><<<(scount + kTlDepTB - 1) / kTlDepTB, kTlDepTB, 0, s>>>(src_slot, src_member, src_first, src_wt,
                                                                       src_order, sbeg, scount, d_tlslots, ratio, cs,
                                                                       cs + tl_chunk[c], keep, active, qw, fscale,
                                                                       acc_in[c]);
                            PYFOF_LAUNCH_CHECK();
                        }
                        C.fft_begin();
                        fft_conv3d(*plan_in[c], C.fft_lines_for(n_in), grid_in[c], acc_in[c], d_tl_ls,
                                   (int32_t)ls.size(), d_tl_sh, fscale, Gin[c], s);
                        C.fft_end(n_in, (int)ls.size());
                        k_tl_interp<<<blocks(M), kB, 0, s>>>(tgt_slot, tgt_first, tgt_wt, self_in, d_tlslots, ratio,
                                                             cs, cs + tl_chunk[c], keep, active, qw, grid_in[c], M,
                                                             A.a, phi_in);
                        PYFOF_LAUNCH_CHECK();
                    }
                }
                if (pass == 1) {
                    float4* qs = C.buf("c2qs").as<float4>(M);
                    k_c2_sorted_pos<<<blocks(M), kB, 0, s>>>(qw, keep, clist, M, qs);
                    PYFOF_LAUNCH_CHECK();
                    k_c2_full<<<blocks(M), kB, 0, s>>>(qw, qs, A.mhalo, d_pmof, d_pm, keep, active, cell_of, cstart,
                                                       clist, M, A.a, A.app, pr, phi_c2);
                    PYFOF_LAUNCH_CHECK();
                } else {
                    k_c2_del_flags<<<blocks(M), kB, 0, s>>>(A.mhalo, d_pmof, active, keep, keep_prev, M, dflag);
                    PYFOF_LAUNCH_CHECK();
                    exclusive_scan_u32(dflag, didx, M, dtot, tmp, s);
                    uint32_t nd = 0;
                    PYFOF_HIP(hipMemcpyAsync(&nd, dtot, 4, hipMemcpyDeviceToHost, s));
                    PYFOF_HIP(hipStreamSynchronize(s));
                    if (nd > 0) {
                        k_c2_del_compact<<<blocks(M), kB, 0, s>>>(dflag, didx, cell_of, M, dk, dv);
                        PYFOF_LAUNCH_CHECK();
                        const bool dalt = radix_sort_pairs(dk, dv, dk2, dv2, nd, bits_for(c2cells), tmp, s);
                        const uint32_t* dks = dalt ? dk2 : dk;
                        const uint32_t* dvs = dalt ? dv2 : dv;
                        k_fill32<<<blocks(c2cells + 1), kB, 0, s>>>(dstart, c2cells + 1, 0u);
                        PYFOF_LAUNCH_CHECK();
                        k_run_ends<<<blocks(nd), kB, 0, s>>>(dks, nd, dstart + 1);
                        PYFOF_LAUNCH_CHECK();
                        inclusive_max_scan_u32(dstart + 1, dstart + 1, c2cells, tmp, s);
                        k_c2_update<<<blocks(M), kB, 0, s>>>(qw, A.mhalo, d_pmof, d_pm, keep, active, cell_of,
                                                             dstart, dvs, clist, M, A.a, A.app, pr, phi_c2);
                        PYFOF_LAUNCH_CHECK();
                    }
                }
                k_phi_combine<<<blocks(M), kB, 0, s>>>(A.mhalo, d_pmof, d_pm, keep, active, npb, phi_c1, phi_c2,
                                                       phi_in, M, phi, err);
                PYFOF_LAUNCH_CHECK();
            }
            // ---- large direct halos
            if (any_dl) {
                std::vector<int2> tiles;
                for (int32_t h : act_list) {
                    if (kind[h] != 1) continue;
                    for (int64_t i0 = 0; i0 < hsize(h); i0 += kB) tiles.push_back(make_int2(h, (int)i0));
                }
                d_tiles = C.buf("dltiles").as<int2>(tiles.size());
                PYFOF_HIP(hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, s));
                if (pass == 1) {
                    if (A.direct_f32) k_dl_full<true><<<(uint32_t)tiles.size(), kB, 0, s>>>(d_tiles, A.hoff, qw, phid, P);
                    else k_dl_full<false><<<(uint32_t)tiles.size(), kB, 0, s>>>(d_tiles, A.hoff, qw, phid, P);
                    PYFOF_LAUNCH_CHECK();
                } else {
                    k_dl_del_flags<<<blocks(M), kB, 0, s>>>(A.mhalo, d_kind, active, keep, keep_prev, M, dflag);
                    PYFOF_LAUNCH_CHECK();
                    exclusive_scan_u32(dflag, didx, M, dtot, tmp, s);
                    uint32_t nd = 0;
                    PYFOF_HIP(hipMemcpyAsync(&nd, dtot, 4, hipMemcpyDeviceToHost, s));
                    PYFOF_HIP(hipStreamSynchronize(s));
                    if (nd > 0) {
                        // members are grouped by halo, so the list is too
                        k_dl_del_compact<<<blocks(M), kB, 0, s>>>(dflag, didx, A.mhalo, M, dk, dv);
                        PYFOF_LAUNCH_CHECK();
                        k_fill32<<<blocks(nh + 1), kB, 0, s>>>(hstart, nh + 1, 0u);
                        PYFOF_LAUNCH_CHECK();
                        k_run_ends<<<blocks(nd), kB, 0, s>>>(dk, nd, hstart + 1);
                        PYFOF_LAUNCH_CHECK();
                        inclusive_max_scan_u32(hstart + 1, hstart + 1, nh, tmp, s);
                        if (A.direct_f32)
                            k_dl_update<true><<<(uint32_t)tiles.size(), kB, 0, s>>>(d_tiles, A.hoff, qw, keep, hstart, dv, phid, P);
                        else
                            k_dl_update<false><<<(uint32_t)tiles.size(), kB, 0, s>>>(d_tiles, A.hoff, qw, keep, hstart, dv, phid, P);
                        PYFOF_LAUNCH_CHECK();
                    }
                }
                k_dl_phi<<<blocks(M), kB, 0, s>>>(A.mhalo, d_kind, keep, active, npb, phid, M, phi, err);
                PYFOF_LAUNCH_CHECK();
            }
            k_copy_u8<<<blocks(M), kB, 0, s>>>(keep, keep_prev, M);
            PYFOF_LAUNCH_CHECK();
            if (pass == 1 && phi_rec) {
                k_record_phi<<<blocks(M), kB, 0, s>>>(A.mhalo, active, phi, M, phi_rec);
                PYFOF_LAUNCH_CHECK();
            }
            // removal loop: one workgroup per candidate, except large ones,
            // which get several (k_pm_big, cooperative launches)
            std::vector<int32_t> small_list, big_list;
            for (int32_t h : act_list)
                (big_cap > 0 && hsize(h) >= kBigMin ? big_list : small_list).push_back(h);
            if (!small_list.empty()) {
                PYFOF_HIP(hipMemcpyAsync(d_alist, small_list.data(), 4 * small_list.size(), hipMemcpyHostToDevice, s));
                k_pm_inner<<<(uint32_t)small_list.size(), kPB, 0, s>>>(d_alist, A.hoff, qw, vv, phi, ulist, ukey,
                                                                      keep, status, npb, active, mkeep, pass == 20,
                                                                      err, P);
                PYFOF_LAUNCH_CHECK();
            }
            std::vector<int2> bmap;
            std::vector<int32_t> bh, bnb;
            for (size_t i = 0; i < big_list.size();) {
                // as many candidates per launch as stay resident together
                bmap.clear();
                bh.clear();
                bnb.clear();
                for (; i < big_list.size(); ++i) {
                    const int64_t n = hsize(big_list[i]);
                    const int32_t nbk = (int32_t)big_blocks((uint32_t)n);
                    if (!bh.empty() && (int64_t)bmap.size() + nbk > big_cap) break;
                    for (int32_t b = 0; b < nbk; ++b) bmap.push_back(make_int2((int32_t)bh.size(), b));
                    bh.push_back(big_list[i]);
                    bnb.push_back(nbk);
                }
                int2* d_bmap = C.buf("bigmap").as<int2>(bmap.size());
                int32_t* d_bh = C.buf("bigh").as<int32_t>(bh.size());
                int32_t* d_bnb = C.buf("bignb").as<int32_t>(bnb.size());
                BigScratch* d_scr = C.buf("bigscr").as<BigScratch>(bh.size());
                const uint64_t words = sizeof(BigScratch) / 4 * bh.size();
                k_fill32<<<blocks(words), kB, 0, s>>>(reinterpret_cast<uint32_t*>(d_scr), words, 0u);
                PYFOF_LAUNCH_CHECK();
                PYFOF_HIP(hipMemcpyAsync(d_bmap, bmap.data(), sizeof(int2) * bmap.size(), hipMemcpyHostToDevice, s));
                PYFOF_HIP(hipMemcpyAsync(d_bh, bh.data(), 4 * bh.size(), hipMemcpyHostToDevice, s));
                PYFOF_HIP(hipMemcpyAsync(d_bnb, bnb.data(), 4 * bnb.size(), hipMemcpyHostToDevice, s));
                int32_t last = pass == 20;
                void* kargs[] = {&d_bmap, &d_bh, &d_bnb, &d_scr, &hoff_p, &qw, &vv, &phi,
                                 &ulist, &ukey, &keep, &status, &npb, &active, &mkeep, &last, &err, &P};
                PYFOF_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pm_big),
                                                     dim3((uint32_t)bmap.size()), dim3(kB), kargs, 0, s));
                // the host vectors are read by the copies above: wait before reuse
                PYFOF_HIP(hipStreamSynchronize(s));
            }
            int32_t herr = 0;
            PYFOF_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
            PYFOF_HIP(hipMemcpyAsync(on.data(), active, nh, hipMemcpyDeviceToHost, s));
            PYFOF_HIP(hipStreamSynchronize(s));
            if (herr) return herr;
            std::vector<int32_t> next;
            for (int32_t h : act_list) if (on[h]) next.push_back(h);
            act_list.swap(next);
        }
        return IERR_OK;
    };

    // largest candidates first: workgroups of one launch are scheduled in
    // order, so the long ones start early and the launch ends sooner (the
    // order changes no result: every candidate is solved on its own)
    auto by_size = [&](std::vector<int32_t> v) {
        std::stable_sort(v.begin(), v.end(), [&](int32_t x, int32_t y) { return hsize(x) > hsize(y); });
        return v;
    };
    int32_t* d_dl = C.buf("dlist").as<int32_t>(std::max<size_t>(dsmall.size(), 1));
    if (A.batched) {
        // every halo of a kind in one launch per step
        if (!dsmall.empty()) {
            const std::vector<int32_t> order = by_size(dsmall);
            PYFOF_HIP(hipMemcpyAsync(d_dl, order.data(), 4 * order.size(), hipMemcpyHostToDevice, s));
            direct_small(d_dl, (uint32_t)order.size());
        }
        std::vector<int32_t> it;
        for (uint32_t h = 0; h < nh; ++h) if (kind[h] != 0) it.push_back((int32_t)h);
        it = by_size(it);
        const int32_t e = iterate(it);
        if (e != IERR_OK) return e;
    } else {
        // the plain form: one halo at a time, in candidate order
        if (!dsmall.empty())
            PYFOF_HIP(hipMemcpyAsync(d_dl, dsmall.data(), 4 * dsmall.size(), hipMemcpyHostToDevice, s));
        size_t k = 0;
        for (uint32_t h = 0; h < nh; ++h) {
            if (kind[h] == 0) {
                direct_small(d_dl + k, 1);
                ++k;
            } else {
                const int32_t e = iterate(std::vector<int32_t>{(int32_t)h});
                if (e != IERR_OK) return e;
            }
        }
    }
    res.passes = passes;
    int32_t herr = 0;
    PYFOF_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
    PYFOF_HIP(hipStreamSynchronize(s));
    if (herr) return herr;
    // ---- write-back: valid halos renumbered in candidate order
    uint32_t* vflag = C.buf("vflag").as<uint32_t>(nh);
    uint32_t* vidx = C.buf("vidx").as<uint32_t>(nh);
    uint32_t* vtot = C.buf("vtot").as<uint32_t>(1);
    k_valid_flags<<<blocks(nh), kB, 0, s>>>(status, npb, nh, A.np_min, vflag);
    PYFOF_LAUNCH_CHECK();
    void* stmp = C.buf("vtmp").get(scan_temp_bytes(nh) + 256);
    exclusive_scan_u32(vflag, vidx, nh, vtot, stmp, s);
    k_bound_write<<<blocks(M), kB, 0, s>>>(A.members, A.mhalo, keep, vflag, vidx, M, A.in_halo, A.halo_id, keep_out);
    PYFOF_LAUNCH_CHECK();
    std::vector<int32_t> st_h(nh);
    PYFOF_HIP(hipMemcpyAsync(&res.nvalid, vtot, 4, hipMemcpyDeviceToHost, s));
    PYFOF_HIP(hipMemcpyAsync(st_h.data(), status, 4 * (size_t)nh, hipMemcpyDeviceToHost, s));
    PYFOF_HIP(hipStreamSynchronize(s));
    for (uint32_t h = 0; h < nh; ++h) {
        if (st_h[h] == 0) ++res.n_ok;
        else if (st_h[h] == 1) ++res.n_too_few;
        else ++res.n_max_iter;
    }
    res.n_direct_large = (uint32_t)dlarge.size();
    // bound_fortran.f90:2861: printed for every candidate that ends at the
    // outer iteration cap, whatever the verbosity (candidate numbers from 1)
    if (res.n_max_iter) {
        std::vector<int32_t> npb_h(nh);
        PYFOF_HIP(hipMemcpyAsync(npb_h.data(), npb, 4 * (size_t)nh, hipMemcpyDeviceToHost, s));
        PYFOF_HIP(hipStreamSynchronize(s));
        for (uint32_t h = 0; h < nh; ++h)
            if (st_h[h] == 5)
                printf(" bound_check: MAX_ITER, candidate =%12u retained members =%12d\n", h + 1, npb_h[h]);
    }
    // bound_fortran.f90:1797-1799, 1808-1811
    if (A.verbose >= 1) {
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        printf("[BOUND_PHASE] phase=1 total_sec=%12.4f\n", sec);
        printf("[BOUND_STATUS] OK/TOO_FEW/MAX_ITER=%u %u %u  total_sec=%12.4f\n", res.n_ok, res.n_too_few,
               res.n_max_iter, sec);
    }
    if (res.n_max_iter || A.verbose >= 1) fflush(stdout);
    return IERR_OK;
}

int32_t run_bound(Ctx& C, const BoundArgs& A, uint8_t* keep_out, BoundResult& res) {
    const int32_t e = run_bound_impl(C, A, keep_out, res);
    if (e != IERR_OK && res.err_msg.empty()) {
        // found on the device, where the error word holds the code only
        const char* msg = e == IERR_NONFINITE ? "non-finite potential or energy"
                          : e == IERR_INVALID ? "a member or source node outside its grid, or no removal threshold"
                          : e == IERR_HIP     ? "workgroups of a large candidate did not meet at a barrier"
                                              : "device error";
        bound_fail(A, res, e, msg);
    }
    return e;
}

}  // namespace pyfof
