// Batched mixed-radix Stockham FFT along one axis of n^3 complex grids.
#include <math.h>

#include <vector>

#include "fft.hpp"

namespace pyfof {

namespace {

constexpr int kFftBlock = 256;
constexpr int kLines = 8;  // lines per workgroup
constexpr int kMaxN = 256;

struct AxisDesc {
    // line l of group g: base(g) + c * line_stride + e * elem_stride, c < kLines
    int64_t s0, s1;  // strides of the group decomposition (i0 fastest)
    int32_t c0, c1;  // counts of i0 and i1 (i2 = slot-local batch index)
    int64_t line_stride, elem_stride;
    int32_t x_contig;  // 1 when lines are contiguous (x axis)
};

struct Radices {
    int32_t nrad;
    int32_t r[16];
};

template <bool INV>
__device__ __forceinline__ float2 cmul_tw(float2 a, float2 w) {
    // w is exp(-i theta); the inverse uses its conjugate
    const float wi = INV ? -w.y : w.y;
    return make_float2(a.x * w.x - a.y * wi, a.x * wi + a.y * w.x);
}

template <bool INV>
__device__ __forceinline__ void dft_small(float2* v, int r) {
    if (r == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    } else if (r == 4) {
        const float2 a = v[0], b = v[1], c = v[2], d = v[3];
        const float2 s0 = make_float2(a.x + c.x, a.y + c.y), d0 = make_float2(a.x - c.x, a.y - c.y);
        const float2 s1 = make_float2(b.x + d.x, b.y + d.y), d1 = make_float2(b.x - d.x, b.y - d.y);
        // forward: X1 = d0 - i d1, X3 = d0 + i d1; inverse swaps them
        const float2 mid = INV ? make_float2(-d1.y, d1.x) : make_float2(d1.y, -d1.x);  // -/+ i d1
        v[0] = make_float2(s0.x + s1.x, s0.y + s1.y);
        v[2] = make_float2(s0.x - s1.x, s0.y - s1.y);
        v[1] = make_float2(d0.x + mid.x, d0.y + mid.y);
        v[3] = make_float2(d0.x - mid.x, d0.y - mid.y);
    } else {  // r == 3
        const float2 a = v[0], b = v[1], c = v[2];
        const float c3 = -0.5f;
        const float s3 = INV ? 0.86602540378443864676f : -0.86602540378443864676f;
        const float2 t = make_float2(b.x + c.x, b.y + c.y);
        const float2 u = make_float2(b.x - c.x, b.y - c.y);
        const float2 m = make_float2(a.x + c3 * t.x, a.y + c3 * t.y);
        v[0] = make_float2(a.x + t.x, a.y + t.y);
        // X1 = m + i s3 u, X2 = m - i s3 u
        v[1] = make_float2(m.x - s3 * u.y, m.y + s3 * u.x);
        v[2] = make_float2(m.x + s3 * u.y, m.y - s3 * u.x);
    }
}

// Stockham autosort stages on the kLines lines held in buf[start]; returns the
// buffer holding the result. Every thread of the workgroup must call it.
template <bool INV, int NL = kLines>
__device__ __forceinline__ int stockham(float2 (*buf)[NL * kMaxN], int start, int n,
                                        const Radices& rad, const float2* __restrict__ tw) {
    int cur = start;
    int ns = 1;
    const int tpl = kFftBlock / NL;  // threads per line
    const int c = threadIdx.x / tpl;
    const int lt = threadIdx.x - c * tpl;
    for (int st = 0; st < rad.nrad; ++st) {
        const int r = rad.r[st];
        const int nb = n / r;
        const float2* in = &buf[cur][c * n];
        float2* out = &buf[cur ^ 1][c * n];
        const int twstep = n / (ns * r);
        for (int j = lt; j < nb; j += tpl) {
            const int k = j % ns;
            float2 v[4];
            for (int q = 0; q < r; ++q) {
                float2 a = in[j + q * nb];
                if (q > 0 && k > 0) a = cmul_tw<INV>(a, tw[(k * q * twstep) % n]);
                v[q] = a;
            }
            dft_small<INV>(v, r);
            const int dst = (j / ns) * ns * r + k;
            for (int p = 0; p < r; ++p) out[dst + p * ns] = v[p];
        }
        ns *= r;
        cur ^= 1;
        __syncthreads();
    }
    return cur;
}

template <bool INV>
__global__ void __launch_bounds__(kFftBlock) k_fft_axis(float2* __restrict__ data, int32_t n,
                                                        const int32_t* __restrict__ slots,
                                                        Radices rad, const float2* __restrict__ tw,
                                                        AxisDesc ax) {
    __shared__ float2 buf[2][kLines * kMaxN];
    const int64_t grid = (int64_t)n * n * n;
    const int32_t g = blockIdx.x;
    const int32_t i0 = g % ax.c0;
    const int32_t i1 = (g / ax.c0) % ax.c1;
    const int32_t ib = g / (ax.c0 * ax.c1);
    float2* base = data + (int64_t)slots[ib] * grid + (int64_t)i0 * ax.s0 + (int64_t)i1 * ax.s1;
    const int tid = threadIdx.x;
    const int total = kLines * n;
    // load
    for (int idx = tid; idx < total; idx += kFftBlock) {
        int c, e;
        if (ax.x_contig) { c = idx / n; e = idx - c * n; }
        else { e = idx / kLines; c = idx - e * kLines; }
        buf[0][c * n + e] = base[(int64_t)c * ax.line_stride + (int64_t)e * ax.elem_stride];
    }
    __syncthreads();
    const int cur = stockham<INV>(buf, 0, n, rad, tw);
    // store
    for (int idx = tid; idx < total; idx += kFftBlock) {
        int cc, e;
        if (ax.x_contig) { cc = idx / n; e = idx - cc * n; }
        else { e = idx / kLines; cc = idx - e * kLines; }
        base[(int64_t)cc * ax.line_stride + (int64_t)e * ax.elem_stride] = buf[cur][cc * n + e];
    }
}


// ---------------------------------------------------------------------------
// Compile-time radix plans of the grid sizes in use (the same factor order as
// fft_make_plan: 4s, then 2s, then 3s). With N, the radix and the stage span
// known at compile time, every index division folds to shifts or multiplies.

template <int N> struct RadixPlan;
template <> struct RadixPlan<48> { static constexpr int nrad = 3; static constexpr int r[5] = {4, 4, 3, 1, 1}; };
template <> struct RadixPlan<64> { static constexpr int nrad = 3; static constexpr int r[5] = {4, 4, 4, 1, 1}; };
template <> struct RadixPlan<128> { static constexpr int nrad = 4; static constexpr int r[5] = {4, 4, 4, 2, 1}; };
template <> struct RadixPlan<216> { static constexpr int nrad = 5; static constexpr int r[5] = {4, 2, 3, 3, 3}; };
template <> struct RadixPlan<256> { static constexpr int nrad = 4; static constexpr int r[5] = {4, 4, 4, 4, 1}; };

template <bool INV, int NL, int N, int R, int NS>
__device__ __forceinline__ void stage_ct(const float2* __restrict__ in, float2* __restrict__ out, int lt,
                                         const float2* __restrict__ tw) {
    constexpr int nb = N / R, tpl = kFftBlock / NL, twstep = N / (NS * R);
    for (int j = lt; j < nb; j += tpl) {
        const int k = j % NS;
        float2 v[4];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            float2 a = in[j + q * nb];
            if (q > 0 && k > 0) a = cmul_tw<INV>(a, tw[(k * q * twstep) % N]);
            v[q] = a;
        }
        dft_small<INV>(v, R);
        const int dst = (j / NS) * NS * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) out[dst + q * NS] = v[q];
    }
}

template <bool INV, int NL, int N, int ST, int NS>
__device__ __forceinline__ int stages_ct(float2 (*buf)[NL * N], int cur, const float2* __restrict__ tw) {
    if constexpr (ST == RadixPlan<N>::nrad) {
        return cur;
    } else {
        constexpr int R = RadixPlan<N>::r[ST];
        constexpr int tpl = kFftBlock / NL;
        const int c = threadIdx.x / tpl, lt = threadIdx.x - c * tpl;
        stage_ct<INV, NL, N, R, NS>(&buf[cur][c * N], &buf[cur ^ 1][c * N], lt, tw);
        __syncthreads();
        return stages_ct<INV, NL, N, ST + 1, NS * R>(buf, cur ^ 1, tw);
    }
}

// ---------------------------------------------------------------------------
// Pruned convolution of zero-padded grids. Sources occupy nodes [0, m]^3 with
// m = n/2 + 1 and only those nodes are read back, so each pass touches only
// the lines that can be non-zero or are needed later, and zero-fills the rest
// in LDS:
//   x forward  lines y, z <= m, input x <= m read from the int64 deposit;
//              output kx <= N/2 (the grid is real: X(N - k) = conj X(k))
//   y forward  lines kx <= N/2, z <= m, input y <= m
//   z          lines kx <= N/2: forward, x G(k), inverse; output z <= m
//   y inverse  lines kx <= N/2, z <= m; output y <= m
//   x inverse  lines y, z <= m; input kx <= N/2, the rest by conjugation;
//              output x <= m

enum { kConvXFwd = 0, kConvYFwd = 1, kConvZ = 2, kConvYInv = 3, kConvXInv = 4 };

struct ConvDesc {
    AxisDesc ax;
    int32_t line_lim;  // x passes: lines y > line_lim are skipped (zero)
    int32_t in_lim;    // elements e > in_lim load as zero
    int32_t out_lim;   // elements e > out_lim are not stored
};

// NL lines per workgroup. The z pass runs every listed slot in one
// workgroup: the spectrum G of its lines is read once into registers and
// reused for all slots.
template <int MODE, int NL, int N>
__global__ void __launch_bounds__(kFftBlock) k_fft_conv(float2* __restrict__ data,
                                                        const unsigned long long* __restrict__ acc,
                                                        const int32_t* __restrict__ slots, int32_t nslots,
                                                        const int32_t* __restrict__ slot_halo,
                                                        const double* __restrict__ fscale,
                                                        const float* __restrict__ G,
                                                        const float2* __restrict__ tw, ConvDesc d) {
    __shared__ float2 buf[2][NL * N];  // LDS sized for this grid: more workgroups per CU for small N
    __shared__ float2 tws[N];          // twiddles, read by every stage
    constexpr int total = NL * N;
    constexpr int kPer = (total + kFftBlock - 1) / kFftBlock;  // elements per thread at most
    const AxisDesc& ax = d.ax;
    constexpr int64_t grid = (int64_t)N * N * N;
    const int32_t g = blockIdx.x;
    const int32_t i0 = g % ax.c0;
    const int32_t i1 = (g / ax.c0) % ax.c1;
    // z pass: blockIdx.y picks a group of slots (several groups when the
    // columns alone would not fill the GPU)
    const int32_t zper = (nslots + (int32_t)gridDim.y - 1) / (int32_t)gridDim.y;
    const int32_t ib0 = MODE == kConvZ ? (int32_t)blockIdx.y * zper : g / (ax.c0 * ax.c1);
    const int32_t ib1 = MODE == kConvZ ? min(nslots, ib0 + zper) : ib0 + 1;
    const int64_t goff = (int64_t)i0 * ax.s0 + (int64_t)i1 * ax.s1;
    // first line of the workgroup: y in the x passes, x in the y and z passes
    const int line0 = i0 * NL;
    const int tid = threadIdx.x;
    // visible to the stages after the barrier that follows the first load
    for (int i = tid; i < N; i += kFftBlock) tws[i] = tw[i];
    float gk[kPer];
    if (MODE == kConvZ) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int idx = tid + k * kFftBlock;
            gk[k] = 0.0f;
            if (idx < total) {
                const int e = idx / NL, c = idx - e * NL;
                gk[k] = G[goff + (int64_t)c * ax.line_stride + (int64_t)e * ax.elem_stride];
            }
        }
    }
    // z pass: the next slot's columns are loaded into registers while this
    // slot is transformed (the slot loop is the only loop over memory here)
    float2 pre[MODE == kConvZ ? kPer : 1];
    auto load_z = [&](int32_t ib) {
        const int64_t sb = (int64_t)slots[ib] * grid;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int idx = tid + k * kFftBlock;
            if (idx < total) {
                const int e = idx / NL, c = idx - e * NL;
                // elements above in_lim are the zero padding and lines above
                // line_lim the conjugate half: never stored, so not read
                pre[k] = e <= d.in_lim && line0 + c <= d.line_lim
                             ? data[sb + goff + (int64_t)c * ax.line_stride + (int64_t)e * ax.elem_stride]
                             : make_float2(0.0f, 0.0f);
            }
        }
    };
    if (MODE == kConvZ && ib0 < ib1) load_z(ib0);
    for (int32_t ib = ib0; ib < ib1; ++ib) {
        const int64_t sbase = (int64_t)slots[ib] * grid;
        double inv = 0.0;
        if (MODE == kConvXFwd) inv = 1.0 / fscale[slot_halo[slots[ib]]];
        if (MODE == kConvZ) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int idx = tid + k * kFftBlock;
                if (idx < total) {
                    const int e = idx / NL, c = idx - e * NL;
                    buf[0][c * N + e] = pre[k];
                }
            }
            __syncthreads();
            if (ib + 1 < ib1) load_z(ib + 1);
        } else {
        for (int idx = tid; idx < total; idx += kFftBlock) {
            int c, e;
            if (ax.x_contig) { c = idx / N; e = idx - c * N; }
            else { e = idx / NL; c = idx - e * NL; }
            float2 v = make_float2(0.0f, 0.0f);
            if (e <= d.in_lim && line0 + c <= d.line_lim) {
                const int64_t o = sbase + goff + (int64_t)c * ax.line_stride + (int64_t)e * ax.elem_stride;
                if (MODE == kConvXFwd) v = make_float2((float)((double)(long long)acc[o] * inv), 0.0f);
                else v = data[o];
            }
            buf[0][c * N + e] = v;
        }
        __syncthreads();
        if (MODE == kConvXInv) {
            // the stored half kx <= N/2 of a Hermitian line: X(N - k) = conj X(k)
            for (int idx = tid; idx < total; idx += kFftBlock) {
                const int c = idx / N, e = idx - c * N;
                if (e > N / 2) {
                    const float2 v = buf[0][c * N + N - e];
                    buf[0][c * N + e] = make_float2(v.x, -v.y);
                }
            }
            __syncthreads();
        }
        }
        int cur;
        if (MODE == kConvXFwd || MODE == kConvYFwd) {
            cur = stages_ct<false, NL, N, 0, 1>(buf, 0, tws);
        } else if (MODE == kConvZ) {
            cur = stages_ct<false, NL, N, 0, 1>(buf, 0, tws);
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int idx = tid + k * kFftBlock;
                if (idx < total) {
                    const int e = idx / NL, c = idx - e * NL;
                    float2& v = buf[cur][c * N + e];
                    v = make_float2(v.x * gk[k], v.y * gk[k]);
                }
            }
            __syncthreads();
            cur = stages_ct<true, NL, N, 0, 1>(buf, cur, tws);
        } else {
            cur = stages_ct<true, NL, N, 0, 1>(buf, 0, tws);
        }
        for (int idx = tid; idx < total; idx += kFftBlock) {
            int c, e;
            if (ax.x_contig) { c = idx / N; e = idx - c * N; }
            else { e = idx / NL; c = idx - e * NL; }
            if (e > d.out_lim || line0 + c > d.line_lim) continue;
            data[sbase + goff + (int64_t)c * ax.line_stride + (int64_t)e * ax.elem_stride] = buf[cur][c * N + e];
        }
        __syncthreads();  // buf is reused by the next slot
    }
}

// NL lines per workgroup in every pass (4, 8 or 16; the transform of a line
// does not depend on it, so neither does the result). More lines read longer
// row segments per load and fewer workgroups stay resident (LDS 16 NL N B);
// the best count depends on N and is measured (py_fof_amd.tune).
template <int N, int NL>
void conv_launch(const FftPlan& p, float2* grid, const unsigned long long* acc, const int32_t* slots,
                 int32_t nslots, const int32_t* slot_halo, const double* fscale, const float* G, hipStream_t s) {
    constexpr int NLY = NL;
    constexpr int NLZ = NL;
    constexpr int m = N / 2 + 1, mc = m + 1;
    constexpr int64_t n2 = (int64_t)N * N;
    constexpr int ygroups = (mc + NL - 1) / NL;
    // the grids are real and G is real and even, so every spectrum is
    // Hermitian along x: the y and z passes keep kx <= N/2 only and the x
    // inverse rebuilds the other half
    constexpr int hx = N / 2;
    constexpr int xgroups_y = (hx + 1 + NLY - 1) / NLY, xgroups_z = (hx + 1 + NLZ - 1) / NLZ;
    const AxisDesc ax_x{(int64_t)NL * N, n2, ygroups, mc, N, 1, 1};
    const AxisDesc ax_y{NLY, n2, xgroups_y, mc, 1, N, 0};
    const AxisDesc ax_z{NLZ, N, xgroups_z, N, 1, n2, 0};
    // {axis, line_lim, in_lim, out_lim}
    const ConvDesc dxf{ax_x, m, m, hx}, dxi{ax_x, m, hx, m};
    const ConvDesc dyf{ax_y, hx, m, N - 1}, dyi{ax_y, hx, N - 1, m}, dz{ax_z, hx, m, m};
    const uint32_t gx = (uint32_t)(ygroups * mc) * (uint32_t)nslots;
    const uint32_t gy = (uint32_t)(xgroups_y * mc) * (uint32_t)nslots;
    const uint32_t gzz = (uint32_t)(xgroups_z * N);
    // about 2048 workgroups in the z pass: slot groups when the columns are few
    const uint32_t zgroups = (uint32_t)std::max(1, std::min(nslots, (int32_t)((2048 + gzz - 1) / gzz)));
    k_fft_conv<kConvXFwd, NL, N><<<gx, kFftBlock, 0, s>>>(grid, acc, slots, nslots, slot_halo, fscale, G, p.tw, dxf);
    PYFOF_LAUNCH_CHECK();
    k_fft_conv<kConvYFwd, NLY, N><<<gy, kFftBlock, 0, s>>>(grid, acc, slots, nslots, slot_halo, fscale, G, p.tw, dyf);
    PYFOF_LAUNCH_CHECK();
    k_fft_conv<kConvZ, NLZ, N><<<dim3(gzz, zgroups), kFftBlock, 0, s>>>(grid, acc, slots, nslots, slot_halo, fscale, G, p.tw, dz);
    PYFOF_LAUNCH_CHECK();
    k_fft_conv<kConvYInv, NLY, N><<<gy, kFftBlock, 0, s>>>(grid, acc, slots, nslots, slot_halo, fscale, G, p.tw, dyi);
    PYFOF_LAUNCH_CHECK();
    k_fft_conv<kConvXInv, NL, N><<<gx, kFftBlock, 0, s>>>(grid, acc, slots, nslots, slot_halo, fscale, G, p.tw, dxi);
    PYFOF_LAUNCH_CHECK();
}

}  // namespace

void fft_make_plan(FftPlan& p, int n, float2* tw_dev, hipStream_t s) {
    if (n > kMaxN || n % kLines) throw HipError(hipErrorInvalidValue, "unsupported FFT size");
    p.n = n;
    p.nrad = 0;
    int m = n;
    while (m % 4 == 0) { p.radix[p.nrad++] = 4; m /= 4; }
    while (m % 2 == 0) { p.radix[p.nrad++] = 2; m /= 2; }
    while (m % 3 == 0) { p.radix[p.nrad++] = 3; m /= 3; }
    if (m != 1) throw HipError(hipErrorInvalidValue, "FFT size must be 2^a 3^b");
    std::vector<float2> t(n);
    for (int k = 0; k < n; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)n;
        t[k] = make_float2((float)cos(a), (float)sin(a));
    }
    PYFOF_HIP(hipMemcpyAsync(tw_dev, t.data(), sizeof(float2) * n, hipMemcpyHostToDevice, s));
    PYFOF_HIP(hipStreamSynchronize(s));
    p.tw = tw_dev;
}

void fft3d(const FftPlan& p, float2* data, const int32_t* slots, int32_t nslots, bool inverse,
           hipStream_t s) {
    if (nslots <= 0) return;
    const int n = p.n;
    Radices rad;
    rad.nrad = p.nrad;
    for (int i = 0; i < 16; ++i) rad.r[i] = p.radix[i];
    const int64_t n2 = (int64_t)n * n;
    AxisDesc ax[3];
    // x: lines (z, y) contiguous; groups of 8 consecutive y
    ax[0] = AxisDesc{(int64_t)kLines * n, n2, n / kLines, n, n, 1, 1};
    // y: lines (z, x), elements strided by n; groups of 8 consecutive x
    ax[1] = AxisDesc{kLines, n2, n / kLines, n, 1, n, 0};
    // z: lines (y, x), elements strided by n^2; groups of 8 consecutive x
    ax[2] = AxisDesc{kLines, n, n / kLines, n, 1, n2, 0};
    const uint32_t groups = (uint32_t)((n / kLines) * n) * (uint32_t)nslots;
    for (int a = 0; a < 3; ++a) {
        if (inverse)
            k_fft_axis<true><<<groups, kFftBlock, 0, s>>>(data, n, slots, rad, p.tw, ax[a]);
        else
            k_fft_axis<false><<<groups, kFftBlock, 0, s>>>(data, n, slots, rad, p.tw, ax[a]);
        PYFOF_LAUNCH_CHECK();
    }
}

namespace {
template <int N>
void conv_lines(const FftPlan& p, int lines, float2* grid, const unsigned long long* acc,
                const int32_t* slots, int32_t nslots, const int32_t* slot_halo, const double* fscale,
                const float* G, hipStream_t s) {
    switch (lines) {
        case 4: conv_launch<N, 4>(p, grid, acc, slots, nslots, slot_halo, fscale, G, s); break;
        case 8: conv_launch<N, 8>(p, grid, acc, slots, nslots, slot_halo, fscale, G, s); break;
        case 16: conv_launch<N, 16>(p, grid, acc, slots, nslots, slot_halo, fscale, G, s); break;
        default: throw HipError(hipErrorInvalidValue, "FFT lines per workgroup must be 4, 8 or 16");
    }
}
}  // namespace

bool fft_conv_size_ok(int n) { return n == 48 || n == 64 || n == 128 || n == 216 || n == 256; }

void fft_conv3d(const FftPlan& p, int lines, float2* grid, const unsigned long long* acc,
                const int32_t* slots, int32_t nslots, const int32_t* slot_halo, const double* fscale,
                const float* G, hipStream_t s) {
    if (nslots <= 0) return;
    switch (p.n) {
        case 48: conv_lines<48>(p, lines, grid, acc, slots, nslots, slot_halo, fscale, G, s); break;
        case 64: conv_lines<64>(p, lines, grid, acc, slots, nslots, slot_halo, fscale, G, s); break;
        case 128: conv_lines<128>(p, lines, grid, acc, slots, nslots, slot_halo, fscale, G, s); break;
        case 216: conv_lines<216>(p, lines, grid, acc, slots, nslots, slot_halo, fscale, G, s); break;
        case 256: conv_lines<256>(p, lines, grid, acc, slots, nslots, slot_halo, fscale, G, s); break;
        default: throw HipError(hipErrorInvalidValue, "no convolution kernel for this grid size");
    }
}

}  // namespace pyfof
