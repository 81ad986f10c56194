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
    // stages
    int cur = 0;
    int ns = 1;
    const int tpl = kFftBlock / kLines;  // threads per line
    const int c = tid / tpl;
    const int lt = tid - c * tpl;
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
    // store
    for (int idx = tid; idx < total; idx += kFftBlock) {
        int cc, e;
        if (ax.x_contig) { cc = idx / n; e = idx - cc * n; }
        else { e = idx / kLines; cc = idx - e * kLines; }
        base[(int64_t)cc * ax.line_stride + (int64_t)e * ax.elem_stride] = buf[cur][cc * n + e];
    }
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

}  // namespace pyfof
