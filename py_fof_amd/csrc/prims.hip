// Device-wide exclusive scan and stable LSD radix sort (keys + values, u32).
//
// The scans work on tiles of 4096 elements per 256-thread workgroup (4 waves
// of 64), the sort on tiles of 8192 keys per 512-thread workgroup. The radix
// sort is the classic reduce-then-scan form: a per-tile digit histogram, one
// device scan of the digit-major [digit][tile] table, then a scatter in
// which each tile ranks its items stably: each wave ranks its own contiguous
// share with ballots and a per-wave running count, and one barrier combines
// the waves in order.
#include "prims.hpp"

#include <algorithm>

namespace pyfof {

namespace {

constexpr int kBlock = 256;
constexpr int kIpt = 16;
constexpr int kTile = kBlock * kIpt;
// Digits of at most 8 bits. A sort of `bits` key bits makes the fewest
// passes of 8-bit digits, ceil(bits / 8), with the digit width spread evenly
// over them (28 bits: four 7-bit passes): fewer digits per pass mean shorter
// histograms and longer output runs of one digit per tile.
constexpr int kMaxRadixBits = 8;
// The radix sort's tiles: 8192 keys per 512-thread workgroup, so that a
// tile's output runs of one digit are 32 keys (128 B) long on average.
constexpr int kSBlock = 512;
constexpr int kSTile = kSBlock * kIpt;

struct OpSum {
    static __device__ __host__ uint32_t id() { return 0u; }
    static __device__ uint32_t f(uint32_t a, uint32_t b) { return a + b; }
};
struct OpMax {
    static __device__ __host__ uint32_t id() { return 0u; }
    static __device__ uint32_t f(uint32_t a, uint32_t b) { return a > b ? a : b; }
};

// Scan input: the array itself.
struct InPlain {
    const uint32_t* p;
    __device__ uint32_t operator()(uint64_t i) const { return p[i]; }
};

template <class Op>
__device__ inline uint32_t wave_reduce(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = Op::f(v, __shfl_down(v, o, kWave));
    return v;
}

template <class Op>
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = __lane_id();
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= o) v = Op::f(v, t);
    }
    return v;
}

template <class Op, class In>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(const In in, uint64_t n,
                                                        uint32_t* __restrict__ partial) {
    __shared__ uint32_t wsum[kBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    uint32_t s = Op::id();
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        if (i < n) s = Op::f(s, in(i));
    }
    s = wave_reduce<Op>(s);
    if (__lane_id() == 0) wsum[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = Op::id();
        for (int w = 0; w < kBlock / kWave; ++w) t = Op::f(t, wsum[w]);
        partial[blockIdx.x] = t;
    }
}

// LDS index with one pad word per 16 to keep the blocked reads conflict-free.
__device__ inline int pad_idx(int e) { return e + (e >> 4); }

// Scans one tile per block. offsets[block] (exclusive prefix of the tiles
// before it) is folded in; inclusive selects the inclusive form.
template <class Op, class In>
__global__ void __launch_bounds__(kBlock) k_scan_down(const In in, uint32_t* out, uint64_t n,
                                                      const uint32_t* __restrict__ offsets,
                                                      uint32_t* __restrict__ total, int inclusive) {
    __shared__ uint32_t tile[kTile + kTile / 16];
    __shared__ uint32_t wtot[kBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        int e = j * kBlock + t;
        uint64_t i = base + e;
        tile[pad_idx(e)] = i < n ? in(i) : Op::id();
    }
    __syncthreads();
    uint32_t loc[kIpt];
    uint32_t s = Op::id();
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        const uint32_t v = tile[pad_idx(t * kIpt + j)];
        loc[j] = inclusive ? Op::f(s, v) : s;
        s = Op::f(s, v);
    }
    const uint32_t incl = wave_inclusive_scan<Op>(s);
    const uint32_t excl_in_wave = __shfl_up(incl, 1, kWave);
    const int w = t / kWave;
    if (__lane_id() == kWave - 1) wtot[w] = incl;
    __syncthreads();
    uint32_t woff = offsets ? offsets[blockIdx.x] : Op::id();
    for (int k = 0; k < w; ++k) woff = Op::f(woff, wtot[k]);
    const uint32_t texcl = __lane_id() == 0 ? woff : Op::f(woff, excl_in_wave);
#pragma unroll
    for (int j = 0; j < kIpt; ++j) tile[pad_idx(t * kIpt + j)] = Op::f(texcl, loc[j]);
    if (total && blockIdx.x == gridDim.x - 1 && t == kBlock - 1) *total = Op::f(texcl, s);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        int e = j * kBlock + t;
        uint64_t i = base + e;
        if (i < n) out[i] = tile[pad_idx(e)];
    }
}

template <int kRadixBits>
__global__ void __launch_bounds__(kSBlock) k_radix_hist(const uint32_t* __restrict__ keys, uint64_t n,
                                                       int shift, uint32_t* __restrict__ hist,
                                                       uint32_t ntiles) {
    constexpr int kRadix = 1 << kRadixBits;
    __shared__ uint32_t h[kSBlock / kWave][kRadix];
    uint32_t* hf = &h[0][0];
    for (int i = threadIdx.x; i < (kSBlock / kWave) * kRadix; i += kSBlock) hf[i] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSTile;
    const int w = threadIdx.x / kWave;
    if (base + kSTile <= n && ((uintptr_t)(keys + base) & 15u) == 0) {
        // a full tile: four keys per 16-byte load, all loads issued first
        const uint4* k4 = reinterpret_cast<const uint4*>(keys + base);
        uint4 v[kIpt / 4];
#pragma unroll
        for (int j = 0; j < kIpt / 4; ++j) v[j] = k4[j * kSBlock + threadIdx.x];
#pragma unroll
        for (int j = 0; j < kIpt / 4; ++j) {
            atomicAdd(&h[w][(v[j].x >> shift) & (kRadix - 1)], 1u);
            atomicAdd(&h[w][(v[j].y >> shift) & (kRadix - 1)], 1u);
            atomicAdd(&h[w][(v[j].z >> shift) & (kRadix - 1)], 1u);
            atomicAdd(&h[w][(v[j].w >> shift) & (kRadix - 1)], 1u);
        }
    } else {
#pragma unroll 4
        for (int j = 0; j < kIpt; ++j) {
            uint64_t i = base + (uint64_t)j * kSBlock + threadIdx.x;
            if (i < n) atomicAdd(&h[w][(keys[i] >> shift) & (kRadix - 1)], 1u);
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kSBlock) {
        uint32_t c = 0;
        for (int k = 0; k < kSBlock / kWave; ++k) c += h[k][d];
        hist[(uint64_t)d * ntiles + blockIdx.x] = c;
    }
}

// Each wave ranks a contiguous eighth of the tile, 64 items per round: lanes
// with equal digits are matched by one ballot per digit bit, and a running
// count per digit in the wave's own LDS row gives each item its rank among
// the wave's earlier items (the rows need no barrier: a wave's lanes run in
// lockstep). One barrier then turns the eight rows into prefixes in wave
// order, so the rank of an item follows (wave, round, lane) = item order:
// the sort is stable. The ranked tile is then laid out in LDS in digit order
// and written from there, so that neighbouring lanes store neighbouring
// addresses (runs of one digit) instead of scattering lane by lane.
template <int kRadixBits>
__global__ void __launch_bounds__(kSBlock) k_radix_scatter(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals,
                                                           uint32_t* __restrict__ okeys,
                                                           uint32_t* __restrict__ ovals, uint64_t n,
                                                           int shift,
                                                           const uint32_t* __restrict__ hist,
                                                           uint32_t ntiles) {
    constexpr int kRadix = 1 << kRadixBits;
    static_assert(kSBlock >= kRadix, "one thread per digit in the digit scans");
    constexpr int kWaves = kSBlock / kWave;
    constexpr int kRounds = kSTile / kSBlock;  // rounds of 64 items per wave
    __shared__ uint32_t wc[kWaves][kRadix];  // per-wave counts, then per-wave offsets in a digit
    __shared__ uint32_t lstart[kRadix];      // first tile-local position of each digit
    __shared__ uint32_t gstart[kRadix];      // first output position of each digit of this tile
    __shared__ uint32_t wsum[kRadix > kWave ? kRadix / kWave : 1];
    __shared__ uint32_t sk[kSTile], sv[kSTile];
    const int t = threadIdx.x, w = t / kWave;
    uint32_t* wf = &wc[0][0];
    for (int i = t; i < kWaves * kRadix; i += kSBlock) wf[i] = 0;
    __syncthreads();
    // consecutive tiles on one XCD: the two ends of a digit's run of output
    // positions are written by neighbouring tiles, whose partial lines then
    // meet in the same L2
    const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t tile0 = (uint64_t)tile * kSTile;
    const uint64_t base = tile0 + (uint64_t)w * (kRounds * kWave);
    const uint64_t lt = lanemask_lt();
    uint32_t kk[kRounds], vv[kRounds], rk[kRounds];
#pragma unroll
    for (int j = 0; j < kRounds; ++j) {
        const uint64_t idx = base + (uint64_t)j * kWave + __lane_id();
        const bool valid = idx < n;
        const uint32_t k = valid ? keys[idx] : 0u;
        const uint32_t d = (k >> shift) & (kRadix - 1);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bb = 0; bb < kRadixBits; ++bb) {
            const bool bit = (d >> bb) & 1u;
            const uint64_t bal = __ballot(valid && bit);
            m &= bit ? bal : ~bal;
        }
        const uint32_t r = __popcll(m & lt);
        const uint32_t before = wc[w][d];
        if (valid && r == 0) wc[w][d] = before + (uint32_t)__popcll(m);
        kk[j] = k;
        vv[j] = valid ? (vals ? vals[idx] : (uint32_t)idx) : 0u;  // no vals: the input index
        rk[j] = before + r;
    }
    __syncthreads();
    {
        // threads 0 .. kRadix - 1: one digit each (waves 0 .. kRadix / kWave - 1)
        uint32_t tot = 0;
        if (t < kRadix) {
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < kWaves; ++k) {
                const uint32_t c = wc[k][t];
                wc[k][t] = acc;
                acc += c;
            }
            tot = acc;
            gstart[t] = hist[(uint64_t)t * ntiles + tile];
        }
        const uint32_t inc = wave_inclusive_scan<OpSum>(tot);
        if (t < kRadix && __lane_id() == kWave - 1) wsum[w] = inc;
        __syncthreads();
        if (t < kRadix) {
            uint32_t off = 0;
            for (int k = 0; k < w; ++k) off += wsum[k];
            lstart[t] = off + inc - tot;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRounds; ++j) {
        const uint64_t idx = base + (uint64_t)j * kWave + __lane_id();
        if (idx >= n) continue;
        const uint32_t d = (kk[j] >> shift) & (kRadix - 1);
        const uint32_t lp = lstart[d] + wc[w][d] + rk[j];
        sk[lp] = kk[j];
        sv[lp] = vv[j];
    }
    __syncthreads();
    const uint64_t rem = n - tile0;
    const uint32_t cnt = rem < (uint64_t)kSTile ? (uint32_t)rem : (uint32_t)kSTile;
#pragma unroll
    for (int j = 0; j < kRounds; ++j) {
        const uint32_t lp = (uint32_t)(j * kSBlock + t);
        if (lp >= cnt) continue;
        const uint32_t k = sk[lp];
        const uint32_t d = (k >> shift) & (kRadix - 1);
        const uint32_t dst = gstart[d] + (lp - lstart[d]);
        okeys[dst] = k;
        ovals[dst] = sv[lp];
    }
}

size_t scan_levels_bytes(uint64_t n) {
    size_t b = 0;
    while (n > (uint64_t)kTile) {
        n = div_up(n, kTile);
        b += ((n + 63) / 64) * 64 * sizeof(uint32_t);
    }
    return b + 256;
}

template <class Op, class In>
void scan_rec(const In in, uint32_t* out, uint64_t n, uint32_t* total, uint32_t* temp,
              int inclusive, hipStream_t s) {
    if (n == 0) return;
    const uint32_t nb = div_up(n, kTile);
    if (nb == 1) {
        k_scan_down<Op, In><<<1, kBlock, 0, s>>>(in, out, n, nullptr, total, inclusive);
        PYFOF_LAUNCH_CHECK();
        return;
    }
    uint32_t* partial = temp;
    uint32_t* next = temp + ((nb + 63) / 64) * 64;
    k_scan_reduce<Op, In><<<nb, kBlock, 0, s>>>(in, n, partial);
    PYFOF_LAUNCH_CHECK();
    scan_rec<Op>(InPlain{partial}, partial, nb, nullptr, next, 0, s);
    k_scan_down<Op, In><<<nb, kBlock, 0, s>>>(in, out, n, partial, total, inclusive);
    PYFOF_LAUNCH_CHECK();
}

}  // namespace

int bits_for(uint64_t max_key) {
    int b = 0;
    while (b < 64 && (max_key >> b) != 0) ++b;
    return b;
}

size_t scan_temp_bytes(uint64_t n) { return scan_levels_bytes(n); }

size_t radix_temp_bytes(uint64_t n) {
    const uint64_t ntiles = n == 0 ? 1 : div_up(n, kSTile);
    const uint64_t m = ntiles << kMaxRadixBits;
    return ((m * sizeof(uint32_t) + 255) / 256) * 256 + scan_levels_bytes(m);
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* temp,
                        hipStream_t s) {
    scan_rec<OpSum>(InPlain{in}, out, n, total, static_cast<uint32_t*>(temp), 0, s);
}

void inclusive_max_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* temp,
                            hipStream_t s) {
    scan_rec<OpMax>(InPlain{in}, out, n, nullptr, static_cast<uint32_t*>(temp), 1, s);
}

namespace {

__global__ void __launch_bounds__(kBlock) k_iota_u32(uint32_t* __restrict__ out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = (uint32_t)i;
}

}  // namespace

static_assert(radix_tile_keys == (uint32_t)kSTile, "tile of the radix passes");

int radix_first_width(int bits) {
    if (bits <= 0) return 0;
    const int passes = (bits + kMaxRadixBits - 1) / kMaxRadixBits;
    // digits narrower than 4 bits only occur in a single pass (bits <= 3),
    // where the wider digit's extra bits are zero in every key
    return std::max(4, (bits + passes - 1) / passes);
}

uint32_t* radix_first_hist(void* temp) { return static_cast<uint32_t*>(temp); }

bool radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* alt_keys, uint32_t* alt_vals,
                      uint64_t n, int bits, void* temp, hipStream_t s, bool vals_iota,
                      bool first_hist_ready) {
    if (n <= 1 || bits <= 0) {
        if (vals_iota && n > 0) {
            k_iota_u32<<<div_up(n, kBlock), kBlock, 0, s>>>(vals, n);
            PYFOF_LAUNCH_CHECK();
        }
        return false;
    }
    const uint32_t ntiles = div_up(n, kSTile);
    const uint64_t m_max = (uint64_t)ntiles << kMaxRadixBits;
    uint32_t* hist = static_cast<uint32_t*>(temp);
    uint32_t* stemp = hist + ((m_max * sizeof(uint32_t) + 255) / 256) * 256 / sizeof(uint32_t);
    // with vals_iota the first pass reads no values: each pair's value is
    // its input index
    uint32_t *ik = keys, *iv = vals_iota ? nullptr : vals, *ok = alt_keys, *ov = alt_vals;
    bool in_alt = false;
    const int width = radix_first_width(bits);
    for (int shift = 0; shift < bits; shift += width) {
        const uint64_t m = (uint64_t)ntiles << width;
        switch (width) {
#define PYFOF_RADIX_PASS(W)                                                                          \
    case W:                                                                                          \
        if (shift > 0 || !first_hist_ready) {                                                        \
            k_radix_hist<W><<<ntiles, kSBlock, 0, s>>>(ik, n, shift, hist, ntiles);                  \
            PYFOF_LAUNCH_CHECK();                                                                    \
        }                                                                                            \
        scan_rec<OpSum>(InPlain{hist}, hist, m, nullptr, stemp, 0, s);                               \
        k_radix_scatter<W><<<ntiles, kSBlock, 0, s>>>(ik, iv, ok, ov, n, shift, hist, ntiles);       \
        PYFOF_LAUNCH_CHECK();                                                                        \
        break;
            PYFOF_RADIX_PASS(4)
            PYFOF_RADIX_PASS(5)
            PYFOF_RADIX_PASS(6)
            PYFOF_RADIX_PASS(7)
            PYFOF_RADIX_PASS(8)
#undef PYFOF_RADIX_PASS
            default:
                throw std::runtime_error("radix_sort_pairs: digit width out of range");
        }
        uint32_t* tk = ik; ik = ok; ok = tk;
        uint32_t* tv = iv ? iv : vals; iv = ov; ov = tv;
        in_alt = !in_alt;
    }
    return in_alt;
}

}  // namespace pyfof
