// Friends-of-Friends on MI355X: cell binning, linking, relabel and catalog.
//
// Pipeline of one call (on the context's streams):
//   k_bbox         per-region bounding box of the wrapped coordinates
//   (host)         reference cell geometry per region (fof_fortran.f90:518-570)
//   k_keys         reference cell id, float coordinates and face bits of every
//                  particle, and the sort's first digit histogram
//   radix sort     stable sort by cell id = the reference's CRS order
//   k_cell_starts  cell_ptr from the sorted keys (beside k_gather)
//   k_gather       float CRS coordinates (the reference's x_crs) and the
//                  cells' face masks (second stream)
//   k_dense_*      cell classes; sub-cells (side <= (b - m)/sqrt 3) inside
//                  dense cells
//   linking        k_link (sparse cells), k_cell_intra / k_cell_pair (mid
//                  cells), k_dense_intra / k_dense_cell_cross / k_dense_cross
//                  (dense cells), k_link_dense: reference predicate +
//                  lock-free union-find, on three streams
//   k_label_count / k_cand_bits / k_write   roots, sizes, candidate
//                  numbering, outputs
//   catalog        halo members grouped by halo id (fof_with_info, ang. mom.)
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pyfof.h"
#include "common.hpp"
#include "fof_geom.hpp"
#include "fof_kernels.hpp"
#include "prims.hpp"
#include "ctx.hpp"
#include "bound.hpp"

namespace pyfof {

constexpr int kBlk = 256;

// ---------------------------------------------------------------------------
// small fills
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_iota_u32(uint32_t* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// Face bits of a coordinate x (the f32 CRS coordinate widened) in a cell
// [lo, lo + cell) along one axis: bit 0 when x lies within R = b + m of the
// lower face, bit 1 within R of the upper face, both with the rounding
// margin 2m. A partner across a face can only be linked from a member that
// has its bit (a linked pair is within b + float rounding < R per axis, and
// the cell coordinates are monotone in x). Computed once per particle in
// k_keys, with the same f64 expressions the linking kernels would use.
__device__ __forceinline__ uint32_t face_bits(double x, double lo, double cell, double R, double m) {
    return (x - R <= lo + 2.0 * m ? 1u : 0u) | (x + R >= lo + cell - 2.0 * m ? 2u : 0u);
}
constexpr uint32_t kEdgeBit = 64u;  // beside the six face bits: the cell lies on the grid's boundary

// ---------------------------------------------------------------------------
// per-region bounding box of x_rel (fof_fortran.f90:524-531)
__global__ void __launch_bounds__(kBlk) k_bbox(const double* __restrict__ xp,
                                               const Chunk* __restrict__ chunks,
                                               const int32_t* __restrict__ reg_start,
                                               unsigned long long* __restrict__ bb, double L,
                                               double Lh) {
    const Chunk ch = chunks[blockIdx.x];
    const size_t a = (size_t)reg_start[ch.region];
    const double xa0 = xp[3 * a], xa1 = xp[3 * a + 1], xa2 = xp[3 * a + 2];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    // kBboxRows rows per thread loaded before any is used (rel_coord's rare
    // fmod branch would otherwise keep each load waiting for the previous
    // row's arithmetic); a row past the chunk's end repeats the thread's
    // first row, which leaves the box unchanged
    constexpr int kBboxRows = 8;
    for (int32_t i0 = ch.begin + threadIdx.x; i0 < ch.end; i0 += kBboxRows * kBlk) {
        double v[kBboxRows][3];
#pragma unroll
        for (int u = 0; u < kBboxRows; ++u) {
            const int32_t i = i0 + u * kBlk;
            const size_t j = 3 * (size_t)(i < ch.end ? i : i0);
            v[u][0] = xp[j];
            v[u][1] = xp[j + 1];
            v[u][2] = xp[j + 2];
        }
#pragma unroll
        for (int u = 0; u < kBboxRows; ++u) {
            const double v0 = rel_coord(v[u][0], xa0, L, Lh);
            const double v1 = rel_coord(v[u][1], xa1, L, Lh);
            const double v2 = rel_coord(v[u][2], xa2, L, Lh);
            mn[0] = fmin(mn[0], v0); mx[0] = fmax(mx[0], v0);
            mn[1] = fmin(mn[1], v1); mx[1] = fmax(mx[1], v1);
            mn[2] = fmin(mn[2], v2); mx[2] = fmax(mx[2], v2);
        }
    }
    __shared__ double red[6][kBlk / kWave];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[d] = fmin(mn[d], __shfl_down(mn[d], o, kWave));
            mx[d] = fmax(mx[d], __shfl_down(mx[d], o, kWave));
        }
    }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0) {
        for (int d = 0; d < 3; ++d) { red[d][w] = mn[d]; red[3 + d][w] = mx[d]; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        double v = red[d][0];
        for (int k = 1; k < kBlk / kWave; ++k) v = d < 3 ? fmin(v, red[d][k]) : fmax(v, red[d][k]);
        unsigned long long* dst = bb + 6 * (size_t)ch.region + d;
        // most chunks do not extend the box: read first, so that the
        // thousands of workgroups of a region do not all queue atomics on
        // the same six words
        const unsigned long long u = dbl_to_ordered(v);
        const unsigned long long cur = __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (d < 3) {
            if (u < cur) atomicMin(dst, u);
        } else if (u > cur) {
            atomicMax(dst, u);
        }
    }
}

// k_keys' rows of a chunk of a region with cells: cell id, float coordinates
// and face bits of each particle; count(i, key) sees every key written.
template <class Count>
__device__ __forceinline__ void keys_rows(const double* __restrict__ xp, const Chunk& ch, const RegionDev& g,
                                          uint32_t* __restrict__ keys, float4* __restrict__ xr, double L,
                                          double Lh, double R, double m, double xa0, double xa1, double xa2,
                                          Count& count) {
    const uint32_t off = g.cell_off;
    const double o0 = g.origin[0], o1 = g.origin[1], o2 = g.origin[2];
    const double c0 = g.cell[0], c1 = g.cell[1], c2 = g.cell[2];
    const int32_t n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    // kKeyRows rows per thread loaded before any is used (as in k_bbox); a
    // row past the chunk's end is loaded from the thread's first row and
    // not stored
    constexpr int kKeyRows = 4;
    for (int32_t i0 = ch.begin + threadIdx.x; i0 < ch.end; i0 += kKeyRows * kBlk) {
        double xv[kKeyRows][3];
#pragma unroll
        for (int u = 0; u < kKeyRows; ++u) {
            const int32_t i = i0 + u * kBlk;
            const size_t j = 3 * (size_t)(i < ch.end ? i : i0);
            xv[u][0] = xp[j];
            xv[u][1] = xp[j + 1];
            xv[u][2] = xp[j + 2];
        }
#pragma unroll
        for (int u = 0; u < kKeyRows; ++u) {
            const int32_t i = i0 + u * kBlk;
            if (i >= ch.end) break;
            const double v0 = rel_coord(xv[u][0], xa0, L, Lh);
            const double v1 = rel_coord(xv[u][1], xa1, L, Lh);
            const double v2 = rel_coord(xv[u][2], xa2, L, Lh);
            const int32_t ix = ref_cell_coord(v0, o0, c0, n0);
            const int32_t iy = ref_cell_coord(v1, o1, c1, n1);
            const int32_t iz = ref_cell_coord(v2, o2, c2, n2);
            const uint32_t key = off + (uint32_t)(ix - 1) + (uint32_t)n0 * ((uint32_t)(iy - 1) + (uint32_t)n1 * (uint32_t)(iz - 1));
            keys[i] = key;
            count(i, key);
            const float f0 = (float)v0, f1 = (float)v1, f2 = (float)v2;
            // bit 6: the cell lies on the grid's boundary (a neighbour may wrap,
            // or not exist)
            const bool edge = ix == 1 || ix == n0 || iy == 1 || iy == n1 || iz == 1 || iz == n2;
            const uint32_t fb = face_bits((double)f0, o0 + (double)(ix - 1) * c0, c0, R, m) |
                                (face_bits((double)f1, o1 + (double)(iy - 1) * c1, c1, R, m) << 2) |
                                (face_bits((double)f2, o2 + (double)(iz - 1) * c2, c2, R, m) << 4) |
                                (edge ? kEdgeBit : 0u);
            xr[i] = make_float4(f0, f1, f2, __uint_as_float(fb));
        }
    }
}

// reference cell id per particle (fof_fortran.f90:580-585, 0-based global id)
// and, in input order, its x_crs = real(x_rel, 4) (fof_fortran.f90:608): the
// one coalesced read of xp the binning makes after the bounding box; k_gather
// then moves 16 B per particle instead of re-reading the 24 B of xp at random.
// The fourth component holds the particle's face bits (face_bits) as an
// integer; the linking kernels find a particle's region from its cell.
// With hist, the kernel also counts the keys' low digits (key & dmask) per
// tile of the radix sort's first pass (radix_first_hist's layout, entries
// zeroed before): a chunk (kKeyChunk positions) spans at most two tiles,
// counted in LDS and added once per non-zero count, so the sort reads the
// keys once less.
constexpr int32_t kKeyChunk = 4096;
static_assert(kKeyChunk <= (int32_t)radix_tile_keys, "a chunk spans at most two sort tiles");
__global__ void __launch_bounds__(kBlk) k_keys(const double* __restrict__ xp,
                                               const Chunk* __restrict__ chunks,
                                               const RegionDev* __restrict__ reg,
                                               uint32_t* __restrict__ keys,
                                               float4* __restrict__ xr, double L, double Lh, double R,
                                               double m, uint32_t* __restrict__ hist, uint32_t ntiles,
                                               uint32_t dmask) {
    __shared__ uint32_t th[2][256];
    const Chunk ch = chunks[blockIdx.x];
    const RegionDev& g = reg[ch.region];
    const uint32_t off = g.cell_off;
    const size_t a = (size_t)g.start;
    const double xa0 = xp[3 * a], xa1 = xp[3 * a + 1], xa2 = xp[3 * a + 2];
    const uint32_t t0 = (uint32_t)ch.begin / radix_tile_keys;
    if (hist) {
        for (int e = threadIdx.x; e < 2 * 256; e += kBlk) (&th[0][0])[e] = 0u;
        __syncthreads();
    }
    auto count = [&](int32_t i, uint32_t key) {
        if (hist) atomicAdd(&th[(uint32_t)i / radix_tile_keys - t0][key & dmask], 1u);
    };
    if (g.skip) {
        for (int32_t i = ch.begin + threadIdx.x; i < ch.end; i += kBlk) {
            keys[i] = off;
            count(i, off);
            xr[i] = make_float4((float)rel_coord(xp[3 * (size_t)i], xa0, L, Lh),
                                (float)rel_coord(xp[3 * (size_t)i + 1], xa1, L, Lh),
                                (float)rel_coord(xp[3 * (size_t)i + 2], xa2, L, Lh), 0.f);
        }
    } else {
        keys_rows(xp, ch, g, keys, xr, L, Lh, R, m, xa0, xa1, xa2, count);
    }
    if (hist) {
        __syncthreads();
        const uint32_t nd = dmask + 1u;
        for (uint32_t e = threadIdx.x; e < 2u * nd; e += kBlk) {
            const uint32_t tt = e / nd, d = e - tt * nd;
            const uint32_t c = th[tt][d];
            if (c) atomicAdd(&hist[(size_t)d * ntiles + t0 + tt], c);
        }
    }
}
// ends[c] = 1 + last position of cell c (0 for empty cells, zero-filled before)
__global__ void k_run_ends(const uint32_t* __restrict__ skeys, uint32_t n, uint32_t* __restrict__ ends) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = skeys[p];
    if (p + 1 == n || skeys[p + 1] != k) ends[k] = p + 1;
}

// cell_ptr[c] = the first CRS position whose cell is c or later, for
// c = 0..ncell (cell_ptr[ncell] = n), straight from the sorted keys:
// position p (0 <= p <= n) owns the entries key[p-1] + 1 .. key[p] (from 0
// for p = 0, up to ncell for p = n), which are p. Each lane reads four
// consecutive positions (one 16-byte load), so a wave's 256 positions own
// one contiguous range of entries; up to kWaveSpan of them are written by
// all lanes in turn, each entry's owner found by a binary search over the
// lanes' last entries and then among the owner lane's four. (With one
// position per lane the kernel had four times the waves and, beside the
// gather, finished 2 ms later, holding back k_dense_mark.) In a wave with a
// longer range each lane writes the runs of its positions, up to kGapInline
// entries each, and a longer run (empty space) is cut into
// records of kGapChunk entries that k_fill_gaps writes with whole
// workgroups. Records needed: at most (ncell + 1) / kGapInline.
constexpr uint32_t kGapInline = 64, kGapChunk = 4096, kWaveSpan = 1024;

__host__ __device__ inline uint64_t gap_records(uint64_t ncell) { return (ncell + 1) / kGapInline + 2; }

__global__ void k_cell_starts(const uint32_t* __restrict__ skeys, uint32_t n, uint64_t ncell,
                              uint32_t* __restrict__ cell_ptr, uint4* __restrict__ gaps,
                              uint32_t* __restrict__ ngaps) {
    // four consecutive positions per lane, 256 per wave
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = __lane_id();
    const uint64_t q0 = 4 * t;
    const uint64_t p0 = q0 - 4 * (uint64_t)lane;  // the wave's first position
    if (p0 > n) return;  // a wave past position n owns nothing
    // last entry owned by each position; positions past n own nothing
    // (ncell again); ncell < 2^32 - 1, so every entry index fits 32 bits
    uint32_t h[4];
    if (q0 + 3 < n) {
        const uint4 v = *reinterpret_cast<const uint4*>(skeys + q0);
        h[0] = v.x;
        h[1] = v.y;
        h[2] = v.z;
        h[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) h[k] = q0 + k < n ? skeys[q0 + k] : (uint32_t)ncell;
    }
    const uint64_t first = p0 == 0 ? 0 : (uint64_t)skeys[p0 - 1] + 1;  // wave-uniform
    const uint64_t last = (uint64_t)__shfl(h[3], kWave - 1, kWave);
    if (first > last) return;  // the whole wave is in one cell
    if (last - first + 1 <= kWaveSpan) {
        // a wave-uniform trip count: the shuffles read every lane's values,
        // which a lane that had left the loop would no longer provide
        for (uint64_t base = first; base <= last; base += kWave) {
            const uint64_t e = base + (uint64_t)lane;
            // owner: the first position whose last entry is e or later; the
            // lane first (by its fourth position), then within the lane
            int j = 0;
            for (int step = kWave / 2; step > 0; step >>= 1) {
                const uint32_t hv = (uint32_t)__shfl(h[3], j + step - 1, kWave);
                if ((uint64_t)hv < e) j += step;
            }
            const uint32_t a0 = (uint32_t)__shfl(h[0], j, kWave);
            const uint32_t a1 = (uint32_t)__shfl(h[1], j, kWave);
            const uint32_t a2 = (uint32_t)__shfl(h[2], j, kWave);
            const uint32_t k = (uint64_t)a0 >= e ? 0u : ((uint64_t)a1 >= e ? 1u : ((uint64_t)a2 >= e ? 2u : 3u));
            if (e <= last) cell_ptr[e] = (uint32_t)(p0 + 4 * (uint64_t)j + k);
        }
        return;
    }
    // the lane's positions in turn: the entry before each is the previous
    // position's last (the previous lane's fourth, for the first)
    const uint32_t up = (uint32_t)__shfl_up(h[3], 1, kWave);
    uint64_t prev_hi = lane == 0 ? first - 1 : (uint64_t)up;  // first - 1 wraps only for p0 == 0
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const uint64_t p = q0 + (uint64_t)kk;
        const uint64_t lo = p == 0 ? 0 : prev_hi + 1;
        const uint64_t hi = h[kk];
        prev_hi = hi;
        if (p > n || lo > hi) continue;  // past n; or the same cell as position p - 1
        const uint64_t len = hi - lo + 1;
        if (len <= kGapInline) {
            for (uint64_t c = lo; c <= hi; ++c) cell_ptr[c] = (uint32_t)p;
            continue;
        }
        const uint32_t nch = (uint32_t)((len + kGapChunk - 1) / kGapChunk);
        const uint32_t at = atomicAdd(ngaps, nch);
        for (uint32_t k = 0; k < nch; ++k) {
            const uint64_t a = lo + (uint64_t)k * kGapChunk;
            const uint64_t b = a + kGapChunk < hi + 1 ? a + kGapChunk : hi + 1;
            gaps[at + k] = make_uint4((uint32_t)a, (uint32_t)b, (uint32_t)p, 0u);
        }
    }
}

__global__ void __launch_bounds__(kBlk) k_fill_gaps(const uint4* __restrict__ gaps,
                                                    const uint32_t* __restrict__ ngaps,
                                                    uint32_t* __restrict__ cell_ptr) {
    const uint32_t m = *ngaps;
    for (uint32_t g = blockIdx.x; g < m; g += gridDim.x) {
        const uint4 r = gaps[g];
        for (uint32_t c = r.x + threadIdx.x; c < r.y; c += kBlk) cell_ptr[c] = r.z;
    }
}

// x_crs in CRS order (fof_fortran.f90:608): the input-order copy permuted;
// the initial forest parent[p] = p; and per non-empty cell the OR of its
// members' six face bits (fmask; bit 2d: a member lies within R of the
// cell's lower face of axis d, bit 2d + 1: of the upper face), so that a
// neighbour cell with no member near the shared face can be skipped
// unread. The wave holding a cell's first position writes its mask: the OR
// over the wave's positions of that cell, or every bit for a cell that goes
// on past the wave's end (a conservative mask).
__global__ void __launch_bounds__(kBlk) k_gather(const float4* __restrict__ xr,
                                                 const uint32_t* __restrict__ perm,
                                                 const uint32_t* __restrict__ skeys,
                                                 float4* __restrict__ xc, int32_t* __restrict__ parent,
                                                 uint8_t* __restrict__ fmask, uint32_t n) {
    const uint32_t p = blockIdx.x * kBlk + threadIdx.x;
    const int lane = __lane_id();
    const bool valid = p < n;
    uint32_t key = 0xffffffffu, fb = 0;
    if (valid) {
        const float4 x = xr[perm[p]];
        key = skeys[p];
        xc[p] = x;
        parent[p] = (int32_t)p;
        fb = __float_as_uint(x.w) & 63u;
    }
    // does a run of this cell start at p (the previous position is another
    // cell)? (every lane takes part in the shuffle: a lane masked off would
    // read as zero to the others)
    const uint32_t kup = __shfl_up(key, 1, kWave);
    const uint32_t kprev = lane == 0 ? (valid && p > 0 ? skeys[p - 1] : 0xffffffffu) : kup;
    uint32_t head = valid && (p == 0 || kprev != key) ? 1u : 0u;
    // segmented inclusive OR over the runs of equal keys (a run's lanes are
    // consecutive), with whether the run's first position is in this wave
    uint32_t acc = fb;
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t ku = __shfl_up(key, off, kWave), au = __shfl_up(acc, off, kWave);
        const uint32_t hu = __shfl_up(head, off, kWave);
        if (lane >= off && ku == key) {
            acc |= au;
            head |= hu;
        }
    }
    const uint32_t knext = __shfl_down(key, 1, kWave);
    const bool tail = lane == kWave - 1 || knext != key;
    if (!valid || !tail || !head) return;
    if (lane == kWave - 1 && p + 1 < n && skeys[p + 1] == key) acc = 63u;  // the cell goes on
    fmask[key] = (uint8_t)acc;
}

// marks cells with at least kDenseMin particles in regions with sub-cells
// aux[c]: the dense index of a dense cell (>= 0); kAuxMid for a cell whose
// own pairs k_cell_intra resolves; kAuxNearDense for any other cell with a
// dense neighbour in a clique region (k_mark_near_dense); -1 otherwise.
// Negative values read as "not dense" everywhere.
constexpr int32_t kAuxMid = -2;
constexpr int32_t kAuxNearDense = -3;
constexpr uint32_t kMidMin = 7, kMidMax = 255;  // particles of a k_cell_intra cell
constexpr int kMarkBlk = 1024;                    // threads of k_dense_mark

constexpr int kMarkPer = 4;                       // consecutive cells per thread

__global__ void __launch_bounds__(kMarkBlk) k_dense_mark(const uint32_t* __restrict__ cell_ptr, uint32_t ncell,
                             const uint32_t* __restrict__ cell_off, const RegionDev* __restrict__ reg,
                             int32_t nreg, int32_t* __restrict__ aux, uint32_t* __restrict__ dense_cells,
                             uint32_t* __restrict__ ndense, uint32_t* __restrict__ mid_cells,
                             uint32_t* __restrict__ nmid) {
    const uint64_t c0 = ((uint64_t)blockIdx.x * kMarkBlk + threadIdx.x) * kMarkPer;
    uint32_t dbits = 0, mbits = 0;  // bit k: cell c0 + k is dense / mid
    if (c0 < ncell) {
        uint32_t prev = cell_ptr[c0];
        for (int k = 0; k < kMarkPer && c0 + k < ncell; ++k) {
            const uint32_t c = (uint32_t)(c0 + k);
            const uint32_t next = cell_ptr[c + 1];
            const uint32_t cnt = next - prev;
            prev = next;
            if (cnt >= kMidMin) {
                const RegionDev& g = reg[nreg == 1 ? 0u : region_of_cell(cell_off, nreg, c)];
                const bool dense = !g.skip && cnt >= (uint32_t)kDenseMin && g.prune && g.k[0] * g.k[1] * g.k[2] > 1;
                const bool mid = !dense && !g.skip && g.prune && cnt <= kMidMax;
                dbits |= (uint32_t)dense << k;
                mbits |= (uint32_t)mid << k;
            }
        }
    }
    // one counter update per workgroup and list: the threads' counts are
    // scanned within each wave, the waves' totals in LDS
    const uint32_t nd = (uint32_t)__popc(dbits), nm = (uint32_t)__popc(mbits);
    uint32_t sd = nd, sm = nm;
    const int lane = __lane_id();
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t td = __shfl_up(sd, o, kWave), tm = __shfl_up(sm, o, kWave);
        if (lane >= o) {
            sd += td;
            sm += tm;
        }
    }
    __shared__ uint32_t wd[kMarkBlk / kWave], wm[kMarkBlk / kWave], bases[2];
    const int w = threadIdx.x / kWave;
    if (lane == kWave - 1) {
        wd[w] = sd;
        wm[w] = sm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t td = 0, tm = 0;
        for (int k = 0; k < kMarkBlk / kWave; ++k) {
            const uint32_t a = wd[k], b = wm[k];
            wd[k] = td;
            wm[k] = tm;
            td += a;
            tm += b;
        }
        bases[0] = td ? atomicAdd(ndense, td) : 0u;
        bases[1] = tm ? atomicAdd(nmid, tm) : 0u;
    }
    __syncthreads();
    if (c0 >= ncell) return;
    uint32_t bd = bases[0] + wd[w] + sd - nd, bm = bases[1] + wm[w] + sm - nm;
    for (int k = 0; k < kMarkPer && c0 + k < ncell; ++k) {
        const uint32_t c = (uint32_t)(c0 + k);
        int32_t v = -1;
        if ((dbits >> k) & 1u) {
            v = (int32_t)bd;
            dense_cells[bd++] = c;
        } else if ((mbits >> k) & 1u) {
            v = kAuxMid;
            mid_cells[bm++] = c;
        }
        aux[c] = v;
    }
}

__device__ inline void cell_lower_corner(const RegionDev& g, uint32_t lc, double lo[3], int32_t ijk[3]) {
    const uint32_t n0 = (uint32_t)g.n[0], n1 = (uint32_t)g.n[1];
    ijk[0] = (int32_t)(lc % n0);
    ijk[1] = (int32_t)((lc / n0) % n1);
    ijk[2] = (int32_t)(lc / (n0 * n1));
    for (int d = 0; d < 3; ++d) lo[d] = g.origin[d] + (double)ijk[d] * g.cell[d];
}

// Sub-cell lists of the dense cells: per dense cell a counting sort of its
// positions by sub-cell (sub_ptr: offsets relative to the cell start,
// sub_list: positions) and, when sub-cells are cliques, the pre-union of
// each sub-cell onto its smallest position.
__global__ void __launch_bounds__(kBlk) k_dense_build(
    const uint32_t* __restrict__ dense_cells, const uint32_t* __restrict__ cell_ptr,
    const uint32_t* __restrict__ cell_off, const RegionDev* __restrict__ reg, int32_t nreg,
    const float4* __restrict__ xc, uint32_t* __restrict__ sub_ptr, int32_t sub_stride,
    uint32_t* __restrict__ sub_list, int32_t* __restrict__ parent, uint4* __restrict__ subs,
    uint32_t* __restrict__ nsubs) {
    // per sub-cell counts (then offsets) and smallest positions: sub_stride
    // (the largest sub-cell count of any region, plus one) words each
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    uint32_t* cnt = reinterpret_cast<uint32_t*>(dyn_lds);
    uint32_t* minp = cnt + sub_stride;
    __shared__ uint32_t nonempty, base;
    const uint32_t c = dense_cells[blockIdx.x];
    const uint32_t r = nreg == 1 ? 0u : region_of_cell(cell_off, nreg, c);
    const RegionDev& g = reg[r];
    double lo[3];
    int32_t ijk[3];
    cell_lower_corner(g, c - g.cell_off, lo, ijk);
    const int32_t k0 = g.k[0], k1 = g.k[1], k2 = g.k[2];
    const int32_t K = k0 * k1 * k2;
    const uint32_t s = cell_ptr[c], e = cell_ptr[c + 1];
    for (int i = threadIdx.x; i < K; i += kBlk) { cnt[i] = 0; minp[i] = 0xffffffffu; }
    if (threadIdx.x == 0) nonempty = 0;
    __syncthreads();
    for (uint32_t p = s + threadIdx.x; p < e; p += kBlk) {
        const float4 v = xc[p];
        const int32_t sx = sub_coord((double)v.x, lo[0], g.sub[0], k0);
        const int32_t sy = sub_coord((double)v.y, lo[1], g.sub[1], k1);
        const int32_t sz = sub_coord((double)v.z, lo[2], g.sub[2], k2);
        const int32_t si = sx + k0 * (sy + k1 * sz);
        if (atomicAdd(&cnt[si], 1u) == 0u) atomicAdd(&nonempty, 1u);
        atomicMin(&minp[si], p);
    }
    __syncthreads();
    // the block's slots in the list of non-empty clique sub-cells
    if (threadIdx.x == 0) base = g.clique ? atomicAdd(nsubs, nonempty) : 0u;
    // exclusive scan of cnt[0..K) by one wave per 64-entry stripe, serially over stripes
    if (threadIdx.x < kWave) {
        uint32_t carry = 0;
        for (int b0 = 0; b0 < K; b0 += kWave) {
            const int i = b0 + threadIdx.x;
            const uint32_t v = i < K ? cnt[i] : 0u;
            uint32_t incl = v;
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o, kWave);
                if ((int)__lane_id() >= o) incl += t;
            }
            if (i < K) cnt[i] = carry + incl - v;
            carry += __shfl(incl, kWave - 1, kWave);
        }
        if (threadIdx.x == 0) cnt[K] = carry;
    }
    __syncthreads();
    uint32_t* row = sub_ptr + (size_t)blockIdx.x * sub_stride;
    for (int i = threadIdx.x; i <= K; i += kBlk) row[i] = cnt[i];
    __syncthreads();
    if (threadIdx.x == 0) nonempty = 0;
    __syncthreads();
    if (g.clique) {
        // one entry per non-empty sub-cell (any order)
        for (int i = threadIdx.x; i < K; i += kBlk)
            if (cnt[i + 1] > cnt[i])
                subs[base + atomicAdd(&nonempty, 1u)] = make_uint4(blockIdx.x, (uint32_t)i, minp[i], r);
    }
    __syncthreads();
    for (uint32_t p = s + threadIdx.x; p < e; p += kBlk) {
        const float4 v = xc[p];
        const int32_t sx = sub_coord((double)v.x, lo[0], g.sub[0], k0);
        const int32_t sy = sub_coord((double)v.y, lo[1], g.sub[1], k1);
        const int32_t sz = sub_coord((double)v.z, lo[2], g.sub[2], k2);
        const int32_t si = sx + k0 * (sy + k1 * sz);
        const uint32_t slot = atomicAdd(&cnt[si], 1u);
        sub_list[s + slot] = p;
        if (g.clique) parent[p] = (int32_t)minp[si];
    }
}

// Per-thread linking context.
struct LinkCtx {
    int32_t* parent;
    int32_t rp;  // last known root of the thread's particle
    uint32_t dmask;  // neighbour offsets (bit o) whose dense cells are left to k_link_dense
    __device__ __forceinline__ void link(int32_t q) {
        const int32_t rq = dsu_find_near(parent, q);
        if (rq != rp) rp = dsu_unite(parent, rp, rq);
    }
};

template <bool PERIODIC>
__device__ __forceinline__ bool pair_linked(const float4& a, const float4& q, const LinkParams& P) {
    if (PERIODIC) return pred_periodic(a, q, P.L, P.Lh, P.rp2);
    return pred_flat(a, q, P.rp2);
}

// pair_linked with an f32 screen for the periodic predicate. Its coordinate
// differences are the f32 differences pred_periodic widens; the f32 sum of
// their squares is within 2e-7 (relative) of the exact sum, so outside
// b^2 (1 -+ 1e-6) it decides as the f64 test does. Near the threshold, or
// where a difference may wrap (>= Lh / 2), the f64 test decides.
template <bool PERIODIC>
__device__ __forceinline__ bool pair_linked_screened(const float4& a, const float4& q, const LinkParams& P) {
    if (PERIODIC) {
        const float fx = fabsf(a.x - q.x), fy = fabsf(a.y - q.y), fz = fabsf(a.z - q.z);
        if (fmaxf(fx, fmaxf(fy, fz)) < P.scr_wrap) {
            const float s = __fadd_rn(__fadd_rn(__fmul_rn(fx, fx), __fmul_rn(fy, fy)), __fmul_rn(fz, fz));
            if (s < P.scr_lo) return true;
            if (s > P.scr_hi) return false;
        }
    }
    return pair_linked<PERIODIC>(a, q, P);
}

// Tests the particle `me` against the members [qq, e) of one cell. Two
// partners at a time: their positions, and then the parents of the linked
// ones, are loaded together; a partner whose parent is already the thread's
// known root needs no union (pairs beat fours here: the wider batch spills
// registers at this kernel's occupancy target).
template <bool PERIODIC>
__device__ __forceinline__ void scan_members(const float4& me, uint32_t qq, uint32_t e, const LinkParams& P,
                                             const float4* __restrict__ xc, LinkCtx& L) {
    for (; qq + 2 <= e; qq += 2) {
        const float4 y0 = xc[qq], y1 = xc[qq + 1];
        const bool h0 = pair_linked<PERIODIC>(me, y0, P), h1 = pair_linked<PERIODIC>(me, y1, P);
        if (!(h0 || h1)) continue;
        const int32_t q0 = (int32_t)qq;
        const int32_t a0 = h0 ? load_near(&L.parent[q0]) : 0;
        const int32_t a1 = h1 ? load_near(&L.parent[q0 + 1]) : 0;
        if (h0 && a0 != L.rp) L.link(q0);
        if (h1 && a1 != L.rp) L.link(q0 + 1);
    }
    for (; qq < e; ++qq) {
        if (pair_linked<PERIODIC>(me, xc[qq], P)) L.link((int32_t)qq);
    }
}

// Tests the particle `me` (position p) against cell jc, reached by neighbour
// offset o (13: its own cell). For the own cell only partners q > p are
// tested. The caller visits only the neighbours the query box can reach
// (link_particle); every pair is decided by the predicate, so a cell visited
// needlessly costs time only. A cell with sub-cells (dense) is not scanned
// here: bit o of the thread's mask is set and k_link_dense scans it, which
// keeps the register footprint of this kernel, and so its occupancy, low. A
// neighbouring mid cell is skipped: k_cell_pair takes its pairs with this
// cell.
template <bool PERIODIC>
__device__ __forceinline__ void visit_cell(const float4& me, uint32_t p, uint32_t jc, int o, bool own,
                                           bool back, const LinkParams& P,
                                           const uint32_t* __restrict__ cell_ptr,
                                           const int32_t* __restrict__ aux,
                                           const float4* __restrict__ xc, LinkCtx& L) {
    // class and member range loaded together, before any branch on them
    const int32_t cls = aux[jc];
    const uint32_t s = cell_ptr[jc], e = cell_ptr[jc + 1];
    if (!own && cls == kAuxMid) return;
    if (back && cls < 0) return;  // a backward cell is visited for its sub-cells only
    if (s >= e) return;
    if (cls >= 0) {
        L.dmask |= 1u << o;
        return;
    }
    scan_members<PERIODIC>(me, own ? p + 1 : s, e, P, xc, L);
}

// Neighbour cell of (ijk + off) under the reference's rule (fof_fortran.f90:
// 729-755 periodic, 684-704 non-periodic): wrap only along periodic axes,
// keep jcell > icell. Returns false if the neighbour does not exist. shift
// receives the periodic image offset that brings this particle next to it.
__device__ __forceinline__ bool neighbour_cell(const RegionDev& g, const int32_t ijk[3], int o,
                                               int32_t jijk[3], double shift[3], double L) {
    const int32_t off[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
    for (int d = 0; d < 3; ++d) {
        int32_t j = ijk[d] + off[d];
        shift[d] = 0.0;
        if (j < 0 || j >= g.n[d]) {
            if (!g.periodic[d]) return false;
            if (j < 0) { j += g.n[d]; shift[d] = L; } else { j -= g.n[d]; shift[d] = -L; }
        }
        jijk[d] = j;
    }
    return true;
}

__device__ __forceinline__ uint32_t local_cell(const RegionDev& g, const int32_t jijk[3]) {
    return (uint32_t)jijk[0] + (uint32_t)g.n[0] * ((uint32_t)jijk[1] + (uint32_t)g.n[1] * (uint32_t)jijk[2]);
}

// Marks the plain cells (aux -1) next to a dense cell of a clique region as
// kAuxNearDense: only their particles need k_link's visits of backward
// neighbour cells, which look for dense cells alone (k_link_dense takes
// those pairs from the sparse side). Every writer stores the same value.
__global__ void __launch_bounds__(kBlk) k_mark_near_dense(const uint32_t* __restrict__ dense_cells, uint32_t ndense,
                                                          const uint32_t* __restrict__ cell_off,
                                                          const RegionDev* __restrict__ reg, int32_t nreg,
                                                          int32_t* __restrict__ aux, double L) {
    const uint32_t d = blockIdx.x * kBlk + threadIdx.x;
    if (d >= ndense) return;
    const uint32_t c = dense_cells[d];
    const RegionDev& g = reg[nreg == 1 ? 0u : region_of_cell(cell_off, nreg, c)];
    if (!g.clique) return;
    double lo[3];
    int32_t ijk[3];
    cell_lower_corner(g, c - g.cell_off, lo, ijk);
    for (int o = 0; o < 27; ++o) {
        if (o == 13) continue;
        int32_t jijk[3];
        double sh[3];
        if (!neighbour_cell(g, ijk, o, jijk, sh, L)) continue;
        const uint32_t jc = g.cell_off + local_cell(g, jijk);
        if (aux[jc] == -1) aux[jc] = kAuxNearDense;
    }
}

template <bool PERIODIC>
__device__ __forceinline__ uint32_t link_particle(uint32_t p, const float4& me, const RegionDev& g,
                                                  uint32_t c, const LinkParams& P,
                                                  const uint32_t* __restrict__ cell_ptr,
                                                  const int32_t* __restrict__ aux,
                                                  const uint8_t* __restrict__ fmask,
                                                  const float4* __restrict__ xc, int32_t* parent) {
    // a mid cell's pairs are all k_cell_intra's and k_cell_pair's
    const int32_t own_cls = aux[c];
    if (own_cls == kAuxMid) return 0u;
    // backward neighbours are visited only for their dense cells, which a
    // cell not marked kAuxNearDense does not have (clique regions only)
    const bool near_dense = g.clique && own_cls == kAuxNearDense;
    LinkCtx L{parent, (int32_t)p, 0u};  // rp becomes a root at the first unite
    const uint32_t lc = c - g.cell_off;
    // (the face bits and edge bit k_keys stored in .w)
    const uint32_t fb = __float_as_uint(me.w);
    if (own_cls < 0 && g.prune && !(fb & kEdgeBit) && (fb & 3u) != 3u && (fb & 12u) != 12u && (fb & 48u) != 48u) {
        // Inside the grid every neighbour exists and none wraps: the one at
        // offset (ox, oy, oz) is cell lc + ox + n0 (oy + n1 oz), forward
        // exactly when the offset is lexicographically positive. With the
        // box reaching at most one side per axis (the face bits), the cells
        // it reaches are the seven subsets of the axes it crosses; their
        // classes and member ranges are all loaded before any is scanned,
        // so the loads of one particle overlap instead of queueing. (A dense
        // own cell, in a region without cliques, takes the general path.)
        const int32_t s1 = g.n[0], s2 = g.n[0] * g.n[1];
        const int32_t dx = (fb & 1u) ? -1 : ((fb & 2u) ? 1 : 0);
        const int32_t dy = (fb & 4u) ? -1 : ((fb & 8u) ? 1 : 0);
        const int32_t dz = (fb & 16u) ? -1 : ((fb & 32u) ? 1 : 0);
        int32_t cl[7];
        uint32_t sb[7], eb[7], fm[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int ax = k + 1;  // axes crossed: bit 0 x, bit 1 y, bit 2 z
            const int32_t ox = (ax & 1) ? dx : 0, oy = (ax & 2) ? dy : 0, oz = (ax & 4) ? dz : 0;
            const bool reach = ((ax & 1) == 0 || dx != 0) && ((ax & 2) == 0 || dy != 0) && ((ax & 4) == 0 || dz != 0);
            const int o = (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1);
            cl[k] = -1;
            sb[k] = eb[k] = fm[k] = 0;
            if (reach && (o > 13 || near_dense)) {
                const uint32_t jc = g.cell_off + (uint32_t)((int32_t)lc + ox + s1 * oy + s2 * oz);
                cl[k] = aux[jc];
                sb[k] = cell_ptr[jc];
                eb[k] = cell_ptr[jc + 1];
                fm[k] = fmask[jc];
            }
        }
        scan_members<PERIODIC>(me, p + 1, cell_ptr[c + 1], P, xc, L);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int ax = k + 1;
            const int32_t ox = (ax & 1) ? dx : 0, oy = (ax & 2) ? dy : 0, oz = (ax & 4) ? dz : 0;
            const int o = (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1);
            const int32_t cls = cl[k];
            if (sb[k] >= eb[k] || cls == kAuxMid) continue;  // empty or not reached; mid: k_cell_pair
            if (cls >= 0) {
                L.dmask |= 1u << o;  // dense: k_link_dense
                continue;
            }
            if (o < 13) continue;  // a backward cell is visited for its sub-cells only
            // a partner lies within R of the neighbour's faces toward this
            // cell on every axis crossed; skip a cell with no such member
            const uint32_t need = ((ax & 1) ? (dx < 0 ? 2u : 1u) : 0u) | ((ax & 2) ? (dy < 0 ? 8u : 4u) : 0u) |
                                  ((ax & 4) ? (dz < 0 ? 32u : 16u) : 0u);
            if ((fm[k] & need) != need) continue;
            scan_members<PERIODIC>(me, sb[k], eb[k], P, xc, L);
        }
        return L.dmask;
    }
    visit_cell<PERIODIC>(me, p, c, 13, true, false, P, cell_ptr, aux, xc, L);
    int32_t ijk[3];
    double lo_unused[3];
    cell_lower_corner(g, lc, lo_unused, ijk);
    // a cell off the periodic edges has no wrapped neighbours: offsets below
    // 13 (lexicographically negative) are then exactly the backward cells
    bool edge = false;
    for (int d = 0; d < 3; ++d) edge = edge || (g.periodic[d] && (ijk[d] == 0 || ijk[d] == g.n[d] - 1));
    auto neighbour = [&](int o) {
        if (o < 13 && !near_dense && !edge) return;
        int32_t jijk[3];
        double sh[3];
        if (!neighbour_cell(g, ijk, o, jijk, sh, P.L)) return;
        const uint32_t jl = local_cell(g, jijk);
        // sparse-sparse pairs once (forward cells); every pair with a dense
        // clique cell is taken here, from the sparse side, in both directions
        const bool fwd = jl > lc;
        if (!fwd && !near_dense) return;
        if (!g.prune) {
            // tiny periodic grids: the same neighbour can be reached twice
            for (int o2 = 0; o2 < o; ++o2) {
                if (o2 == 13) continue;
                int32_t j2[3];
                double s2[3];
                if (neighbour_cell(g, ijk, o2, j2, s2, P.L) && local_cell(g, j2) == jl) return;
            }
        }
        visit_cell<PERIODIC>(me, p, g.cell_off + jl, o, false, !fwd, P, cell_ptr, aux, xc, L);
    };
    if (!g.prune) {
        for (int o = 0; o < 27; ++o)
            if (o != 13) neighbour(o);
        return L.dmask;
    }
    // Offsets along each axis whose neighbour the query box [x - R, x + R] can
    // reach: a pair passing the predicate is within b (+ float rounding < m)
    // per axis, so its partner lies inside the box with R = b + m, and the
    // cell coordinates are monotone in x; the margin 2m covers the f64
    // rounding of the cell corners. So only about (1 + 2R / cell)^3 - 1
    // neighbours are visited instead of 26. A neighbour across a periodic
    // edge is always visited.
    int32_t olo[3], ohi[3];
    for (int d = 0; d < 3; ++d) {
        const bool wrap_lo = g.periodic[d] && ijk[d] == 0;
        const bool wrap_hi = g.periodic[d] && ijk[d] == g.n[d] - 1;
        olo[d] = (wrap_lo || ((fb >> (2 * d)) & 1u)) ? -1 : 0;
        ohi[d] = (wrap_hi || ((fb >> (2 * d)) & 2u)) ? 1 : 0;
    }
    for (int oz = olo[2]; oz <= ohi[2]; ++oz)
        for (int oy = olo[1]; oy <= ohi[1]; ++oy)
            for (int ox = olo[0]; ox <= ohi[0]; ++ox)
                if (ox != 0 || oy != 0 || oz != 0) neighbour((ox + 1) + 3 * (oy + 1) + 9 * (oz + 1));
    return L.dmask;
}

// One thread per particle outside dense clique cells: pairs with sparse cells
// (own cell and forward neighbours), and the mask of dense cells to scan,
// written for every particle (0 when none). Every thread also clears its
// entry of the relabel's component sizes and halo ids (zero0, zero1), which
// nothing in the linking reads: two coalesced stores here cost less than two
// fill kernels beside the linking (0.13 ms per call on the 134M set).
// Occupancy target (5 waves per SIMD) chosen by measurement on the
// 134M-particle clump set: the kernel waits on chains of dependent loads and
// DSU atomics, but the face-mask loads of its neighbour cells need 92
// registers, and spilling them (at 6 waves) wrote scratch lines to HBM and
// cost more than the sixth wave gained.
__global__ void __launch_bounds__(kBlk, 5) k_link(const float4* __restrict__ xc,
                                               const uint32_t* __restrict__ skeys,
                                               const uint32_t* __restrict__ cell_ptr,
                                               const int32_t* __restrict__ aux,
                                               const uint8_t* __restrict__ fmask,
                                               const RegionDev* __restrict__ reg,
                                               const uint32_t* __restrict__ cell_off, int32_t* parent,
                                               uint32_t* __restrict__ dmask, uint32_t* __restrict__ zero0,
                                               int32_t* __restrict__ zero1, LinkParams P) {
    const uint32_t p = xcd_block(blockIdx.x, gridDim.x) * kBlk + threadIdx.x;
    if (p >= P.N) return;
    const float4 me = xc[p];
    const uint32_t c = skeys[p];
    const int32_t rid = P.nregions == 1 ? 0 : (int32_t)region_of_cell(cell_off, P.nregions, c);
    uint32_t mask = 0;
    // One pass per distinct region in the wave (almost always one): the
    // region record is then wave-uniform and read with scalar loads.
    bool pending = true;
    while (pending) {
        const int32_t r0 = __builtin_amdgcn_readfirstlane(rid);
        if (rid != r0) continue;
        pending = false;
        const RegionDev& g = reg[r0];
        if (g.skip) break;
        if (g.clique && aux[c] >= 0) break;  // dense clique cells: k_dense_* kernels
        mask = g.is_periodic ? link_particle<true>(p, me, g, c, P, cell_ptr, aux, fmask, xc, parent)
                             : link_particle<false>(p, me, g, c, P, cell_ptr, aux, fmask, xc, parent);
    }
    dmask[p] = mask;
    zero0[p] = 0u;
    zero1[p] = 0;
}

// One entry of a mask: particle p (wave-uniform) against the dense cell at
// neighbour offset o (13: its own cell). The sub-cells inside the query box
// are spread over the wave's lanes, one sub-cell per lane. In a clique region
// a sub-cell needs one link at most; a cell that is one component needs none
// if it already shares p's root, else one link joins all of it. In other
// regions (sub-cells only prune) every linked pair is joined.
template <bool PERIODIC>
__device__ void dense_entry(uint32_t p, int o, const float4& me, const RegionDev& g, uint32_t c,
                            const LinkParams& P, const uint32_t* __restrict__ cell_ptr,
                            const int32_t* __restrict__ aux, const uint32_t* __restrict__ sub_ptr,
                            const uint32_t* __restrict__ sub_list, const int32_t* __restrict__ ncomp,
                            const int32_t* __restrict__ croot, const float4* __restrict__ xc,
                            int32_t* parent) {
    const int lane = __lane_id();
    int32_t ijk[3], jijk[3];
    double lo[3], q[3];
    cell_lower_corner(g, c - g.cell_off, lo, ijk);
    const double x0[3] = {(double)me.x, (double)me.y, (double)me.z};
    const bool own = o == 13;
    uint32_t jc = c;
    if (own) {
        for (int d = 0; d < 3; ++d) { jijk[d] = ijk[d]; q[d] = x0[d]; }
    } else {
        double sh[3];
        neighbour_cell(g, ijk, o, jijk, sh, P.L);  // exists: k_link reached it
        for (int d = 0; d < 3; ++d) {
            q[d] = x0[d] + sh[d];
            lo[d] = g.origin[d] + (double)jijk[d] * g.cell[d];
        }
        jc = g.cell_off + local_cell(g, jijk);
    }
    const uint32_t s = cell_ptr[jc];
    const int32_t dcell = aux[jc];
    const bool single = g.clique && !own && ncomp[dcell] == 1;
    if (single && dsu_find_near(parent, croot[dcell]) == dsu_find_near(parent, (int32_t)p)) return;
    const int32_t kk[3] = {g.k[0], g.k[1], g.k[2]};
    int32_t a[3], nb[3], mine[3] = {-1, -1, -1};
    for (int d = 0; d < 3; ++d) {
        a[d] = sub_coord(q[d] - P.R, lo[d], g.sub[d], kk[d]);
        nb[d] = sub_coord(q[d] + P.R, lo[d], g.sub[d], kk[d]) - a[d] + 1;
    }
    if (own && g.clique) {
        mine[0] = sub_coord((double)me.x, lo[0], g.sub[0], kk[0]);
        mine[1] = sub_coord((double)me.y, lo[1], g.sub[1], kk[1]);
        mine[2] = sub_coord((double)me.z, lo[2], g.sub[2], kk[2]);
    }
    const uint32_t total = (uint32_t)(nb[0] * nb[1] * nb[2]);
    const uint32_t* row0 = sub_ptr + (size_t)dcell * P.sub_stride;
    LinkCtx L{parent, (int32_t)p, 0u};
    for (uint32_t base = 0; base < total; base += kWave) {
        const uint32_t i = base + (uint32_t)lane;
        int32_t hit = -1;
        if (i < total) {
            const int32_t sx = a[0] + (int32_t)(i % (uint32_t)nb[0]);
            const int32_t sy = a[1] + (int32_t)((i / (uint32_t)nb[0]) % (uint32_t)nb[1]);
            const int32_t sz = a[2] + (int32_t)(i / (uint32_t)(nb[0] * nb[1]));
            if (!(sx == mine[0] && sy == mine[1] && sz == mine[2])) {
                const uint32_t* row = row0 + (size_t)kk[0] * (uint32_t)(sy + kk[1] * sz);
                const uint32_t t0 = row[sx], t1 = row[sx + 1];
                for (uint32_t t = t0; t < t1; ++t) {
                    const uint32_t qq = sub_list[s + t];
                    if (own && qq <= p) continue;
                    if (pair_linked<PERIODIC>(me, xc[qq], P)) {
                        if (g.clique) { hit = (int32_t)qq; break; }
                        L.link((int32_t)qq);
                    }
                }
            }
        }
        if (!g.clique) continue;
        const uint64_t hits = __ballot(hit >= 0);
        if (single) {
            if (hits) {
                if (lane == __ffsll((long long)hits) - 1) L.link(hit);
                return;
            }
        } else if (hit >= 0) {
            L.link(hit);
        }
    }
}

// The dense-cell scans left by k_link. Each wave takes the entries (particle,
// offset) of its 256 particles (four consecutive positions per lane: most
// masks are 0, and a quarter of the waves read them 0.13 ms faster on the
// 134M set) one at a time, lowest lane, position and offset first, and
// scans each with all its lanes.
__global__ void __launch_bounds__(kBlk) k_link_dense(const float4* __restrict__ xc,
                                                     const uint32_t* __restrict__ skeys,
                                                     const uint32_t* __restrict__ cell_ptr,
                                                     const int32_t* __restrict__ aux,
                                                     const uint32_t* __restrict__ sub_ptr,
                                                     const uint32_t* __restrict__ sub_list,
                                                     const int32_t* __restrict__ ncomp,
                                                     const int32_t* __restrict__ croot,
                                                     const RegionDev* __restrict__ reg,
                                                     const uint32_t* __restrict__ cell_off, int32_t* parent,
                                                     const uint32_t* __restrict__ dmask, LinkParams P) {
    // four consecutive positions per thread (most masks are 0)
    const uint32_t q0 = 4u * (xcd_block(blockIdx.x, gridDim.x) * kBlk + threadIdx.x);
    uint32_t mk[4] = {0u, 0u, 0u, 0u};
    if (q0 + 3u < P.N) {
        const uint4 v = *reinterpret_cast<const uint4*>(dmask + q0);
        mk[0] = v.x;
        mk[1] = v.y;
        mk[2] = v.z;
        mk[3] = v.w;
    } else {
        for (int j = 0; j < 4; ++j)
            if (q0 + (uint32_t)j < P.N) mk[j] = dmask[q0 + j];
    }
    const int lane = __lane_id();
    for (;;) {
        const bool any = (mk[0] | mk[1] | mk[2] | mk[3]) != 0u;
        const uint64_t todo = __ballot(any);
        if (todo == 0) break;
        const int src = __ffsll((long long)todo) - 1;
        // this lane's first position with a mask left, and that mask
        const int jj = mk[0] ? 0 : (mk[1] ? 1 : (mk[2] ? 2 : 3));
        const uint32_t mm = jj == 0 ? mk[0] : (jj == 1 ? mk[1] : (jj == 2 ? mk[2] : mk[3]));
        const int js = __builtin_amdgcn_readlane(jj, src);
        const uint32_t p = (uint32_t)__builtin_amdgcn_readlane((int)q0, src) + (uint32_t)js;
        const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)mm, src);
        const int o = __builtin_ctz(m);
        if (lane == src) {
            if (jj == 0) mk[0] &= mk[0] - 1u;
            else if (jj == 1) mk[1] &= mk[1] - 1u;
            else if (jj == 2) mk[2] &= mk[2] - 1u;
            else mk[3] &= mk[3] - 1u;
        }
        const float4 me = xc[p];
        const uint32_t c = skeys[p];
        const RegionDev& g = reg[P.nregions == 1 ? 0u : region_of_cell(cell_off, P.nregions, c)];
        if (g.is_periodic)
            dense_entry<true>(p, o, me, g, c, P, cell_ptr, aux, sub_ptr, sub_list, ncomp, croot, xc, parent);
        else
            dense_entry<false>(p, o, me, g, c, P, cell_ptr, aux, sub_ptr, sub_list, ncomp, croot, xc, parent);
    }
}

// Linking of dense cells whose sub-cells are cliques.
//
// Sub-cells tile the region on a fine grid (k per cell and axis, side
// s <= (b - 2m)/sqrt 3), so a pair passing the predicate lies in fine cells at
// most 2 apart per axis (2s > b + m). Each unordered pair of sub-cells is
// visited once: a lexicographically positive offset inside one cell
// (k_dense_intra), any offset into a coarse cell that is forward under the
// reference's neighbour rule (k_dense_cross). A visit tests member pairs until
// the first one that passes and then unites the two cliques.

// forward offsets in [-2,2]^3 (lexicographically positive), as (dx,dy,dz)
__constant__ int8_t c_fwd_off[62][3];

__device__ __forceinline__ int32_t lds_find(int32_t* lp, int32_t x) {
    for (;;) {
        const int32_t p = __hip_atomic_load(&lp[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (p == x) return x;
        const int32_t gp = __hip_atomic_load(&lp[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (gp == p) return p;
        __hip_atomic_store(&lp[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = gp;
    }
}

__device__ __forceinline__ void lds_unite(int32_t* lp, int32_t a, int32_t b) {
    for (;;) {
        a = lds_find(lp, a);
        b = lds_find(lp, b);
        if (a == b) return;
        const int32_t lo = a < b ? a : b, hi = a < b ? b : a;
        if (atomicCAS(&lp[hi], hi, lo) == hi) return;
    }
}

// One wave per cell of kMidMin..kMidMax particles that is not dense: every
// pair of its members is tested (the same predicate as k_link's own-cell
// loop) and joined in a union-find in LDS; each member's parent becomes the
// smallest position of its local component, a valid state for the global
// union-find (parents point to smaller positions). k_link then skips these
// cells' own pairs.
//
// A 64-pair round links about 2 pairs in the clump cells of the benchmark,
// so joining them where they are found would run the union-find loops with
// a lane or two of 64. The linked pairs are queued in LDS instead and joined
// 64 at a time, one per lane.
constexpr uint32_t kIntraQueue = 3 * kWave;

template <bool PERIODIC>
__device__ void cell_intra(uint32_t s, uint32_t n, const LinkParams& P, const float4* __restrict__ xc,
                           int32_t* parent, float4* __restrict__ pos, int32_t* __restrict__ lp,
                           uint16_t* __restrict__ hq) {
    const int lane = __lane_id();
    const uint64_t below = lanemask_lt();
    for (uint32_t i = lane; i < n; i += kWave) {
        pos[i] = xc[s + i];
        lp[i] = (int32_t)i;
    }
    __syncthreads();
    // every unordered pair once, all lanes busy: pair t = (m - 1) n + k is
    // (k, (k + m) mod n) for m = 1 .. (n - 1) / 2 and every k, then, for even
    // n, (k, k + n / 2) for k < n / 2. Pair t + kWave is (k + r, m + q) with
    // kWave = q n + r (one more m where k + r wraps).
    const uint32_t total = n * ((n - 1) / 2) + ((n & 1u) ? 0u : n / 2);
    const uint32_t q64 = (uint32_t)kWave / n, r64 = (uint32_t)kWave - q64 * n;
    uint32_t k = (uint32_t)lane % n, m = 1 + (uint32_t)lane / n;
    auto advance = [&](uint32_t& kk, uint32_t& mm) {
        kk += r64;
        mm += q64;
        if (kk >= n) {
            kk -= n;
            ++mm;
        }
    };
    uint32_t nq = 0;  // queued pairs (wave-uniform)
    // two rounds of kWave pairs per trip, their LDS loads issued together
    for (uint32_t t0 = 0; t0 < total; t0 += 2 * kWave) {
        uint32_t k1 = k, m1 = m;
        advance(k1, m1);
        bool hit0 = false, hit1 = false;
        uint32_t j0 = 0, j1 = 0;
        if (t0 + (uint32_t)lane < total) {
            j0 = k + m;
            if (j0 >= n) j0 -= n;
            hit0 = pair_linked_screened<PERIODIC>(pos[k], pos[j0], P);
        }
        if (t0 + (uint32_t)(kWave + lane) < total) {
            j1 = k1 + m1;
            if (j1 >= n) j1 -= n;
            hit1 = pair_linked_screened<PERIODIC>(pos[k1], pos[j1], P);
        }
        const uint64_t h0 = __ballot(hit0), h1 = __ballot(hit1);
        const uint32_t c0 = (uint32_t)__popcll(h0);
        if (hit0) hq[nq + (uint32_t)__popcll(h0 & below)] = (uint16_t)(k | (j0 << 8));
        if (hit1) hq[nq + c0 + (uint32_t)__popcll(h1 & below)] = (uint16_t)(k1 | (j1 << 8));
        nq += c0 + (uint32_t)__popcll(h1);
        while (nq >= (uint32_t)kWave) {
            // one queued pair per lane; the rest moves to the front
            __syncthreads();
            const uint32_t e = hq[lane];
            const uint32_t rest = nq - (uint32_t)kWave;
            const uint32_t moved0 = (uint32_t)lane < rest ? hq[kWave + lane] : 0u;
            const uint32_t moved1 = (uint32_t)(kWave + lane) < rest ? hq[2 * kWave + lane] : 0u;
            lds_unite(lp, (int32_t)(e & 0xffu), (int32_t)(e >> 8));
            __syncthreads();
            if ((uint32_t)lane < rest) hq[lane] = (uint16_t)moved0;
            if ((uint32_t)(kWave + lane) < rest) hq[kWave + lane] = (uint16_t)moved1;
            nq = rest;
            __syncthreads();
        }
        k = k1;
        m = m1;
        advance(k, m);
    }
    __syncthreads();
    if ((uint32_t)lane < nq) {
        const uint32_t e = hq[lane];
        lds_unite(lp, (int32_t)(e & 0xffu), (int32_t)(e >> 8));
    }
    __syncthreads();
    for (uint32_t i = lane; i < n; i += kWave) parent[s + i] = (int32_t)s + lds_find(lp, (int32_t)i);
}

// k_cell_intra runs beside k_link, which ends the linking phase (13 ms there
// against 3.9 ms alone on the 134M-particle clump set). At the occupancy its
// own 5.5 KB of LDS allows (29 one-wave workgroups per CU) it holds CU slots
// k_link needs; launched with 11 KiB of LDS in all (14 per CU), the call was
// 0.15-0.2 ms faster in each of three paired runs (3 to 5 waves per SIMD
// gained 0.0-0.28 ms in eight more). The padding is dynamic LDS the kernel
// does not touch.
constexpr size_t kIntraLdsStatic = sizeof(float4) * kMidMax + sizeof(int32_t) * kMidMax +
                                   sizeof(uint16_t) * kIntraQueue;
constexpr size_t kIntraLdsPad = 11264 - kIntraLdsStatic;

__global__ void __launch_bounds__(kWave) k_cell_intra(const uint32_t* __restrict__ mid_cells,
                                                      const uint32_t* __restrict__ cell_ptr,
                                                      const uint32_t* __restrict__ cell_off,
                                                      const RegionDev* __restrict__ reg, int32_t nreg,
                                                      const float4* __restrict__ xc, int32_t* parent,
                                                      LinkParams P) {
    __shared__ float4 pos[kMidMax];
    __shared__ int32_t lp[kMidMax];
    __shared__ uint16_t hq[kIntraQueue];
    const uint32_t c = mid_cells[xcd_block(blockIdx.x, gridDim.x)];
    const RegionDev& g = reg[nreg == 1 ? 0u : region_of_cell(cell_off, nreg, c)];
    const uint32_t s = cell_ptr[c], n = cell_ptr[c + 1] - s;
    if (g.is_periodic) cell_intra<true>(s, n, P, xc, parent, pos, lp, hq);
    else cell_intra<false>(s, n, P, xc, parent, pos, lp, hq);
}

// One wave per mid cell A (kAuxMid): every pair of A's members with a
// neighbour cell B: a forward mid cell (the lower cell takes the pair), and
// any other cell in either direction (k_link skips mid neighbours, and the
// dense-cell kernels take no pairs with mid cells). Only members within R = b + m
// of the shared face, edge or corner can link, so both sides are filtered to
// those first (the same margins as k_link's pruning) and the remaining pairs
// are tested by all lanes. Each staged member carries its union-find parent
// (in .w, which the predicates do not read; k_cell_intra and k_dense_intra
// leave every member pointing at its local root), so most linked pairs of two
// cells share one pair of parents: a pair with equal parents, or with the two
// parents this wave has just joined, is connected already and is not tested,
// and a passing pair joins the two parents, each distinct pair of a round
// once. k_link does nothing for A's particles.
template <bool PERIODIC>
__device__ void cell_pair(uint32_t c, const RegionDev& g, const LinkParams& P,
                          const uint32_t* __restrict__ cell_ptr, const int32_t* __restrict__ aux,
                          const uint8_t* __restrict__ fmask,
                          const uint32_t* __restrict__ sub_ptr, const uint32_t* __restrict__ sub_list,
                          const float4* __restrict__ xc, int32_t* parent, float4* pa, float4* pb,
                          uint16_t* ia, uint8_t* fa, uint2* uq) {
    const int lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    // linked pairs, queued and joined one per lane once kWave are waiting (a
    // round of pair tests links a few pairs of 64)
    uint32_t nq = 0;
    auto drain = [&](bool all) {
        __syncthreads();
        const uint32_t take = all ? nq : (uint32_t)kWave;
        const uint32_t rest = nq - take;
        const uint2 e = (uint32_t)lane < take ? uq[lane] : make_uint2(0u, 0u);
        const uint2 moved = (uint32_t)lane < rest ? uq[take + lane] : make_uint2(0u, 0u);
        if ((uint32_t)lane < take) dsu_unite(parent, (int32_t)e.x, (int32_t)e.y);
        __syncthreads();
        if ((uint32_t)lane < rest) uq[lane] = moved;
        nq = rest;
        __syncthreads();
    };
    const uint32_t lc = c - g.cell_off;
    int32_t ijk[3];
    double lo[3];
    cell_lower_corner(g, lc, lo, ijk);
    const uint32_t sA = cell_ptr[c], nA = cell_ptr[c + 1] - sA;
    // lane o < 27 looks up neighbour o: its cell, class and member range, in
    // one round of loads together with A's first kWave members and their
    // parents. Pairs with another mid cell are taken once (forward cells),
    // pairs with any other cell in both directions.
    uint32_t n_s = 0, n_e = 0, n_fm = 0;
    int32_t n_cls = -1;
    bool n_fwd = false;
    if (lane < 27 && lane != 13) {
        int32_t jijk[3];
        double sh[3];
        if (neighbour_cell(g, ijk, lane, jijk, sh, P.L)) {
            const uint32_t jl = local_cell(g, jijk);
            const uint32_t cB = g.cell_off + jl;
            n_cls = aux[cB];
            n_s = cell_ptr[cB];
            n_e = cell_ptr[cB + 1];
            n_fm = fmask[cB];
            n_fwd = jl > lc;
        }
    }
    // A's members and their face bits: bit 2d when the member lies within
    // R = b + m of the lower face of axis d (with the 2m pruning margin),
    // bit 2d+1 for the upper face
    uint32_t fa0 = 0;  // face bits of member `lane` (the first kWave members)
    for (uint32_t i = lane; i < nA; i += kWave) {
        const float4 a = xc[sA + i];
        pa[i] = make_float4(a.x, a.y, a.z, __int_as_float(load_near(&parent[sA + i])));
        const uint32_t f = __float_as_uint(a.w) & 0xffu;  // face bits (k_keys)
        fa[i] = (uint8_t)f;
        if (i < (uint32_t)kWave) fa0 = f;
    }
    const uint32_t n_n = n_e - n_s;
    const bool n_dense = n_cls >= 0;
    // B's members toward A: a cell none of whose members lies near the
    // shared face(s) holds no partner (the cell's face mask, k_gather)
    uint32_t n_needB = 0;
    if (lane < 27) {
        const int32_t off[3] = {lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1};
        for (int d = 0; d < 3; ++d) {
            if (off[d] < 0) n_needB |= 2u << (2 * d);
            else if (off[d] > 0) n_needB |= 1u << (2 * d);
        }
    }
    const bool n_work = n_n > 0 && !(!n_fwd && n_cls == kAuxMid) && (n_fm & n_needB) == n_needB;
    __syncthreads();
    uint64_t work = __ballot(n_work);
    while (work) {
        const int o = __ffsll((long long)work) - 1;
        work &= work - 1ull;
        const uint32_t sB = __shfl(n_s, o, kWave), nB = __shfl(n_n, o, kWave);
        const bool denseB = __shfl((int32_t)n_dense, o, kWave) != 0;
        const int32_t off[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
        // face bits a member needs to reach the neighbour: A's toward B, and
        // B's toward A (the opposite faces)
        uint32_t need = 0, needB = 0;
        for (int d = 0; d < 3; ++d) {
            if (off[d] < 0) {
                need |= 1u << (2 * d);
                needB |= 2u << (2 * d);
            } else if (off[d] > 0) {
                need |= 2u << (2 * d);
                needB |= 1u << (2 * d);
            }
        }
        // A's members that can reach B, in member order (the first kWave
        // from registers)
        uint32_t na;
        {
            const bool ok = (uint32_t)lane < nA && (fa0 & need) == need;
            const uint64_t m = __ballot(ok);
            if (ok) ia[(uint32_t)__popcll(m & below)] = (uint16_t)lane;
            na = (uint32_t)__popcll(m);
        }
        for (uint32_t i0 = kWave; i0 < nA; i0 += kWave) {
            const uint32_t i = i0 + (uint32_t)lane;
            const bool ok = i < nA && (fa[i] & need) == need;
            const uint64_t m = __ballot(ok);
            if (ok) ia[na + (uint32_t)__popcll(m & below)] = (uint16_t)i;
            na += (uint32_t)__popcll(m);
        }
        if (na == 0) continue;
        __syncthreads();
        // B's members that can reach A are staged kWave at a time (a
        // non-dense B can be large); each full window is tested against A
        int32_t ja = -1, jb = -1;  // the parent pair this wave joined last
        auto flush = [&](uint32_t nb) {
            __syncthreads();
            for (uint32_t t0 = 0; t0 < na * nb; t0 += kWave) {
                const uint32_t t = t0 + (uint32_t)lane;
                int32_t xa = 0, xb = 0;
                bool hit = false;
                if (t < na * nb) {
                    const uint32_t i = t / nb, j = t - i * nb;
                    const float4 a = pa[ia[i]], q = pb[j];
                    xa = __float_as_int(a.w);
                    xb = __float_as_int(q.w);
                    if (xa != xb && !(xa == ja && xb == jb)) hit = pair_linked_screened<PERIODIC>(a, q, P);
                }
                // each distinct parent pair of the round is queued once
                uint64_t m = __ballot(hit);
                while (m) {
                    const int l = __ffsll((long long)m) - 1;
                    const int32_t la = __shfl(xa, l, kWave), lb = __shfl(xb, l, kWave);
                    if (lane == 0) uq[nq] = make_uint2((uint32_t)la, (uint32_t)lb);
                    ++nq;
                    ja = la;
                    jb = lb;
                    hit = hit && !(xa == la && xb == lb);
                    m = __ballot(hit);
                }
                if (nq >= (uint32_t)kWave) drain(false);
            }
            __syncthreads();
        };
        // stages the lanes' candidates (a member of B at position q with
        // parent pr, loaded together; ok when its face bits reach A),
        // flushing full windows
        uint32_t nb = 0;
        auto stage = [&](bool ok, float4 q, int32_t pr) {
            ok = ok && (__float_as_uint(q.w) & needB) == needB;  // B's face bits toward A
            const uint64_t m = __ballot(ok);
            const uint32_t cnt = (uint32_t)__popcll(m);
            if (nb + cnt > (uint32_t)kWave) {
                flush(nb);
                nb = 0;
            }
            if (ok) {
                const uint32_t k = nb + (uint32_t)__popcll(m & below);
                q.w = __int_as_float(pr);
                pb[k] = q;
            }
            nb += cnt;
        };
        if (!denseB) {
            for (uint32_t i0 = 0; i0 < nB; i0 += kWave) {
                const uint32_t i = i0 + (uint32_t)lane;
                const bool ok = i < nB;
                stage(ok, ok ? xc[sB + i] : make_float4(0.f, 0.f, 0.f, 0.f), ok ? load_near(&parent[sB + i]) : 0);
            }
        } else {
            // A dense B holds its members by sub-cell: only the sub-cells
            // within R + 4m of the face toward A can hold members with those
            // face bits (sub-cell coordinates are monotone in x; the margin
            // exceeds the face bits' 2m and the f64 rounding), so only they
            // are loaded. Lanes take the box's sub-cells, a wave scan of
            // their member counts spreads the members over the lanes.
            const int32_t dB = __shfl(n_cls, o, kWave);
            const uint32_t* row = sub_ptr + (size_t)dB * P.sub_stride;
            int32_t blo[3], bn[3];
            for (int d = 0; d < 3; ++d) {
                const int32_t k = g.k[d];
                const double reach = P.R + 4.0 * P.m;
                if (off[d] > 0) {
                    blo[d] = 0;
                    bn[d] = (int32_t)fmin((double)k, floor(reach / g.sub[d]) + 1.0);
                } else if (off[d] < 0) {
                    blo[d] = (int32_t)fmax(0.0, floor((g.cell[d] - reach) / g.sub[d]));
                    bn[d] = k - blo[d];
                } else {
                    blo[d] = 0;
                    bn[d] = k;
                }
            }
            const uint32_t nbox = (uint32_t)(bn[0] * bn[1] * bn[2]);
            for (uint32_t b0 = 0; b0 < nbox; b0 += kWave) {
                const uint32_t bi = b0 + (uint32_t)lane;
                uint32_t t0 = 0, cnt = 0;
                if (bi < nbox) {
                    const uint32_t sx = (uint32_t)blo[0] + bi % (uint32_t)bn[0];
                    const uint32_t sy = (uint32_t)blo[1] + (bi / (uint32_t)bn[0]) % (uint32_t)bn[1];
                    const uint32_t sz = (uint32_t)blo[2] + bi / (uint32_t)(bn[0] * bn[1]);
                    const uint32_t idx = sx + (uint32_t)g.k[0] * (sy + (uint32_t)g.k[1] * sz);
                    t0 = row[idx];
                    cnt = row[idx + 1] - t0;
                }
                uint32_t incl = cnt;
                for (int sh = 1; sh < kWave; sh <<= 1) {
                    const uint32_t v = __shfl_up(incl, sh, kWave);
                    if (lane >= sh) incl += v;
                }
                const uint32_t total = __shfl(incl, kWave - 1, kWave);
                for (uint32_t o0 = 0; o0 < total; o0 += kWave) {
                    const uint32_t oo = o0 + (uint32_t)lane;
                    // the box lane holding member oo: the first with incl > oo
                    int L = 0;
                    for (int step = kWave / 2; step > 0; step >>= 1) {
                        const uint32_t v = (uint32_t)__shfl((int)incl, L + step - 1, kWave);
                        if (v <= oo) L += step;
                    }
                    const uint32_t first = (uint32_t)__shfl((int)t0, L, kWave);
                    const uint32_t before = (uint32_t)__shfl((int)(incl - cnt), L, kWave);
                    const bool ok = oo < total;
                    const uint32_t i = ok ? sub_list[sB + first + (oo - before)] - sB : 0u;
                    stage(ok, ok ? xc[sB + i] : make_float4(0.f, 0.f, 0.f, 0.f), ok ? load_near(&parent[sB + i]) : 0);
                }
            }
        }
        if (nb) flush(nb);
    }
    if (nq) drain(true);
}

__global__ void __launch_bounds__(kWave) k_cell_pair(const uint32_t* __restrict__ mid_cells,
                                                     const uint32_t* __restrict__ cell_ptr,
                                                     const uint32_t* __restrict__ cell_off,
                                                     const RegionDev* __restrict__ reg, int32_t nreg,
                                                     const int32_t* __restrict__ aux,
                                                     const uint8_t* __restrict__ fmask,
                                                     const uint32_t* __restrict__ sub_ptr,
                                                     const uint32_t* __restrict__ sub_list,
                                                     const float4* __restrict__ xc, int32_t* parent,
                                                     LinkParams P) {
    __shared__ float4 pa[kMidMax], pb[kWave];
    __shared__ uint16_t ia[kMidMax];
    __shared__ uint8_t fa[kMidMax];
    __shared__ uint2 uq[2 * kWave];
    const uint32_t c = mid_cells[xcd_block(blockIdx.x, gridDim.x)];
    const RegionDev& g = reg[nreg == 1 ? 0u : region_of_cell(cell_off, nreg, c)];
    if (g.is_periodic)
        cell_pair<true>(c, g, P, cell_ptr, aux, fmask, sub_ptr, sub_list, xc, parent, pa, pb, ia, fa, uq);
    else
        cell_pair<false>(c, g, P, cell_ptr, aux, fmask, sub_ptr, sub_list, xc, parent, pa, pb, ia, fa, uq);
}

template <bool PERIODIC>
__device__ __forceinline__ bool cliques_linked(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1,
                                               const uint32_t* __restrict__ sub_list,
                                               const float4* __restrict__ xc, const LinkParams& P) {
    for (uint32_t ta = a0; ta < a1; ++ta) {
        const float4 xa = xc[sub_list[ta]];
        for (uint32_t tb = b0; tb < b1; ++tb)
            if (pair_linked<PERIODIC>(xa, xc[sub_list[tb]], P)) return true;
    }
    return false;
}

// One workgroup per dense cell: union-find over its sub-cells in LDS, then
// every member's parent is set to the smallest position of its local
// component (a valid DSU state: parents point to smaller positions). Offsets
// are processed nearest first in classes separated by barriers, so that once
// the close sub-cells have merged the far offsets are skipped on the root
// check alone. Cells with at most kIntraStage members are staged in LDS.
// Each sub-cell's local root (or -1 for an empty one) is kept in sub_root,
// the row of the cell's sub_ptr layout, for k_dense_cross.
constexpr int kIntraStage = 768;
constexpr int kFwdClassEnd[5] = {3, 9, 13, 16, 62};

template <bool PERIODIC>
__device__ void dense_intra(uint32_t d, const RegionDev& g, const LinkParams& P,
                            const uint32_t* __restrict__ dense_cells, const uint32_t* __restrict__ cell_ptr,
                            const uint32_t* __restrict__ sub_ptr, const uint32_t* __restrict__ sub_list,
                            const float4* __restrict__ xc, int32_t* parent, int32_t* __restrict__ ncomp,
                            int32_t* __restrict__ croot, int32_t* __restrict__ needs_sub,
                            int32_t* __restrict__ sub_root, unsigned char* lds) {
    const int32_t k0 = g.k[0], k1 = g.k[1], k2 = g.k[2];
    const int32_t K = k0 * k1 * k2;
    uint32_t* rp = reinterpret_cast<uint32_t*>(lds);
    int32_t* lp = reinterpret_cast<int32_t*>(rp + P.sub_stride);
    uint32_t* mh = reinterpret_cast<uint32_t*>(lp + P.sub_stride);
    uint32_t* list = mh + P.sub_stride;
    uint32_t* nl = list + P.sub_stride;  // [0] list size, [1] component count, [2] roots counted
    float4* pos = reinterpret_cast<float4*>(lds + (size_t)16 * ((4 * P.sub_stride + 2 + 3) / 4));
    const uint32_t c = dense_cells[d];
    const uint32_t cs = cell_ptr[c], n_c = cell_ptr[c + 1] - cs;
    const bool staged = n_c <= (uint32_t)P.intra_stage;
    const uint32_t* row = sub_ptr + (size_t)d * P.sub_stride;
    if (threadIdx.x == 0) { nl[0] = 0; nl[1] = 0; }
    for (int i = threadIdx.x; i <= K; i += kBlk) rp[i] = row[i];
    if (staged)
        for (uint32_t t = threadIdx.x; t < n_c; t += kBlk) pos[t] = xc[sub_list[cs + t]];
    __syncthreads();
    // list entries: the sub-cell index (12 bits, K <= 4096) and its three
    // coordinates (4 bits each, k <= 16), found once per sub-cell here
    for (int i = threadIdx.x; i < K; i += kBlk) {
        lp[i] = i;
        mh[i] = 0xffffffffu;
        if (rp[i + 1] > rp[i]) {
            const uint32_t sx = (uint32_t)(i % k0), sy = (uint32_t)((i / k0) % k1), sz = (uint32_t)(i / (k0 * k1));
            list[atomicAdd(&nl[0], 1u)] = (uint32_t)i | (sx << 12) | (sy << 16) | (sz << 20);
        }
    }
    __syncthreads();
    const uint32_t na = nl[0];
    // one class of offsets [b0, b1): the class sizes are constants once the
    // calls below are inlined, so the (entry, offset) split costs no division
    auto offset_class = [&](const int b0, const int b1) {
        const uint32_t noff = (uint32_t)(b1 - b0);
        for (uint32_t it = threadIdx.x; it < na * noff; it += kBlk) {
            const uint32_t ent = list[it / noff];
            const int oi = b0 + (int)(it % noff);
            const int32_t A = (int32_t)(ent & 0xfffu);
            const int32_t ax = (int32_t)((ent >> 12) & 15u) + c_fwd_off[oi][0];
            const int32_t ay = (int32_t)((ent >> 16) & 15u) + c_fwd_off[oi][1];
            const int32_t az = (int32_t)((ent >> 20) & 15u) + c_fwd_off[oi][2];
            if (ax < 0 || ax >= k0 || ay < 0 || ay >= k1 || az < 0 || az >= k2) continue;
            const int32_t B = ax + k0 * (ay + k1 * az);
            if (rp[B + 1] == rp[B]) continue;
            if (lds_find(lp, A) == lds_find(lp, B)) continue;
            bool hit = false;
            if (staged) {
                for (uint32_t ta = rp[A]; ta < rp[A + 1] && !hit; ++ta) {
                    const float4 xa = pos[ta];
                    for (uint32_t tb = rp[B]; tb < rp[B + 1]; ++tb)
                        if (pair_linked<PERIODIC>(xa, pos[tb], P)) { hit = true; break; }
                }
            } else {
                hit = cliques_linked<PERIODIC>(cs + rp[A], cs + rp[A + 1], cs + rp[B], cs + rp[B + 1],
                                               sub_list, xc, P);
            }
            if (hit) lds_unite(lp, A, B);
        }
        __syncthreads();
    };
    static_assert(kFwdClassEnd[0] == 3 && kFwdClassEnd[1] == 9 && kFwdClassEnd[2] == 13 &&
                      kFwdClassEnd[3] == 16 && kFwdClassEnd[4] == 62,
                  "offset classes");
    // after each class but the last, a cell that is already one component
    // (one root among its non-empty sub-cells) has nothing left to join:
    // most clump-centre cells are, after the nearest offsets
    auto one_component = [&]() {
        if (threadIdx.x == 0) nl[2] = 0;
        __syncthreads();
        uint32_t roots = 0;
        for (uint32_t i = threadIdx.x; i < na; i += kBlk) {
            const int32_t A = (int32_t)(list[i] & 0xfffu);
            roots += lp[A] == A ? 1u : 0u;
        }
        if (roots) atomicAdd(&nl[2], roots);
        __syncthreads();
        return nl[2] == 1u;
    };
    offset_class(0, 3);
    if (!one_component()) {
        offset_class(3, 9);
        if (!one_component()) {
            offset_class(9, 13);
            if (!one_component()) {
                offset_class(13, 16);
                if (!one_component()) offset_class(16, 62);
            }
        }
    }
    for (uint32_t i = threadIdx.x; i < na; i += kBlk) {
        const int32_t A = (int32_t)(list[i] & 0xfffu);
        // head of A: its smallest position (set as parent of every member)
        const uint32_t h = (uint32_t)parent[sub_list[cs + rp[A]]];
        const int32_t r = lds_find(lp, A);
        atomicMin(&mh[r], h);
        if (r == A) atomicAdd(&nl[1], 1u);
    }
    __syncthreads();
    int32_t* srow = sub_root + (size_t)d * P.sub_stride;
    for (int i = threadIdx.x; i < K; i += kBlk)
        if (rp[i + 1] == rp[i]) srow[i] = -1;
    // Every member's parent is its sub-cell's head (smallest position) since
    // k_dense_build: the heads take their component's root, then every
    // position of the cell (cs .. cs + n_c, in order) its head's new parent.
    for (uint32_t i = threadIdx.x; i < na; i += kBlk) {
        const int32_t A = (int32_t)(list[i] & 0xfffu);
        const int32_t root = (int32_t)mh[lds_find(lp, A)];
        srow[A] = root;
        parent[parent[sub_list[cs + rp[A]]]] = root;
    }
    __syncthreads();
    for (uint32_t p = cs + threadIdx.x; p < cs + n_c; p += kBlk) parent[p] = parent[parent[p]];
    if (threadIdx.x == 0) {
        ncomp[d] = (int32_t)nl[1];
        croot[d] = nl[1] == 1 ? (int32_t)mh[lds_find(lp, (int32_t)(list[0] & 0xfffu))] : -1;
        needs_sub[d] = nl[1] != 1 ? 1 : 0;
    }
}

__global__ void __launch_bounds__(kBlk) k_dense_intra(const uint32_t* __restrict__ dense_cells,
                                                      const uint32_t* __restrict__ cell_ptr,
                                                      const uint32_t* __restrict__ cell_off,
                                                      const RegionDev* __restrict__ reg, int32_t nreg,
                                                      const uint32_t* __restrict__ sub_ptr,
                                                      const uint32_t* __restrict__ sub_list,
                                                      const float4* __restrict__ xc, int32_t* parent,
                                                      int32_t* __restrict__ ncomp,
                                                      int32_t* __restrict__ croot,
                                                      int32_t* __restrict__ needs_sub,
                                                      int32_t* __restrict__ sub_root, LinkParams P) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint32_t d = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t c = dense_cells[d];
    const RegionDev& g = reg[nreg == 1 ? 0u : region_of_cell(cell_off, nreg, c)];
    if (!g.clique) {
        if (threadIdx.x == 0) { ncomp[d] = -1; croot[d] = -1; needs_sub[d] = 0; }
        return;
    }
    if (g.is_periodic)
        dense_intra<true>(d, g, P, dense_cells, cell_ptr, sub_ptr, sub_list, xc, parent, ncomp, croot,
                          needs_sub, sub_root, lds);
    else
        dense_intra<false>(d, g, P, dense_cells, cell_ptr, sub_ptr, sub_list, xc, parent, ncomp, croot,
                           needs_sub, sub_root, lds);
}

// One wave per dense cell of a single local component: for each forward
// neighbour dense cell that is also a single component, decide once whether
// the two components link. If their roots differ, the sub-cell pairs across
// the shared face/edge/corner (fine offset <= 2 per axis) are searched in
// parallel, 64 at a time, until one pair of cliques links.
constexpr int kCrossRound = 16;  // member tests per lane between wave checks in cell_cross

template <bool PERIODIC>
__device__ void cell_cross(uint32_t d, const RegionDev& g, const LinkParams& P,
                           const uint32_t* __restrict__ dense_cells, const uint32_t* __restrict__ cell_ptr,
                           const int32_t* __restrict__ aux, const int32_t* __restrict__ ncomp,
                           const int32_t* __restrict__ croot, int32_t* __restrict__ needs_sub,
                           const uint32_t* __restrict__ sub_ptr,
                           const uint32_t* __restrict__ sub_list, const float4* __restrict__ xc,
                           int32_t* parent) {
    const uint32_t c = dense_cells[d];
    const uint32_t lc = c - g.cell_off;
    int32_t ijk[3];
    double lo_unused[3];
    cell_lower_corner(g, lc, lo_unused, ijk);
    const int32_t kk[3] = {g.k[0], g.k[1], g.k[2]};
    const uint32_t* rowA = sub_ptr + (size_t)d * P.sub_stride;
    const uint32_t csA = cell_ptr[c];
    const int lane = __lane_id();
    for (int o = 0; o < 27; ++o) {
        if (o == 13) continue;
        int32_t jijk[3];
        double sh[3];
        if (!neighbour_cell(g, ijk, o, jijk, sh, P.L)) continue;
        const uint32_t jl = local_cell(g, jijk);
        if (jl <= lc) continue;
        const uint32_t cB = g.cell_off + jl;
        const int32_t dB = aux[cB];
        if (dB < 0) continue;
        if (ncomp[dB] != 1) {
            if (lane == 0) needs_sub[d] = 1;
            continue;
        }
        int32_t rA = dsu_find_near(parent, croot[d]);
        int32_t rB = dsu_find_near(parent, croot[dB]);
        if (rA == rB) continue;
        // Per axis, a sub-cell pair is a position and an option: along an axis
        // parallel to the shared face, A's sub-cell coordinate (kk positions)
        // and B's lateral offset from it (0, -1, +1, -2, +2); across it, one
        // position and the depth pairs (k-1,0) (k-1,1) (k-2,0) toward +1,
        // mirrored toward -1. Positions vary fastest, so the directly opposite
        // sub-cells (all options 0) are searched first: where the two cells
        // touch, one of them links at once.
        const int32_t co[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
        int32_t npos[3], nopt[3];
        for (int ax = 0; ax < 3; ++ax) {
            npos[ax] = co[ax] == 0 ? kk[ax] : 1;
            nopt[ax] = co[ax] == 0 ? 5 : 3;
        }
        const uint32_t total = (uint32_t)(npos[0] * npos[1] * npos[2]) * (uint32_t)(nopt[0] * nopt[1] * nopt[2]);
        const uint32_t* rowB = sub_ptr + (size_t)dB * P.sub_stride;
        const uint32_t csB = cell_ptr[cB];
        for (uint32_t base = 0; base < total; base += kWave) {
            const uint32_t it = base + (uint32_t)lane;
            bool hit = false, active = false;
            uint32_t ta = 0, a1 = 0, tb = 0, b0 = 0, b1 = 0;
            if (it < total) {
                uint32_t rem = it;
                int32_t pos[3], opt[3], sa[3], sb[3];
                bool valid = true;
                for (int ax = 0; ax < 3; ++ax) {
                    pos[ax] = (int32_t)(rem % (uint32_t)npos[ax]);
                    rem /= (uint32_t)npos[ax];
                }
                for (int ax = 0; ax < 3; ++ax) {
                    opt[ax] = (int32_t)(rem % (uint32_t)nopt[ax]);
                    rem /= (uint32_t)nopt[ax];
                }
                for (int ax = 0; ax < 3; ++ax) {
                    if (co[ax] == 0) {
                        const int32_t lat = (opt[ax] + 1) / 2 * ((opt[ax] & 1) ? -1 : 1);  // 0 -1 1 -2 2
                        sa[ax] = pos[ax];
                        sb[ax] = pos[ax] + lat;
                        if (sb[ax] < 0 || sb[ax] >= kk[ax]) valid = false;
                    } else {
                        const int32_t da = opt[ax] == 2 ? kk[ax] - 2 : kk[ax] - 1;
                        const int32_t db = opt[ax] == 1 ? 1 : 0;
                        sa[ax] = co[ax] > 0 ? da : kk[ax] - 1 - da;
                        sb[ax] = co[ax] > 0 ? db : kk[ax] - 1 - db;
                    }
                }
                if (valid) {
                    const uint32_t ia = (uint32_t)sa[0] + (uint32_t)kk[0] * ((uint32_t)sa[1] + (uint32_t)kk[1] * (uint32_t)sa[2]);
                    const uint32_t ib = (uint32_t)sb[0] + (uint32_t)kk[0] * ((uint32_t)sb[1] + (uint32_t)kk[1] * (uint32_t)sb[2]);
                    ta = csA + rowA[ia];
                    a1 = csA + rowA[ia + 1];
                    b0 = csB + rowB[ib];
                    b1 = csB + rowB[ib + 1];
                    active = ta < a1 && b0 < b1;
                    tb = b0;
                }
            }
            // each lane scans its pair of sub-cells member by member, at most
            // kCrossRound tests between checks of the wave: once any lane has
            // linked, the others stop (one link joins the two components)
            float4 xa = active ? xc[sub_list[ta]] : make_float4(0.f, 0.f, 0.f, 0.f);
            for (;;) {
                for (int r = 0; r < kCrossRound && active; ++r) {
                    if (pair_linked<PERIODIC>(xa, xc[sub_list[tb]], P)) {
                        hit = true;
                        active = false;
                    } else if (++tb == b1) {
                        tb = b0;
                        if (++ta == a1) active = false;
                        else xa = xc[sub_list[ta]];
                    }
                }
                if (__ballot(hit) || !__ballot(active)) break;
            }
            const uint64_t hits = __ballot(hit);
            if (hits) {
                if (lane == __ffsll((long long)hits) - 1) dsu_unite(parent, rA, rB);
                break;
            }
        }
    }
}

__global__ void __launch_bounds__(kWave) k_dense_cell_cross(const uint32_t* __restrict__ dense_cells,
                                                            uint32_t ndense,
                                                            const uint32_t* __restrict__ cell_ptr,
                                                            const uint32_t* __restrict__ cell_off,
                                                            const RegionDev* __restrict__ reg, int32_t nreg,
                                                            const int32_t* __restrict__ aux,
                                                            const int32_t* __restrict__ ncomp,
                                                            const int32_t* __restrict__ croot,
                                                            int32_t* __restrict__ needs_sub,
                                                            const uint32_t* __restrict__ sub_ptr,
                                                            const uint32_t* __restrict__ sub_list,
                                                            const float4* __restrict__ xc, int32_t* parent,
                                                            LinkParams P) {
    const uint32_t d = xcd_block(blockIdx.x, gridDim.x);
    if (d >= ndense || ncomp[d] != 1) return;
    const uint32_t c = dense_cells[d];
    const RegionDev& g = reg[nreg == 1 ? 0u : region_of_cell(cell_off, nreg, c)];
    if (g.is_periodic)
        cell_cross<true>(d, g, P, dense_cells, cell_ptr, aux, ncomp, croot, needs_sub, sub_ptr, sub_list, xc, parent);
    else
        cell_cross<false>(d, g, P, dense_cells, cell_ptr, aux, ncomp, croot, needs_sub, sub_ptr, sub_list, xc, parent);
}

constexpr int kJoinedCache = 6;  // local roots remembered as joined in dense_cross

// One thread per non-empty clique sub-cell A of a cell with several local
// components, or next to such a cell: offsets into forward dense coarse cells,
// sub-cell against sub-cell. (Pairs with sparse cells are taken by k_link.)
template <bool PERIODIC>
__device__ __forceinline__ void dense_cross(const uint4 ent, const RegionDev& g, const LinkParams& P,
                                            const uint32_t* __restrict__ dense_cells,
                                            const uint32_t* __restrict__ cell_ptr,
                                            const int32_t* __restrict__ aux,
                                            const int32_t* __restrict__ ncomp,
                                            const int32_t* __restrict__ croot,
                                            const int32_t* __restrict__ needs_sub,
                                            const uint32_t* __restrict__ sub_ptr,
                                            const uint32_t* __restrict__ sub_list,
                                            const int32_t* __restrict__ sub_root,
                                            const float4* __restrict__ xc, int32_t* parent) {
    const uint32_t d = ent.x, si = ent.y;
    const uint32_t c = dense_cells[d];
    const uint32_t lc = c - g.cell_off;
    int32_t ijk[3];
    double lo[3];
    cell_lower_corner(g, lc, lo, ijk);
    const int32_t kk[3] = {g.k[0], g.k[1], g.k[2]};
    const int32_t sa[3] = {(int32_t)(si % (uint32_t)kk[0]), (int32_t)((si / (uint32_t)kk[0]) % (uint32_t)kk[1]),
                           (int32_t)(si / (uint32_t)(kk[0] * kk[1]))};
    // cells whose pairs were all decided per cell pair, and interior
    // sub-cells (they reach no other cell), have nothing left to do
    if (!needs_sub[d]) return;
    bool interior = true;
    for (int ax = 0; ax < 3; ++ax)
        if (sa[ax] < 2 || sa[ax] > kk[ax] - 3) interior = false;
    if (interior) return;
    // Per axis, the window's offsets -2..2 (o = 0..4) fall into up to three
    // parts by the coarse cell they reach (shift -1, 0, +1): the offsets
    // [olo, ohi] of each part and that cell's coordinate (or -1). Cells are
    // then taken one at a time, each checked once, and only their offsets
    // visited.
    int32_t olo[3][3], ohi[3][3], jc3[3][3];
    for (int ax = 0; ax < 3; ++ax) {
        const int32_t k = kk[ax], a = sa[ax];
        // sub-coordinate f = a + o - 2: below 0, in [0, k), at k or above
        olo[ax][0] = 0;
        ohi[ax][0] = min(4, 1 - a);
        olo[ax][1] = max(0, 2 - a);
        ohi[ax][1] = min(4, k + 1 - a);
        olo[ax][2] = max(0, k + 2 - a);
        ohi[ax][2] = 4;
        for (int part = 0; part < 3; ++part) {
            int32_t j = ijk[ax] + part - 1;
            if (j < 0 || j >= g.n[ax]) {
                if (g.periodic[ax]) j = j < 0 ? j + g.n[ax] : j - g.n[ax];
                else j = -1;
            }
            jc3[ax][part] = j;
        }
    }
    const uint32_t* rowA = sub_ptr + (size_t)d * P.sub_stride;
    const uint32_t cs = cell_ptr[c];
    const uint32_t a0 = cs + rowA[si], a1 = cs + rowA[si + 1];
    int32_t rA = dsu_find_near(parent, (int32_t)ent.z);
    // pairs of single-component dense cells are decided per cell pair by
    // k_dense_cell_cross. B's sub-cell is known by its local root (k_dense_
    // intra's component of that cell); the last kJoinedCache local roots found
    // joined to A are remembered, so the sub-cells of a local component
    // already joined are skipped without a find.
    const bool single = ncomp[d] == 1;
    int32_t joined[kJoinedCache];
#pragma unroll
    for (int i = 0; i < kJoinedCache; ++i) joined[i] = -1;
    int next = 0;
    auto remember = [&](int32_t lb) {
#pragma unroll
        for (int i = 0; i < kJoinedCache; ++i)
            if (i == next) joined[i] = lb;
        next = next + 1 == kJoinedCache ? 0 : next + 1;
    };
    // a neighbour cell's class and component count are loaded once for all
    // the offsets into it; each offset then costs one load (its sub-cell's
    // local root)
    for (int pz = 0; pz < 3; ++pz) {
        const int32_t jz = jc3[2][pz];
        if (jz < 0 || olo[2][pz] > ohi[2][pz]) continue;
        for (int py = 0; py < 3; ++py) {
            const int32_t jy = jc3[1][py];
            if (jy < 0 || olo[1][py] > ohi[1][py]) continue;
            for (int px = 0; px < 3; ++px) {
                const int32_t jx = jc3[0][px];
                if (jx < 0 || olo[0][px] > ohi[0][px]) continue;
                if (jx == ijk[0] && jy == ijk[1] && jz == ijk[2]) continue;  // same cell: intra
                const uint32_t jl = (uint32_t)jx + (uint32_t)g.n[0] * ((uint32_t)jy + (uint32_t)g.n[1] * (uint32_t)jz);
                if (jl <= lc) continue;
                const uint32_t cB = g.cell_off + jl;
                const int32_t dB = aux[cB];
                if (dB < 0) continue;  // sparse cell: k_link
                if (single && ncomp[dB] == 1) continue;  // decided by k_dense_cell_cross
                const uint32_t* rowB = sub_ptr + (size_t)dB * P.sub_stride;
                const int32_t* rootB = sub_root + (size_t)dB * P.sub_stride;
                const uint32_t bs = cell_ptr[cB];
                for (int oz = olo[2][pz]; oz <= ohi[2][pz]; ++oz) {
                    const int32_t fz = sa[2] + oz - 2 - (pz - 1) * kk[2];
                    for (int oy = olo[1][py]; oy <= ohi[1][py]; ++oy) {
                        const int32_t fy = sa[1] + oy - 2 - (py - 1) * kk[1];
                        for (int ox = olo[0][px]; ox <= ohi[0][px]; ++ox) {
                            const int32_t fx = sa[0] + ox - 2 - (px - 1) * kk[0];
                            const uint32_t sb = (uint32_t)fx + (uint32_t)kk[0] * ((uint32_t)fy + (uint32_t)kk[1] * (uint32_t)fz);
                            const int32_t lb = rootB[sb];
                            if (lb < 0) continue;  // empty sub-cell
                            bool known = false;
#pragma unroll
                            for (int i = 0; i < kJoinedCache; ++i) known = known || joined[i] == lb;
                            if (known) continue;
                            const int32_t rB = dsu_find_near(parent, lb);
                            if (rB != rA) rA = dsu_find_near(parent, rA);
                            if (rB == rA) {
                                remember(lb);
                                continue;
                            }
                            const uint32_t b0 = bs + rowB[sb], b1 = bs + rowB[sb + 1];
                            if (cliques_linked<PERIODIC>(a0, a1, b0, b1, sub_list, xc, P)) {
                                rA = dsu_unite(parent, rA, rB);
                                remember(lb);
                            }
                        }
                    }
                }
            }
        }
    }
}

__global__ void __launch_bounds__(kBlk) k_dense_cross(const uint4* __restrict__ subs, uint32_t nsubs,
                                                      const uint32_t* __restrict__ dense_cells,
                                                      const uint32_t* __restrict__ cell_ptr,
                                                      const int32_t* __restrict__ aux,
                                                      const int32_t* __restrict__ ncomp,
                                                      const int32_t* __restrict__ croot,
                                                      const int32_t* __restrict__ needs_sub,
                                                      const uint32_t* __restrict__ sub_ptr,
                                                      const uint32_t* __restrict__ sub_list,
                                                      const int32_t* __restrict__ sub_root,
                                                      const RegionDev* __restrict__ reg,
                                                      const float4* __restrict__ xc, int32_t* parent,
                                                      LinkParams P) {
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * kBlk + threadIdx.x;
    if (t >= nsubs) return;
    const uint4 ent = subs[t];
    const RegionDev& g = reg[ent.w];
    if (g.is_periodic)
        dense_cross<true>(ent, g, P, dense_cells, cell_ptr, aux, ncomp, croot, needs_sub, sub_ptr, sub_list,
                          sub_root, xc, parent);
    else
        dense_cross<false>(ent, g, P, dense_cells, cell_ptr, aux, ncomp, croot, needs_sub, sub_ptr, sub_list,
                           sub_root, xc, parent);
}

// Chunks of kWave positions per wave of k_label_count, and chunks whose
// parents are loaded together.
constexpr int kCountChunks = 256;

constexpr int kCountBatch = 4;

// Labels and component sizes in one pass (after linking; nothing changes
// the forest meanwhile). Each wave walks kCountChunks consecutive chunks of
// kWave positions, kCountBatch chunks at a time: the parents of the batch's
// positions are followed to their roots with all the batch's loads of a step
// issued together (most paths are a few steps after the path halving of
// linking), and the labels written. Runs of equal labels are counted per
// chunk (size was zero-filled), and the run that reaches a chunk's last lane
// is carried into the next chunk instead of being added at once, so a large
// component costs one atomic per wave rather than per chunk (its counter is
// the contended one). A position past n ends a run. A root is the smallest
// position of its component, so the run holding it starts there: the root
// itself is counted by the root bitmap (bit p of roots: position p is a
// root), not in size, and an isolated particle costs no atomic at all;
// a component's size is size[r] + 1 at its root r.
__global__ void k_label_count(const int32_t* __restrict__ parent, int32_t* __restrict__ label, uint32_t n,
                              uint32_t* __restrict__ size, unsigned long long* __restrict__ roots) {
    const int lane = __lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t base = wave * kWave * kCountChunks;
    int32_t carry = -1;  // wave-uniform: the open run's label (>= 0), count
    uint32_t ccount = 0, croot = 0;  // and whether it began at its root
    for (int k0 = 0; k0 < kCountChunks; k0 += kCountBatch) {
        int32_t x[kCountBatch];
#pragma unroll
        for (int b = 0; b < kCountBatch; ++b) {
            const uint64_t p = base + (uint64_t)(k0 + b) * kWave + (uint64_t)lane;
            x[b] = p < n ? parent[p] : -1;
        }
        for (;;) {
            int32_t y[kCountBatch];
#pragma unroll
            for (int b = 0; b < kCountBatch; ++b) y[b] = x[b] >= 0 ? parent[x[b]] : -1;
            bool moved = false;
#pragma unroll
            for (int b = 0; b < kCountBatch; ++b) {
                moved = moved || y[b] != x[b];
                x[b] = y[b];
            }
            if (!__ballot(moved)) break;
        }
#pragma unroll
        for (int b = 0; b < kCountBatch; ++b) {
            const int k = k0 + b;
            const uint64_t p = base + (uint64_t)k * kWave + (uint64_t)lane;
            const bool valid = p < n;
            if (valid) label[p] = x[b];
            const int32_t lab = valid ? x[b] : -2 - lane;  // unique: never continues a run
            const uint64_t rbits = __ballot(valid && x[b] == (int32_t)p);
            if (lane == 0 && valid) roots[wave * kCountChunks + (uint64_t)k] = rbits;
            const int32_t prev = __shfl_up(lab, 1, kWave);
            const bool head = lane == 0 || prev != lab;
            const uint64_t heads = __ballot(head);
            const int32_t lab0 = __shfl(lab, 0, kWave);
            const int lasthead = 63 - __clzll((long long)heads);
            const bool last = k + 1 == kCountChunks;
            const bool merge0 = carry >= 0 && lab0 == carry;
            if (carry >= 0 && !merge0 && lane == 0 && ccount > croot) atomicAdd(&size[carry], ccount - croot);
            uint32_t cnt = 0, rs = 0;
            if (valid && head) {
                const uint64_t above = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
                const int next = above ? __ffsll((long long)above) - 1 : 64;
                cnt = (uint32_t)(next - lane) + (lane == 0 && merge0 ? ccount : 0u);
                rs = lane == 0 && merge0 ? croot : (lab == (int32_t)p ? 1u : 0u);
                if ((last || lane != lasthead) && cnt > rs) atomicAdd(&size[lab], cnt - rs);
            }
            // the run at the last head stays open (its label is valid when >= 0)
            const int32_t labL = __shfl(lab, lasthead, kWave);
            const uint32_t cntL = __shfl(cnt, lasthead, kWave);
            const uint32_t rsL = __shfl(rs, lasthead, kWave);
            carry = (!last && labL >= 0) ? labL : -1;
            ccount = carry >= 0 ? cntL : 0u;
            croot = carry >= 0 ? rsL : 0u;
        }
    }
}

// Candidates (fof_fortran.f90:866-872): positions whose component size
// (0 for non-roots) reaches np_min. A component's size is size[r] + 1 at its
// root r (k_label_count) and size is 0 at every other position, so a position
// qualifies when size + its root bit >= np_min; for np_min >= 2 that is
// size >= np_min - 1 (roots == nullptr, thr = np_min - 1). Each wave takes
// kCandWords words of 64 positions (their loads issued together) and writes
// their candidate bits (cbits, bit q of word q / 64) and counts (ccnt,
// scanned afterwards into the count before each word).
constexpr int kCandWords = 4;
__global__ void k_cand_bits(const uint32_t* __restrict__ size, const unsigned long long* __restrict__ roots,
                            int64_t thr, uint32_t n, unsigned long long* __restrict__ cbits,
                            uint32_t* __restrict__ ccnt) {
    const uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave * kCandWords;
    const int lane = __lane_id();
    int64_t v[kCandWords];
#pragma unroll
    for (int j = 0; j < kCandWords; ++j) {
        const uint64_t p = (w0 + j) * kWave + (uint64_t)lane;
        v[j] = p < n ? (int64_t)size[p] : INT64_MIN;
        if (roots && p < n) v[j] += (int64_t)((roots[w0 + j] >> lane) & 1ull);
    }
#pragma unroll
    for (int j = 0; j < kCandWords; ++j) {
        const uint64_t b = __ballot(v[j] >= thr);
        if (lane == 0 && (w0 + j) * kWave < n) {
            cbits[w0 + j] = b;
            ccnt[w0 + j] = (uint32_t)__popcll(b);
        }
    }
}

// Halo number (from 1, in position order) of a candidate position q, 0 for
// any other: the candidates before q's word plus those below q in it.
__device__ __forceinline__ uint32_t halo_number(const unsigned long long* __restrict__ cbits,
                                                const uint32_t* __restrict__ cpre, uint32_t q) {
    // both loads issued together (the count is used only for a candidate)
    const unsigned long long w = cbits[q >> 6];
    const uint32_t pre = cpre[q >> 6];
    const unsigned long long bit = 1ull << (q & 63u);
    return (w & bit) ? pre + (uint32_t)__popcll(w & (bit - 1ull)) + 1u : 0u;
}

// in_halo / halo_id per particle (fof_fortran.f90:209-223); per candidate the
// anchor particle (its first CRS member) and its region.
__global__ void k_write(const int32_t* __restrict__ label, const unsigned long long* __restrict__ cbits,
                        const uint32_t* __restrict__ cpre,
                        const uint32_t* __restrict__ perm, const uint32_t* __restrict__ skeys,
                        const uint32_t* __restrict__ cell_off, int32_t nreg,
                        const RegionDev* __restrict__ reg, uint32_t n,
                        int32_t* __restrict__ halo_id,
                        int32_t* __restrict__ root_out, uint32_t* __restrict__ anchor,
                        int32_t* __restrict__ hregion) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t r = label[p];
    const uint32_t h = halo_number(cbits, cpre, (uint32_t)r);
    const uint32_t i = perm[p];
    // the one scattered write per halo member (halo_id was zero-filled,
    // coalesced, before); in_halo follows from halo_id in a coalesced pass
    // (k_in_halo)
    if (h) halo_id[i] = (int32_t)h;
    const uint32_t rg = nreg == 1 ? 0u : region_of_cell(cell_off, nreg, skeys[p]);
    if (root_out) root_out[i] = reg[rg].skip ? -1 : (int32_t)perm[r];
    // candidates are the positions whose component size (0 for non-roots)
    // reaches np_min (fof_fortran.f90:866-872), numbered in position order
    const uint32_t hp = halo_number(cbits, cpre, p);
    if (hp) {
        anchor[hp - 1] = i;
        hregion[hp - 1] = (int32_t)rg;
    }
}

// four entries per thread when both arrays are 16-byte aligned (device
// outputs of the caller may not be)
__global__ void k_in_halo(const int32_t* __restrict__ halo_id, uint32_t n, int32_t* __restrict__ in_halo) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool vec = (((uintptr_t)halo_id | (uintptr_t)in_halo) & 15u) == 0;
    if (vec) {
        const uint64_t i = 4 * t;
        if (i + 4 <= n) {
            const int4 h = reinterpret_cast<const int4*>(halo_id)[t];
            reinterpret_cast<int4*>(in_halo)[t] = make_int4(h.x > 0, h.y > 0, h.z > 0, h.w > 0);
        } else {
            for (uint64_t j = i; j < n; ++j) in_halo[j] = halo_id[j] > 0 ? 1 : 0;
        }
    } else {
        for (uint64_t j = 4 * t; j < 4 * t + 4 && j < n; ++j) in_halo[j] = halo_id[j] > 0 ? 1 : 0;
    }
}

// anchor and region of each valid halo under its new (bound) number
__global__ void k_remap_valid(const uint32_t* __restrict__ vflag, const uint32_t* __restrict__ vidx,
                              const uint32_t* __restrict__ anchor, const int32_t* __restrict__ hreg,
                              uint32_t nh, uint32_t* __restrict__ anchor2, int32_t* __restrict__ hreg2) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh || !vflag[h]) return;
    anchor2[vidx[h]] = anchor[h];
    hreg2[vidx[h]] = hreg[h];
}

// catalog: in-halo particles compacted in index order, keyed by halo id - 1
__global__ void k_cat_flags(const int32_t* __restrict__ in_halo, uint32_t n, uint32_t* __restrict__ f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = in_halo[i] ? 1u : 0u;
}
__global__ void k_cat_compact(const int32_t* __restrict__ halo_id, const uint32_t* __restrict__ pos,
                              uint32_t n, uint32_t* __restrict__ ckeys, uint32_t* __restrict__ cvals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t h = halo_id[i];
    if (h > 0) {
        const uint32_t j = pos[i];
        ckeys[j] = (uint32_t)(h - 1);
        cvals[j] = i;
    }
}

// Angular momentum of each halo about its f64 centre of mass, relative
// coordinates wrapped about the anchor (fof_fortran.f90:224-248). One block
// per halo; fixed reduction tree, so the result does not depend on timing.
__device__ inline void block_sum7(double v[7], double (*lds)[kBlk / kWave]) {
    for (int k = 0; k < 7; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o, kWave);
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int k = 0; k < 7; ++k) lds[k][w] = v[k];
    __syncthreads();
    for (int k = 0; k < 7; ++k) {
        double s = 0.0;
        for (int j = 0; j < kBlk / kWave; ++j) s += lds[k][j];
        v[k] = s;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kBlk) k_halo_angmom(const double* __restrict__ xp,
                                                      const double* __restrict__ vp,
                                                      const double* __restrict__ mp,
                                                      const uint32_t* __restrict__ hoff,
                                                      const uint32_t* __restrict__ members,
                                                      const uint32_t* __restrict__ anchor,
                                                      double* __restrict__ am_h, double L,
                                                      double Lh) {
    __shared__ double lds[7][kBlk / kWave];
    const uint32_t h = blockIdx.x;
    const uint32_t s = h == 0 ? 0u : hoff[h - 1], e = hoff[h];
    if (s >= e) {  // an empty candidate (np_halo_min <= 0) adds nothing
        if (threadIdx.x < 3) am_h[3 * (size_t)h + threadIdx.x] = 0.0;
        return;
    }
    const size_t a = anchor[h];
    const double xa[3] = {xp[3 * a], xp[3 * a + 1], xp[3 * a + 2]};
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t t = s + threadIdx.x; t < e; t += kBlk) {
        const size_t i = members[t];
        const double w = mp ? mp[i] : 1.0;
        v[0] += w;
        for (int d = 0; d < 3; ++d) {
            v[1 + d] += w * rel_coord(xp[3 * i + d], xa[d], L, Lh);
            v[4 + d] += w * vp[3 * i + d];
        }
    }
    block_sum7(v, lds);
    if (!(v[0] > 0.0)) {
        if (threadIdx.x < 3) am_h[3 * (size_t)h + threadIdx.x] = 0.0;
        return;
    }
    const double qc[3] = {v[1] / v[0], v[2] / v[0], v[3] / v[0]};
    const double vc[3] = {v[4] / v[0], v[5] / v[0], v[6] / v[0]};
    double u[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t t = s + threadIdx.x; t < e; t += kBlk) {
        const size_t i = members[t];
        const double w = mp ? mp[i] : 1.0;
        double dx[3], dv[3];
        for (int d = 0; d < 3; ++d) {
            dx[d] = rel_coord(xp[3 * i + d], xa[d], L, Lh) - qc[d];
            dv[d] = vp[3 * i + d] - vc[d];
        }
        u[0] += w * (dx[1] * dv[2] - dx[2] * dv[1]);
        u[1] += w * (dx[2] * dv[0] - dx[0] * dv[2]);
        u[2] += w * (dx[0] * dv[1] - dx[1] * dv[0]);
    }
    block_sum7(u, lds);
    if (threadIdx.x < 3) am_h[3 * (size_t)h + threadIdx.x] = u[threadIdx.x];
}

// Per-region sums of the halo moments; a region's halos are contiguous in
// halo-id order. One block per region, fixed tree.
__global__ void __launch_bounds__(kBlk) k_region_angmom(const double* __restrict__ am_h,
                                                        const int32_t* __restrict__ hregion,
                                                        uint32_t nh, double* __restrict__ am) {
    __shared__ double lds[7][kBlk / kWave];
    const int32_t r = blockIdx.x;
    // first halo with region >= r, first with region > r
    uint32_t lo = 0, hi = nh;
    while (lo < hi) { const uint32_t mid = (lo + hi) / 2; if (hregion[mid] < r) lo = mid + 1; else hi = mid; }
    const uint32_t s = lo;
    hi = nh;
    while (lo < hi) { const uint32_t mid = (lo + hi) / 2; if (hregion[mid] <= r) lo = mid + 1; else hi = mid; }
    const uint32_t e = lo;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t h = s + threadIdx.x; h < e; h += kBlk)
        for (int d = 0; d < 3; ++d) v[d] += am_h[3 * (size_t)h + d];
    block_sum7(v, lds);
    if (threadIdx.x < 3) am[3 * (size_t)r + threadIdx.x] = v[threadIdx.x];
}

// ---------------------------------------------------------------------------
// host side

void fof_init_device_tables() {
        // once per device (contexts may be created from several threads)
        static std::mutex mu;
        static std::vector<bool> done;
        std::lock_guard<std::mutex> lk(mu);
        int dev = 0;
        PYFOF_HIP(hipGetDevice(&dev));
        if ((int)done.size() <= dev) done.resize(dev + 1, false);
        if (done[dev]) return;
        // forward offsets of [-2,2]^3, nearest first (|o|^2 = 1, 2, 3, 4, ...)
        int8_t off[62][3];
        int n = 0;
        for (int r2 = 1; r2 <= 12; ++r2)
            for (int o = 63; o < 125; ++o) {
                const int dx = o % 5 - 2, dy = (o / 5) % 5 - 2, dz = o / 25 - 2;
                if (dx * dx + dy * dy + dz * dz != r2) continue;
                off[n][0] = (int8_t)dx;
                off[n][1] = (int8_t)dy;
                off[n][2] = (int8_t)dz;
                ++n;
            }
        PYFOF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_fwd_off), off, sizeof(off)));
        done[dev] = true;
}

static thread_local std::string g_last_error;

static void set_error(const std::string& m) { g_last_error = m; }

void set_last_error(const std::string& m) { g_last_error = m; }

const std::string& last_error_message() { return g_last_error; }

static inline uint32_t nblk(uint64_t n) { return div_up(n == 0 ? 1 : n, kBlk); }


static int32_t run_fof(Ctx& C, const CallArgs& A) {
    hipStream_t s = C.stream;
    const uint32_t N = (uint32_t)A.total_np;
    const int32_t H = A.nhalo;
    const bool dev_mem = A.mem_kind == PYFOF_MEM_DEVICE;
    for (auto& r : C.ev_recorded) r = false;
    C.cat_nh = 0;
    C.cat_nmem = 0;
    C.cat_hoff = nullptr;
    C.cat_members = nullptr;
    C.last_bound = BoundStats();
    C.fft_reset();
    C.mark(kEvStart);

    // inputs in HBM
    const double* xp = A.xp;
    const double* vp = A.vp;
    const double* mp = A.mp;
    if (!dev_mem) {
        double* d = C.xp.as<double>(3 * (size_t)N);
        PYFOF_HIP(hipMemcpyAsync(d, A.xp, 24 * (size_t)N, hipMemcpyHostToDevice, s));
        xp = d;
        if (A.vp && (A.ang_mom || A.bound_check)) {
            double* dv = C.vp.as<double>(3 * (size_t)N);
            PYFOF_HIP(hipMemcpyAsync(dv, A.vp, 24 * (size_t)N, hipMemcpyHostToDevice, s));
            vp = dv;
        }
        if (A.mp && (A.ang_mom || A.bound_check)) {
            double* dm = C.mp.as<double>((size_t)N);
            PYFOF_HIP(hipMemcpyAsync(dm, A.mp, 8 * (size_t)N, hipMemcpyHostToDevice, s));
            mp = dm;
        }
    }

    // regions and chunks (host)
    constexpr int32_t kChunk = kKeyChunk;  // k_keys counts a chunk's keys in at most two sort tiles
    std::vector<int32_t> start(H);
    std::vector<Chunk> chunks;
    {
        int64_t acc = 0;
        for (int32_t h = 0; h < H; ++h) {
            start[h] = (int32_t)acc;
            const int32_t n = A.np_halo[h];
            for (int32_t b0 = 0; b0 < n; b0 += kChunk)
                chunks.push_back(Chunk{h, (int32_t)acc + b0, (int32_t)acc + std::min(n, b0 + kChunk)});
            acc += n;
        }
    }
    const uint32_t nch = (uint32_t)chunks.size();
    Chunk* d_chunks = C.chunks.as<Chunk>(nch);
    PYFOF_HIP(hipMemcpyAsync(d_chunks, chunks.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice, s));
    int32_t* d_start = C.regstart.as<int32_t>((size_t)H + 1);
    PYFOF_HIP(hipMemcpyAsync(d_start, start.data(), sizeof(int32_t) * H, hipMemcpyHostToDevice, s));
    unsigned long long* d_bb = C.bbox.as<unsigned long long>(6 * (size_t)H);
    {
        std::vector<unsigned long long> init(6 * (size_t)H);
        for (int32_t h = 0; h < H; ++h)
            for (int d = 0; d < 3; ++d) {
                init[6 * h + d] = dbl_to_ordered(INFINITY);
                init[6 * h + 3 + d] = dbl_to_ordered(-INFINITY);
            }
        PYFOF_HIP(hipMemcpyAsync(d_bb, init.data(), 8 * init.size(), hipMemcpyHostToDevice, s));
        const double Lh = 0.5 * A.L;
        k_bbox<<<nch, kBlk, 0, s>>>(xp, d_chunks, d_start, d_bb, A.L, Lh);
        PYFOF_LAUNCH_CHECK();
        std::vector<unsigned long long> bb(6 * (size_t)H);
        PYFOF_HIP(hipMemcpyAsync(bb.data(), d_bb, 8 * bb.size(), hipMemcpyDeviceToHost, s));
        PYFOF_HIP(hipStreamSynchronize(s));
        init.swap(bb);
        // geometry per region (host), reference rule + sub-cell plan
        std::vector<RegionDev> regs(H);
        std::vector<uint32_t> cell_off(H + 1);
        uint64_t ncell = 0;
        int32_t max_sub = 1;
        double amax = 0.0;
        for (int32_t h = 0; h < H; ++h) {
            for (int d = 0; d < 3; ++d)
                if (A.np_halo[h] > 0)
                    amax = std::max(amax, std::max(fabs(ordered_to_dbl(init[6 * h + d])),
                                                   fabs(ordered_to_dbl(init[6 * h + 3 + d]))));
        }
        // rounding margin: a few float ulps at the largest coordinate plus a
        // relative slack on b; covers every float rounding of the predicate
        const double m = 8.0 * std::max(ulp_f32(amax), ulp_f32(A.L)) + 1e-7 * A.b;
        for (int32_t h = 0; h < H; ++h) {
            RegionDev& g = regs[h];
            memset(&g, 0, sizeof(g));
            g.start = start[h];
            g.np = A.np_halo[h];
            g.skip = !(g.np >= A.np_min && g.np > 0);
            if (g.skip) {
                for (int d = 0; d < 3; ++d) { g.n[d] = 1; g.k[d] = 1; g.cell[d] = 1.0; g.sub[d] = 1.0; }
                g.cell_off = (uint32_t)ncell;
                g.ncell = 1;
                cell_off[h] = (uint32_t)ncell;
                ncell += 1;
                continue;
            }
            double bmin[3], bmax[3];
            for (int d = 0; d < 3; ++d) {
                bmin[d] = ordered_to_dbl(init[6 * h + d]);
                bmax[d] = ordered_to_dbl(init[6 * h + 3 + d]);
            }
            reference_geometry(bmin, bmax, g.np, A.L, A.b, g);
            g.prune = 1;
            for (int d = 0; d < 3; ++d)
                if (g.periodic[d] && g.n[d] < 4) g.prune = 0;
            // sub-cells: side <= (b - m)/sqrt(3) makes each sub-cell a clique
            const double smax = (A.b - 2.0 * m) / sqrt(3.0);
            g.clique = smax > 0.0 ? 1 : 0;
            for (int d = 0; d < 3; ++d) {
                int32_t k = 1;
                if (smax > 0.0) {
                    const double kk = ceil(g.cell[d] / smax * (1.0 + 1e-12));
                    k = kk > kMaxSubPerDim ? kMaxSubPerDim : (int32_t)kk;
                    if (kk > kMaxSubPerDim) g.clique = 0;
                }
                g.k[d] = k < 1 ? 1 : k;
                g.sub[d] = g.cell[d] / (double)g.k[d];
                if (g.sub[d] > smax) g.clique = 0;
            }
            if (!g.prune) { g.k[0] = g.k[1] = g.k[2] = 1; g.clique = 0; }
            const uint64_t nc = (uint64_t)g.n[0] * (uint64_t)g.n[1] * (uint64_t)g.n[2];
            g.cell_off = (uint32_t)ncell;
            g.ncell = (uint32_t)nc;
            cell_off[h] = (uint32_t)ncell;
            ncell += nc;
            if (ncell >= 0xffffffffull) {
                set_error("total cell count exceeds 2^32 - 1");
                return IERR_TOO_LARGE;
            }
            max_sub = std::max(max_sub, g.k[0] * g.k[1] * g.k[2]);
        }
        cell_off[H] = (uint32_t)ncell;
        RegionDev* d_reg = C.regions.as<RegionDev>(H);
        PYFOF_HIP(hipMemcpyAsync(d_reg, regs.data(), sizeof(RegionDev) * H, hipMemcpyHostToDevice, s));
        uint32_t* d_coff = C.cell_off.as<uint32_t>((size_t)H + 1);
        PYFOF_HIP(hipMemcpyAsync(d_coff, cell_off.data(), sizeof(uint32_t) * (H + 1),
                                 hipMemcpyHostToDevice, s));
        C.mark(kEvBbox);

        LinkParams P;
        P.L = A.L;
        P.Lh = 0.5 * A.L;
        P.b = A.b;
        P.rp2 = A.b * A.b;
        P.m = m;
        P.R = A.b + m;
        P.scr_lo = (float)(P.rp2 * (1.0 - 1e-6));
        P.scr_hi = (float)(P.rp2 * (1.0 + 1e-6));
        P.scr_wrap = (float)(0.5 * P.Lh);
        P.np_min = A.np_min;
        P.nregions = H;
        P.N = N;
        P.ncell = (uint32_t)ncell;
        P.sub_stride = max_sub + 1;
        P.intra_stage = 0;

        // ---- binning: keys, stable sort, cell_ptr, float CRS coordinates
        uint32_t* k0 = C.keys0.as<uint32_t>(N);
        uint32_t* k1 = C.keys1.as<uint32_t>(N);
        uint32_t* v0 = C.vals0.as<uint32_t>(N);
        uint32_t* v1 = C.vals1.as<uint32_t>(N);
        float4* xr = C.buf("xrel").as<float4>(N);
        const size_t tmp_bytes = std::max(radix_temp_bytes(N), scan_temp_bytes(ncell + 1));
        void* tmp = C.temp.get(std::max(tmp_bytes, scan_temp_bytes(N) + 64));
        // the sort's first histogram counted by k_keys
        const int kbits = bits_for(ncell - 1), kwidth = radix_first_width(kbits);
        uint32_t* khist = nullptr;
        uint32_t ktiles = 0;
        if (N > 1 && kwidth > 0) {
            ktiles = (uint32_t)((N + radix_tile_keys - 1) / radix_tile_keys);
            khist = radix_first_hist(tmp);
            k_fill_u32<<<nblk((uint64_t)ktiles << kwidth), kBlk, 0, s>>>(khist, (uint64_t)ktiles << kwidth, 0u);
            PYFOF_LAUNCH_CHECK();
        }
        k_keys<<<nch, kBlk, 0, s>>>(xp, d_chunks, d_reg, k0, xr, A.L, P.Lh, P.R, P.m, khist, ktiles,
                                    kwidth > 0 ? (1u << kwidth) - 1u : 0u);
        PYFOF_LAUNCH_CHECK();
        const bool alt = radix_sort_pairs(k0, v0, k1, v1, N, kbits, tmp, s, true, khist != nullptr);
        uint32_t* skeys = alt ? k1 : k0;
        uint32_t* perm = alt ? v1 : v0;
        // the gather of the float coordinates (random reads) runs on stream2
        // beside the cell starts and classes (streaming); the dense-cell
        // build, the first user of xc and parent, waits for it
        hipStream_t s2 = C.stream2, s3 = C.stream3;
        float4* xc = C.xc.as<float4>(N);
        int32_t* parent = C.parent.as<int32_t>(N);
        PYFOF_HIP(hipEventRecord(C.ev_sort, s));
        PYFOF_HIP(hipStreamWaitEvent(s2, C.ev_sort, 0));
        uint8_t* fmask = C.buf("fmask").as<uint8_t>(ncell);
        k_gather<<<nblk(N), kBlk, 0, s2>>>(xr, perm, skeys, xc, parent, fmask, N);
        PYFOF_LAUNCH_CHECK();
        PYFOF_HIP(hipEventRecord(C.ev_gather, s2));
        uint32_t* scal = C.scal.as<uint32_t>(64);
        k_fill_u32<<<1, 64, 0, s>>>(scal, 64, 0u);
        PYFOF_LAUNCH_CHECK();
        uint32_t* cell_ptr = C.ends.as<uint32_t>(ncell + 1);
        uint4* gaps = C.buf("gaps").as<uint4>(gap_records(ncell));
        k_cell_starts<<<nblk(((uint64_t)N + 1 + 3) / 4), kBlk, 0, s>>>(skeys, N, ncell, cell_ptr, gaps, scal + 16);
        PYFOF_LAUNCH_CHECK();
        k_fill_gaps<<<1024, kBlk, 0, s>>>(gaps, scal + 16, cell_ptr);
        PYFOF_LAUNCH_CHECK();
        // dense cells and their sub-cell lists
        int32_t* aux = C.aux.as<int32_t>(ncell);
        uint32_t* dense_cells = C.dense.as<uint32_t>(N / kDenseMin + 1);
        uint32_t* mid_cells = C.mid.as<uint32_t>(N / kMidMin + 1);
        const uint64_t mark_cells = (uint64_t)kMarkBlk * kMarkPer;
        k_dense_mark<<<(uint32_t)((ncell + mark_cells - 1) / mark_cells), kMarkBlk, 0, s>>>(
            cell_ptr, (uint32_t)ncell, d_coff, d_reg, H, aux, dense_cells, scal, mid_cells, scal + 8);
        PYFOF_LAUNCH_CHECK();
        uint32_t cnts[9];
        PYFOF_HIP(hipMemcpyAsync(cnts, scal, sizeof(cnts), hipMemcpyDeviceToHost, s));
        PYFOF_HIP(hipStreamSynchronize(s));
        const uint32_t ndense = cnts[0], nmid = cnts[8];
        if (ndense > 0) {
            k_mark_near_dense<<<nblk(ndense), kBlk, 0, s>>>(dense_cells, ndense, d_coff, d_reg, H, aux, A.L);
            PYFOF_LAUNCH_CHECK();
        }
        PYFOF_HIP(hipStreamWaitEvent(s, C.ev_gather, 0));
        C.mark(kEvBin);

        // ---- linking. k_dense_intra and k_cell_intra write the parents of
        // dense-clique and mid-cell members directly (local components), so
        // every atomic union that may reach those members waits for them.
        // Three streams: the sparse-particle kernels on s2, the mid-cell
        // kernels on s3, the dense-cell kernels on s. k_link touches neither
        // kind of member (it joins sparse particles, and the dense-cell
        // members of regions without cliques, which no plain store writes),
        // and k_cell_intra only mid-cell members, so both start once the
        // cell classes and coordinates are in place, beside k_dense_build
        // (which writes the parents of dense-clique members only). The
        // dense-cell pair kernels join dense clique members only, whose
        // union-find paths stay among dense members until k_cell_pair or
        // k_link_dense joins another kind: they wait for k_dense_intra alone.
        // k_cell_pair and k_link_dense join mid or sparse members to dense
        // ones and wait for both intra-cell kernels.
        // per-particle masks of dense cells to scan, in the buffer that holds
        // the labels after linking
        uint32_t* dmask = reinterpret_cast<uint32_t*>(C.label.as<int32_t>(N));
        PYFOF_HIP(hipEventRecord(C.ev_fork, s));
        PYFOF_HIP(hipStreamWaitEvent(s2, C.ev_fork, 0));
        PYFOF_HIP(hipStreamWaitEvent(s3, C.ev_fork, 0));
        // the relabel's component sizes and halo ids, cleared by k_link
        uint32_t* size = C.size.as<uint32_t>(N);
        int32_t* o_id = dev_mem ? A.halo_id : C.out_id.as<int32_t>(N);
        k_link<<<nblk(N), kBlk, 0, s2>>>(xc, skeys, cell_ptr, aux, fmask, d_reg, d_coff, parent, dmask, size, o_id, P);
        PYFOF_LAUNCH_CHECK();
        if (nmid > 0) {
            k_cell_intra<<<nmid, kWave, kIntraLdsPad, s3>>>(mid_cells, cell_ptr, d_coff, d_reg, H, xc, parent, P);
            PYFOF_LAUNCH_CHECK();
        }
        PYFOF_HIP(hipEventRecord(C.ev_mid, s3));

        // ---- dense cells' sub-cell lists
        uint32_t* sub_ptr = C.sub_ptr.as<uint32_t>((size_t)std::max<uint32_t>(ndense, 1) * P.sub_stride);
        uint32_t* sub_list = C.sub_list.as<uint32_t>(N);
        uint4* subs = C.subs.as<uint4>(ndense > 0 ? N : 1);
        uint32_t nsubs = 0;
        if (ndense > 0) {
            k_dense_build<<<ndense, kBlk, 2 * (size_t)P.sub_stride * sizeof(uint32_t), s>>>(
                dense_cells, cell_ptr, d_coff, d_reg, H, xc, sub_ptr,
                                                  P.sub_stride, sub_list, parent, subs, scal + 3);
            PYFOF_LAUNCH_CHECK();
            PYFOF_HIP(hipMemcpyAsync(&nsubs, scal + 3, 4, hipMemcpyDeviceToHost, s));
            PYFOF_HIP(hipStreamSynchronize(s));
        }
        C.last_n = N;
        C.last_ncell = ncell;
        C.last_ndense = ndense;

        int32_t* ncomp = C.ncomp.as<int32_t>(std::max<uint32_t>(ndense, 1));
        int32_t* croot = C.croot.as<int32_t>(std::max<uint32_t>(ndense, 1));
        int32_t* needs_sub = C.needs_sub.as<int32_t>(std::max<uint32_t>(ndense, 1));
        int32_t* sub_root = C.sub_root.as<int32_t>((size_t)std::max<uint32_t>(ndense, 1) * P.sub_stride);
        if (nsubs > 0) {
            // LDS: sub-cell tables, plus staged positions within a 64 KiB budget
            const size_t base = (size_t)16 * ((4 * P.sub_stride + 2 + 3) / 4);
            P.intra_stage = (int32_t)std::min<size_t>(kIntraStage, base < 65536 ? (65536 - base) / 16 : 0);
            const size_t lds = base + 16 * (size_t)P.intra_stage;
            k_dense_intra<<<ndense, kBlk, lds, s>>>(dense_cells, cell_ptr, d_coff, d_reg, H, sub_ptr,
                                                    sub_list, xc, parent, ncomp, croot, needs_sub, sub_root,
                                                    P);
            PYFOF_LAUNCH_CHECK();
        }
        PYFOF_HIP(hipEventRecord(C.ev_dense, s));
        PYFOF_HIP(hipStreamWaitEvent(s2, C.ev_dense, 0));
        PYFOF_HIP(hipStreamWaitEvent(s2, C.ev_mid, 0));
        if (ndense > 0) {
            k_link_dense<<<nblk(((uint64_t)N + 3) / 4), kBlk, 0, s2>>>(xc, skeys, cell_ptr, aux, sub_ptr, sub_list, ncomp, croot,
                                                   d_reg, d_coff, parent, dmask, P);
            PYFOF_LAUNCH_CHECK();
        }
        PYFOF_HIP(hipEventRecord(C.ev_join, s2));
        PYFOF_HIP(hipStreamWaitEvent(s3, C.ev_dense, 0));
        if (nmid > 0) {
            k_cell_pair<<<nmid, kWave, 0, s3>>>(mid_cells, cell_ptr, d_coff, d_reg, H, aux, fmask, sub_ptr, sub_list,
                                                xc, parent, P);
            PYFOF_LAUNCH_CHECK();
        }
        PYFOF_HIP(hipEventRecord(C.ev_pair, s3));
        if (nsubs > 0) {
            k_dense_cell_cross<<<ndense, kWave, 0, s>>>(dense_cells, ndense, cell_ptr, d_coff, d_reg, H,
                                                        aux, ncomp, croot, needs_sub, sub_ptr, sub_list,
                                                        xc, parent, P);
            PYFOF_LAUNCH_CHECK();
            k_dense_cross<<<nblk(nsubs), kBlk, 0, s>>>(subs, nsubs, dense_cells, cell_ptr, aux, ncomp,
                                                       croot, needs_sub, sub_ptr, sub_list, sub_root, d_reg,
                                                       xc, parent, P);
            PYFOF_LAUNCH_CHECK();
        }
        PYFOF_HIP(hipStreamWaitEvent(s, C.ev_pair, 0));
        PYFOF_HIP(hipStreamWaitEvent(s, C.ev_join, 0));
        C.mark(kEvLink);

        // ---- roots, sizes, candidate numbering, outputs
        int32_t* label = C.label.as<int32_t>(N);
        unsigned long long* roots = C.buf("roots").as<unsigned long long>(((uint64_t)N + kWave - 1) / kWave);
        k_label_count<<<nblk(((uint64_t)N + kCountChunks - 1) / kCountChunks), kBlk, 0, s>>>(parent, label, N, size,
                                                                                              roots);
        PYFOF_LAUNCH_CHECK();
        // candidates (k_cand_bits), numbered from 1 in position order
        const uint64_t nwords = ((uint64_t)N + kWave - 1) / kWave;
        unsigned long long* cbits = C.buf("cbits").as<unsigned long long>(nwords);
        uint32_t* cpre = C.buf("ccnt").as<uint32_t>(nwords);
        const uint32_t cand_blocks = nblk((nwords + kCandWords - 1) / kCandWords * kWave);
        if (A.np_min >= 2)
            k_cand_bits<<<cand_blocks, kBlk, 0, s>>>(size, nullptr, (int64_t)A.np_min - 1, N, cbits, cpre);
        else
            k_cand_bits<<<cand_blocks, kBlk, 0, s>>>(size, roots, A.np_min, N, cbits, cpre);
        PYFOF_LAUNCH_CHECK();
        exclusive_scan_u32(cpre, cpre, nwords, scal + 1, tmp, s);
        int32_t* o_in = dev_mem ? A.in_halo : C.out_in.as<int32_t>(N);
        int32_t* o_root = nullptr;
        if (A.root_out) o_root = dev_mem ? A.root_out : C.out_root.as<int32_t>(N);
        uint32_t* anchor = C.anchor.as<uint32_t>(N);
        int32_t* hreg = C.hregion.as<int32_t>(N);
        k_write<<<nblk(N), kBlk, 0, s>>>(label, cbits, cpre, perm, skeys, d_coff, H, d_reg, N, o_id,
                                         o_root, anchor, hreg);
        PYFOF_LAUNCH_CHECK();
        k_in_halo<<<nblk(((uint64_t)N + 3) / 4), kBlk, 0, s>>>(o_id, N, o_in);
        PYFOF_LAUNCH_CHECK();
        uint32_t nh = 0;
        PYFOF_HIP(hipMemcpyAsync(&nh, scal + 1, 4, hipMemcpyDeviceToHost, s));
        C.mark(kEvOut);

        // ---- catalog (members grouped by halo, ascending index)
        PYFOF_HIP(hipStreamSynchronize(s));
        C.cat_nh = nh;
        struct Cat {
            uint32_t nmem = 0;
            uint32_t* members = nullptr;
            uint32_t* mkeys = nullptr;
            uint32_t* hoff = nullptr;
        };
        auto build_catalog = [&](uint32_t nhalo) {
            Cat c;
            uint32_t* cf = C.flag.as<uint32_t>(N);
            k_cat_flags<<<nblk(N), kBlk, 0, s>>>(o_in, N, cf);
            PYFOF_LAUNCH_CHECK();
            exclusive_scan_u32(cf, cf, N, scal + 2, tmp, s);
            PYFOF_HIP(hipMemcpyAsync(&c.nmem, scal + 2, 4, hipMemcpyDeviceToHost, s));
            PYFOF_HIP(hipStreamSynchronize(s));
            uint32_t* ck = C.cat_keys.as<uint32_t>(c.nmem);
            uint32_t* cv = C.cat_vals.as<uint32_t>(c.nmem);
            uint32_t* ck2 = C.cat_keys2.as<uint32_t>(c.nmem);
            uint32_t* cv2 = C.cat_vals2.as<uint32_t>(c.nmem);
            k_cat_compact<<<nblk(N), kBlk, 0, s>>>(o_id, cf, N, ck, cv);
            PYFOF_LAUNCH_CHECK();
            const bool calt = radix_sort_pairs(ck, cv, ck2, cv2, c.nmem, bits_for(nhalo - 1), tmp, s);
            c.mkeys = calt ? ck2 : ck;
            c.members = calt ? cv2 : cv;
            c.hoff = C.hoff.as<uint32_t>(nhalo);
            k_fill_u32<<<nblk(nhalo), kBlk, 0, s>>>(c.hoff, nhalo, 0u);
            PYFOF_LAUNCH_CHECK();
            k_run_ends<<<nblk(c.nmem), kBlk, 0, s>>>(c.mkeys, c.nmem, c.hoff);
            PYFOF_LAUNCH_CHECK();
            inclusive_max_scan_u32(c.hoff, c.hoff, nhalo, tmp, s);
            return c;
        };
        uint32_t nh_out = nh;  // halos in the output numbering
        if (A.bound_check && nh > 0) {
            // ---- unbinding of every candidate (fof_fortran.f90:253-356)
            Cat cand = build_catalog(nh);
            BoundArgs B;
            B.app = A.softening; B.a = A.a_val; B.b = A.b_val; B.s15 = A.s_15; B.ng = A.ng;
            B.np_min = A.np_min; B.verbose = A.verbose; B.L = A.L; B.Lh = P.Lh; B.ng_global = A.ng_global;
            B.xp = xp; B.vp = vp; B.mp = mp; B.N = N;
            B.nh = nh; B.nmem = cand.nmem; B.hoff = cand.hoff; B.members = cand.members;
            B.mhalo = cand.mkeys; B.anchor = anchor; B.in_halo = o_in; B.halo_id = o_id;
            B.direct_max = C.direct_max;
            B.direct_f32 = C.direct_f32;
            B.batched = C.batched;
            B.multi_wg = C.solo;
            C.rec_n = 0;
            if (C.record_phi) {
                B.phi_rec = C.rec_phi.as<float>(std::max<uint32_t>(cand.nmem, 1));
                uint32_t* ri = C.rec_idx.as<uint32_t>(std::max<uint32_t>(cand.nmem, 1));
                PYFOF_HIP(hipMemcpyAsync(ri, cand.members, 4 * (size_t)cand.nmem, hipMemcpyDeviceToDevice, s));
                C.rec_n = cand.nmem;
            }
            uint8_t* bkeep = C.buf("bkeep_out").as<uint8_t>(std::max<uint32_t>(cand.nmem, 1));
            BoundResult br;
            const int32_t berr = run_bound(C, B, bkeep, br);
            C.last_bound = br;
            if (berr != IERR_OK) {
                set_error("unbinding failed: " + br.err_msg);
                return berr;
            }
            // anchors and regions of the valid halos, in their new numbering
            k_remap_valid<<<nblk(nh), kBlk, 0, s>>>(C.buf("vflag").as<uint32_t>(nh), C.buf("vidx").as<uint32_t>(nh),
                                                    anchor, hreg, nh, C.buf("anchor2").as<uint32_t>(nh),
                                                    C.buf("hreg2").as<int32_t>(nh));
            PYFOF_LAUNCH_CHECK();
            PYFOF_HIP(hipMemcpyAsync(C.anchor.as<uint32_t>(N), C.buf("anchor2").as<uint32_t>(nh), 4 * (size_t)nh,
                                     hipMemcpyDeviceToDevice, s));
            PYFOF_HIP(hipMemcpyAsync(C.hregion.as<int32_t>(N), C.buf("hreg2").as<int32_t>(nh), 4 * (size_t)nh,
                                     hipMemcpyDeviceToDevice, s));
            nh_out = br.nvalid;
            C.cat_nh = nh_out;
        }
        C.mark(kEvBound);
        const bool need_groups = C.want_catalog || (A.ang_mom && vp);
        if (nh_out > 0 && need_groups) {
            Cat cat = build_catalog(nh_out);
            C.cat_nmem = cat.nmem;
            C.cat_members = cat.members;
            C.cat_hoff = cat.hoff;
            if (A.ang_mom && vp) {
                double* am_h = C.am_h.as<double>(3 * (size_t)nh_out);
                double* am = dev_mem ? A.ang_mom : C.am.as<double>(3 * (size_t)H);
                k_halo_angmom<<<nh_out, kBlk, 0, s>>>(xp, vp, mp, cat.hoff, cat.members, anchor, am_h, A.L, P.Lh);
                PYFOF_LAUNCH_CHECK();
                k_region_angmom<<<H, kBlk, 0, s>>>(am_h, hreg, nh_out, am);
                PYFOF_LAUNCH_CHECK();
                if (!dev_mem)
                    PYFOF_HIP(hipMemcpyAsync(A.ang_mom, am, 24 * (size_t)H, hipMemcpyDeviceToHost, s));
            } else if (A.ang_mom) {
                if (dev_mem) PYFOF_HIP(hipMemsetAsync(A.ang_mom, 0, 24 * (size_t)H, s));
                else memset(A.ang_mom, 0, 24 * (size_t)H);
            }
        } else if (A.ang_mom) {
            if (dev_mem) PYFOF_HIP(hipMemsetAsync(A.ang_mom, 0, 24 * (size_t)H, s));
            else memset(A.ang_mom, 0, 24 * (size_t)H);
        }
        C.mark(kEvCat);
        if (!dev_mem) {
            PYFOF_HIP(hipMemcpyAsync(A.in_halo, o_in, 4 * (size_t)N, hipMemcpyDeviceToHost, s));
            PYFOF_HIP(hipMemcpyAsync(A.halo_id, o_id, 4 * (size_t)N, hipMemcpyDeviceToHost, s));
            if (A.root_out)
                PYFOF_HIP(hipMemcpyAsync(A.root_out, o_root, 4 * (size_t)N, hipMemcpyDeviceToHost, s));
        }
        PYFOF_HIP(hipStreamSynchronize(s));
        if (C.time_fft) C.fft_collect();
        if (A.verbose >= 2)
            printf("[pyfof] N=%u regions=%d cells=%llu dense=%u halos=%u\n", N, H,
                   (unsigned long long)ncell, ndense, nh);
    }
    return IERR_OK;
}

static std::mutex g_ctx_mu;
static std::vector<Ctx*> g_default_ctx;

static Ctx* default_ctx() {
    int dev = 0;
    PYFOF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if ((int)g_default_ctx.size() <= dev) g_default_ctx.resize(dev + 1, nullptr);
    if (!g_default_ctx[dev]) g_default_ctx[dev] = new Ctx(dev);
    return g_default_ctx[dev];
}

// fof_fortran.f90:110-113: in_halo, halo_id and ang_mom are zero whenever a
// call ends without a result (validation errors, ierr 10, empty input, a
// failed unbinding). Host arrays are cleared here; device arrays on the
// context's stream. Lengths are the caller's: total_np and 3 * nhalo.
void clear_outputs(Ctx* C, const CallArgs& A) {
    const size_t n = A.total_np > 0 ? (size_t)A.total_np : 0;
    const size_t h = A.nhalo > 0 ? (size_t)A.nhalo : 0;
    if (A.mem_kind == PYFOF_MEM_HOST) {
        if (A.in_halo && n) memset(A.in_halo, 0, 4 * n);
        if (A.halo_id && n) memset(A.halo_id, 0, 4 * n);
        if (A.ang_mom && h) memset(A.ang_mom, 0, 24 * h);
        return;
    }
    if (A.mem_kind != PYFOF_MEM_DEVICE) return;
    (void)hipSetDevice(C->device);
    if (A.in_halo && n) (void)hipMemsetAsync(A.in_halo, 0, 4 * n, C->stream);
    if (A.halo_id && n) (void)hipMemsetAsync(A.halo_id, 0, 4 * n, C->stream);
    if (A.ang_mom && h) (void)hipMemsetAsync(A.ang_mom, 0, 24 * h, C->stream);
    (void)hipStreamSynchronize(C->stream);
}

static int32_t validate(const CallArgs& A, int32_t n_threads) {
    if (n_threads > 256) {
        set_error("n_threads exceeds 256");
        return IERR_THREADS;
    }
    if (A.nhalo <= 0 || A.total_np <= 0) return IERR_OK;
    int64_t tot = 0;
    for (int32_t h = 0; h < A.nhalo; ++h) {
        if (A.np_halo[h] < 0) { set_error("negative region size"); return IERR_INVALID; }
        tot += A.np_halo[h];
    }
    if (tot != A.total_np) {
        set_error("sum(np_halo) != total_np");
        return IERR_INVALID;
    }
    if (!A.xp || !A.in_halo || !A.halo_id) {
        set_error("null xp/in_halo/halo_id");
        return IERR_INVALID;
    }
    if (!(A.L > 0.0) || !(A.b > 0.0)) {
        set_error("L_Box and linking_length must be positive");
        return IERR_INVALID;
    }
    if (A.bound_check && !(A.softening > 0.0)) {
        set_error("softening must be positive for bound_check");
        return IERR_INVALID;
    }
    return IERR_OK;
}

int32_t fof_entry(Ctx* C, const CallArgs& A, int32_t n_threads) {
    set_error("");
    if (A.mem_kind != PYFOF_MEM_HOST && A.mem_kind != PYFOF_MEM_DEVICE) {
        set_error("mem_kind must be PYFOF_MEM_HOST or PYFOF_MEM_DEVICE");
        return IERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(C->mu);
    int32_t err = validate(A, n_threads);
    if (err != IERR_OK || A.nhalo <= 0 || A.total_np <= 0) {
        const std::string msg = g_last_error;
        clear_outputs(C, A);
        set_error(msg);
        return err;
    }
    try {
        PYFOF_HIP(hipSetDevice(C->device));
        err = run_fof(*C, A);
    } catch (const HipError& e) {
        set_error(e.what());
        (void)hipStreamSynchronize(C->stream2);
        (void)hipStreamSynchronize(C->stream3);
        (void)hipStreamSynchronize(C->stream);
        err = e.code == hipErrorOutOfMemory ? IERR_ALLOC : IERR_HIP;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        err = IERR_ALLOC;
    } catch (const std::exception& e) {
        set_error(e.what());
        err = IERR_HIP;
    }
    if (err != IERR_OK) {
        const std::string msg = g_last_error;
        clear_outputs(C, A);
        set_error(msg);
    }
    return err;
}

}  // namespace pyfof

using namespace pyfof;

struct pyfof_ctx {
    Ctx* c;
};

extern "C" {

const char* pyfof_last_error(void) { return g_last_error.c_str(); }

const char* pyfof_version(void) { return "pyfof-mi355x 0.1 (gfx950)"; }

pyfof_ctx* pyfof_ctx_create(int32_t device) {
    try {
        auto* h = new pyfof_ctx;
        h->c = new Ctx(device);
        return h;
    } catch (const std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

void pyfof_ctx_destroy(pyfof_ctx* ctx) {
    if (!ctx) return;
    delete ctx->c;
    delete ctx;
}

void* pyfof_ctx_stream(pyfof_ctx* ctx) { return ctx ? (void*)ctx->c->stream : nullptr; }

int32_t pyfof_fof_ctx(pyfof_ctx* ctx, const double* xp, const double* vp, const double* mp,
                      const int32_t* np_halo, int32_t nhalo, double L_Box, double linking_length,
                      int32_t np_halo_min, int32_t bound_check, int32_t verbose, double softening,
                      double a_val, double b_val, double s_15_val, int32_t n_threads,
                      int32_t total_np, int64_t ng_val, double ng_global_val, int32_t* in_halo,
                      int32_t* halo_id, double* ang_mom, int32_t* root_out, int32_t mem_kind) {
    if (!ctx) { set_error("null context"); return IERR_INVALID; }
    CallArgs A{xp, vp, mp, np_halo, nhalo, L_Box, linking_length, np_halo_min, bound_check,
               verbose, softening, a_val, b_val, s_15_val, ng_val, total_np, in_halo, halo_id,
               ang_mom, root_out, mem_kind};
    A.ng_global = ng_global_val;
    return fof_entry(ctx->c, A, n_threads);
}

int32_t pyfof_fof(const double* xp, const double* vp, const double* mp, const int32_t* np_halo,
                  int32_t nhalo, double L_Box, double linking_length, int32_t np_halo_min,
                  int32_t bound_check, int32_t verbose, double softening, double a_val,
                  double b_val, double s_15_val, int32_t n_threads, int32_t total_np,
                  int64_t ng_val, double ng_global_val, int32_t* in_halo, int32_t* halo_id,
                  double* ang_mom, int32_t mem_kind) {
    Ctx* C = nullptr;
    try {
        C = default_ctx();
    } catch (const std::exception& e) {
        set_error(e.what());
        return IERR_HIP;
    }
    CallArgs A{xp, vp, mp, np_halo, nhalo, L_Box, linking_length, np_halo_min, bound_check,
               verbose, softening, a_val, b_val, s_15_val, ng_val, total_np, in_halo, halo_id,
               ang_mom, nullptr, mem_kind};
    A.ng_global = ng_global_val;
    return fof_entry(C, A, n_threads);
}

void pyfof_release(void) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    for (Ctx*& c : g_default_ctx) {
        delete c;
        c = nullptr;
    }
}

int32_t pyfof_catalog_size(pyfof_ctx* ctx, int32_t* nhalo_found, int64_t* n_members) {
    if (!ctx) return IERR_INVALID;
    if (!ctx->c->cat_hoff && ctx->c->cat_nh > 0) {
        set_error("catalog not built: enable PYFOF_OPT_CATALOG before the call");
        return IERR_INVALID;
    }
    *nhalo_found = (int32_t)ctx->c->cat_nh;
    *n_members = (int64_t)ctx->c->cat_nmem;
    return IERR_OK;
}

int32_t pyfof_catalog_copy(pyfof_ctx* ctx, int64_t* offsets, int32_t* indices, int32_t mem_kind) {
    if (!ctx) return IERR_INVALID;
    Ctx& C = *ctx->c;
    std::lock_guard<std::mutex> lk(C.mu);
    try {
        PYFOF_HIP(hipSetDevice(C.device));
        const uint32_t nh = C.cat_nh;
        std::vector<uint32_t> ends(nh);
        if (nh) {
            PYFOF_HIP(hipMemcpyAsync(ends.data(), C.cat_hoff, 4 * (size_t)nh, hipMemcpyDeviceToHost, C.stream));
            PYFOF_HIP(hipStreamSynchronize(C.stream));
        }
        std::vector<int64_t> off(nh + 1);
        off[0] = 0;
        for (uint32_t h = 0; h < nh; ++h) off[h + 1] = ends[h];
        const hipMemcpyKind k_off = mem_kind == PYFOF_MEM_DEVICE ? hipMemcpyHostToDevice : hipMemcpyHostToHost;
        PYFOF_HIP(hipMemcpyAsync(offsets, off.data(), 8 * (size_t)(nh + 1), k_off, C.stream));
        if (C.cat_nmem) {
            const hipMemcpyKind k_idx = mem_kind == PYFOF_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
            PYFOF_HIP(hipMemcpyAsync(indices, C.cat_members, 4 * (size_t)C.cat_nmem, k_idx, C.stream));
        }
        PYFOF_HIP(hipStreamSynchronize(C.stream));
        return IERR_OK;
    } catch (const HipError& e) {
        set_error(e.what());
        return IERR_HIP;
    }
}

int32_t pyfof_ctx_set_option(pyfof_ctx* ctx, int32_t option, int64_t value) {
    if (!ctx) return IERR_INVALID;
    Ctx& C = *ctx->c;
    switch (option) {
        case PYFOF_OPT_CATALOG: C.want_catalog = value != 0; return IERR_OK;
        case PYFOF_OPT_DIRECT_MAX:
            if (value < 1000 || value > 65536) {
                set_error("direct-summation threshold must be in [1000, 65536]");
                return IERR_INVALID;
            }
            C.direct_max = (int32_t)value;
            return IERR_OK;
        case PYFOF_OPT_DIRECT_FP32: C.direct_f32 = value != 0; return IERR_OK;
        case PYFOF_OPT_BATCHED: C.batched = value != 0; return IERR_OK;
        case PYFOF_OPT_RECORD_PHI: C.record_phi = value != 0; return IERR_OK;
        case PYFOF_OPT_TIME_FFT: C.time_fft = value != 0; return IERR_OK;
        case PYFOF_OPT_FFT_LINES: {
            const int n = (int)(value / 1000), lines = (int)(value % 1000);
            const int i = Ctx::fft_size_index(n);
            if (value < 0 || i < 0 || (lines != 4 && lines != 8 && lines != 16)) {
                set_error("PYFOF_OPT_FFT_LINES takes 1000 * n + lines with n in {48, 64, 128, 216, 256} "
                          "and lines in {4, 8, 16}");
                return IERR_INVALID;
            }
            C.fft_lines[i] = lines;
            return IERR_OK;
        }
        case PYFOF_OPT_POISON_ALLOC:
            if (value < 0 || value > 256) {
                set_error("PYFOF_OPT_POISON_ALLOC takes 0 or 1 + a byte value (at most 256)");
                return IERR_INVALID;
            }
            g_poison_alloc.store((int)value);
            return IERR_OK;
        default: break;
    }
    set_error("unknown option");
    return IERR_INVALID;
}

int64_t pyfof_ctx_potential(pyfof_ctx* ctx, float* phi, int32_t* index, int64_t cap, int32_t mem_kind) {
    if (!ctx) return -1;
    Ctx& C = *ctx->c;
    const uint64_t n = C.rec_n;
    if (!phi && !index) return (int64_t)n;
    if ((uint64_t)cap < n) {
        set_error("capacity below the number of recorded members");
        return -1;
    }
    if (n == 0) return 0;
    try {
        PYFOF_HIP(hipSetDevice(C.device));
        const hipMemcpyKind k = mem_kind == PYFOF_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        if (phi) PYFOF_HIP(hipMemcpyAsync(phi, C.rec_phi.as<float>(n), 4 * n, k, C.stream));
        if (index) PYFOF_HIP(hipMemcpyAsync(index, C.rec_idx.as<uint32_t>(n), 4 * n, k, C.stream));
        PYFOF_HIP(hipStreamSynchronize(C.stream));
    } catch (const HipError& e) {
        set_error(e.what());
        return -1;
    }
    return (int64_t)n;
}

int32_t pyfof_ctx_stats(pyfof_ctx* ctx, int64_t* out, int32_t n) {
    if (!ctx) return 0;
    const Ctx& C = *ctx->c;
    const BoundStats& b = C.last_bound;
    const int64_t v[18] = {(int64_t)C.last_n, (int64_t)C.last_ncell, (int64_t)C.last_ndense,
                           (int64_t)C.cat_nh, (int64_t)C.cat_nmem, (int64_t)b.nvalid,
                           (int64_t)b.n_ok, (int64_t)b.n_too_few, (int64_t)b.n_max_iter,
                           (int64_t)b.n_direct, (int64_t)b.n_pm[0], (int64_t)b.n_pm[1],
                           (int64_t)b.n_pm[2], (int64_t)b.n_pm[3], (int64_t)b.passes,
                           (int64_t)b.n_direct_large, (int64_t)C.fft_slots,
                           (int64_t)C.fft_bytes};
    int32_t w = 0;
    for (; w < n && w < 18; ++w) out[w] = v[w];
    return w;
}

int32_t pyfof_ctx_timings(pyfof_ctx* ctx, double* ms, int32_t n) {
    if (!ctx) return 0;
    Ctx& C = *ctx->c;
    int32_t w = 0;
    for (int e = 1; e < kNumEv && w < n; ++e, ++w) {
        ms[w] = 0.0;
        if (!C.ev_recorded[e] || !C.ev_recorded[e - 1]) continue;
        float t = 0.f;
        if (hipEventElapsedTime(&t, C.ev[e - 1], C.ev[e]) == hipSuccess) ms[w] = t;
    }
    if (w < n) ms[w++] = C.fft_ms;
    return w;
}

}  // extern "C"
