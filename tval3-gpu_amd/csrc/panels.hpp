// panels.hpp -- LDS-staged panel products for gfx950.
//
// A^T r and A x over 128^3 ray matrices are bound by scattered gathers: every
// entry of a row reads x at an unrelated index, and the L1 can look up only
// about one cache line per clock.  The panel form (Panels) groups rows so
// that a workgroup first brings the few thousand x values a panel reads into
// LDS (one coalesced or lightly scattered pass), after which every entry
// reads x from LDS through a 16-bit index.
//
//   A^T r: rows = voxels of one or two 8x8x8 tiles, sources = the rays
//          crossing them.
//   A u:   pass 1 rows = ray segments (the part of one ray inside a box of
//          1, 2, 4 or 8 tiles), sources = the box's voxels (a contiguous
//          range, or for 4 and 8 tiles one in a tile order of its own);
//          pass 2 sums each ray's segment partials (k_seg_sum) and runs the
//          epilogue.
//
// A row block's rows are stored back to back from a 4-aligned start, each
// padded with zero entries to a multiple of the group size G (1, 2 or 4),
// the block's end to a multiple of 4.  Phase 1: each thread loads 4
// consecutive entries (16 bytes; 4 / G whole groups of one row each),
// multiplies with the staged x and stores the 4 / G group sums in LDS;
// phase 2 sums each row's group sums (row_sum, a power-of-two lane group per
// row) and hands the total to the epilogue.  G trades padding (entries read)
// against phase 2's LDS reads: short rows favour small G.
#pragma once

#include <type_traits>

#include "lanes.hpp"
#include "siddon.hpp"
#include "spmv_stream.hpp"

namespace tv3 {

// weight of a packed entry (bits 16..31)
template <typename V> __device__ __forceinline__ float pk_weight(uint32_t u);
template <> __device__ __forceinline__ float pk_weight<__half>(uint32_t u) {
    return __half2float(__ushort_as_half((unsigned short)(u >> 16)));
}
template <> __device__ __forceinline__ float pk_weight<__hip_bfloat16>(uint32_t u) {
    return __uint_as_float(u & 0xffff0000u);
}

// The 4 stored entries at (4-aligned) position q: packed 16-bit weights and
// indices, or 16-bit indices and FP32 weights.
template <typename V> struct Quad { uint4 u; };
template <> struct Quad<float> { uint2 u; float4 w; };

template <typename V>
__device__ __forceinline__ Quad<V> load_quad(const uint32_t *pk, const uint16_t *, const float *,
                                             int64_t q) {
    return Quad<V>{*reinterpret_cast<const uint4 *>(pk + q)};
}
template <>
__device__ __forceinline__ Quad<float> load_quad<float>(const uint32_t *, const uint16_t *ps,
                                                        const float *pv, int64_t q) {
    return Quad<float>{*reinterpret_cast<const uint2 *>(ps + q),
                       *reinterpret_cast<const float4 *>(pv + q)};
}

// their 4 products with the staged sources
template <typename V>
__device__ __forceinline__ float4 quad_prod(const Quad<V> &e, const float *xs) {
    const uint4 u = e.u;
    return make_float4(pk_weight<V>(u.x) * xs[u.x & 0xffffu], pk_weight<V>(u.y) * xs[u.y & 0xffffu],
                       pk_weight<V>(u.z) * xs[u.z & 0xffffu], pk_weight<V>(u.w) * xs[u.w & 0xffffu]);
}
template <>
__device__ __forceinline__ float4 quad_prod<float>(const Quad<float> &e, const float *xs) {
    return make_float4(e.w.x * xs[e.u.x & 0xffffu], e.w.y * xs[e.u.x >> 16],
                       e.w.z * xs[e.u.y & 0xffffu], e.w.w * xs[e.u.y >> 16]);
}

// Epilogue that stores y[row] = sum unless the predicate *active is 0.
struct EpiStoreIf {
    float *y;
    const int32_t *active;     // NULL: always
    struct Pre {};
    __device__ bool skip() const { return active && !*active; }
    __device__ __forceinline__ Pre load(int) const { return {}; }
    __device__ __forceinline__ void row(int r, float s, bool on, const Pre &) const {
        if (on) y[r] = s;
    }
    __device__ __forceinline__ void finish() const {}
};

// Extra workgroups an epilogue asks of a panel launch (Epi::kExtra, default
// 0): blocks 0 .. kExtra - 1 run epi.extra() instead of a panel.
template <typename E, typename = void> struct ExtraBlocks { static constexpr int value = 0; };
template <typename E>
struct ExtraBlocks<E, std::void_t<decltype(E::kExtra)>> {
    static constexpr int value = E::kExtra;
};

// One workgroup per work item: a chunk of a panel's row blocks, whose
// workgroup stages the panel's sources (largest items first, so the launch
// ends with small ones).  Block descriptors: p0 / len = start / length of the block's
// stored entries (both multiples of 4); row r's entries start at pptr[r].
// A panel's blocks are stored back to back, so block b + 1 starts where
// block b ends: its entries are requested before block b's products are
// formed (the first block's before the sources are staged).  Such a
// request covers NB entries from the block's start whatever its length,
// clamped to the panel's last stored quad (so no entries past the panel are
// fetched from memory; qlast, the last stored quad overall, bounds a panel
// without entries): entries past the block belong to the panel's later
// blocks or repeat its last quad, and phase 2 never sums them.
template <int NB, int G, typename V, bool CONTIG, typename Epi>
__global__ __launch_bounds__(256) void k_panel(const int32_t *__restrict__ order,
                                               const int32_t *__restrict__ src_off,
                                               const int32_t *__restrict__ chunk,
                                               const int32_t *__restrict__ src,
                                               const BlkDesc *__restrict__ desc,
                                               const int32_t *__restrict__ pptr,
                                               const uint32_t *__restrict__ pk,
                                               const uint16_t *__restrict__ ps,
                                               const float *__restrict__ pv, int64_t qlast,
                                               const float *__restrict__ x, Epi epi) {
    static_assert(G == 1 || G == 2 || G == 4, "group of 1, 2 or 4 entries");
    constexpr int NQ = NB / 4;                   // 16-byte entry quads per block
    extern __shared__ float xs[];                // the panel's source values
    __shared__ __align__(16) float prod[NB / G]; // the block's group sums
    constexpr int NX = ExtraBlocks<Epi>::value;
    if constexpr (NX > 0) {
        if ((int)blockIdx.x < NX) {
            epi.extra();
            return;
        }
    }
    if (epi.skip()) return;
    const int t = threadIdx.x;
    const int item = order[blockIdx.x - NX];
    const int pan = chunk[3 * item], bb = chunk[3 * item + 1], b1 = chunk[3 * item + 2];
    const int s0 = src_off[pan], ns = src_off[pan + 1] - s0;
    BlkDesc d = desc[bb];
    if constexpr (HasSkipRows<Epi>::value) {
        // a panel's rows lie in one row range of the epilogue (tiles never
        // straddle one): its first row decides, uniformly
        if (bb < b1 && epi.skip_rows(d.r0)) return;
    }
    constexpr int J = NQ / 256;
    const int64_t qhi = min(qlast, max((int64_t)desc[b1].p0 - 4, (int64_t)d.p0));
    Quad<V> cur[J];
#pragma unroll
    for (int j = 0; j < J; ++j)
        cur[j] = load_quad<V>(pk, ps, pv, min((int64_t)d.p0 + 4 * (j * 256 + t), qhi));
    // stage the sources: batches of 8 per thread, all loads in flight
    for (int k0 = 0; k0 < ns; k0 += 8 * 256) {
        float xv[8];
        if (CONTIG && src == nullptr) {
#pragma unroll
            for (int i = 0; i < 8; ++i) xv[i] = x[s0 + min(k0 + i * 256 + t, ns - 1)];
        } else if (CONTIG) {
            // index range in a tile order of its own: src maps its 512-row
            // tiles to the tiles of x
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = s0 + min(k0 + i * 256 + t, ns - 1);
                xv[i] = x[((int64_t)src[k >> 9] << 9) | (k & 511)];
            }
        } else {
            int rv[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) rv[i] = src[s0 + min(k0 + i * 256 + t, ns - 1)];
#pragma unroll
            for (int i = 0; i < 8; ++i) xv[i] = x[rv[i]];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (k0 + i * 256 + t < ns) xs[k0 + i * 256 + t] = xv[i];
    }
    __syncthreads();
    for (int b = bb; b < b1; ++b) {
        const BlkDesc dn = desc[b + 1];      // in flight while this block runs
        Quad<V> nxt[J];
        if (b + 1 < b1) {
            const int64_t pn = (int64_t)d.p0 + d.len;
#pragma unroll
            for (int j = 0; j < J; ++j)
                nxt[j] = load_quad<V>(pk, ps, pv, min(pn + 4 * (j * 256 + t), qhi));
        }
        if (d.len <= NB) {
            const int row = block_row(d);
            const int a = (pptr[row] - d.p0) / G, e = (pptr[row + 1] - d.p0) / G;
            const auto pre = epi.load(row);
            if (d.len > 0) {
                float4 v[J];
#pragma unroll
                for (int j = 0; j < J; ++j) v[j] = quad_prod<V>(cur[j], xs);
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    const int q = j * 256 + t;
                    if constexpr (G == 4) {
                        prod[q] = (v[j].x + v[j].y) + (v[j].z + v[j].w);
                    } else if constexpr (G == 2) {
                        *reinterpret_cast<float2 *>(prod + 2 * q) =
                            make_float2(v[j].x + v[j].y, v[j].z + v[j].w);
                    } else {
                        *reinterpret_cast<float4 *>(prod + 4 * q) = v[j];
                    }
                }
            }
            __syncthreads();
            block_rows<NB / G>(prod, d, a, e, epi, row, pre);
        } else {
            // one row longer than NB
            const auto pre = epi.load(d.r0);
            float s = 0;
            for (int g = t; g < (d.len >> 2); g += 256) {
                const float4 v = quad_prod<V>(load_quad<V>(pk, ps, pv, (int64_t)d.p0 + 4 * g), xs);
                s += (v.x + v.y) + (v.z + v.w);
            }
            s = lane_sum<64>(s);
            if ((t & 63) == 0) prod[t >> 6] = s;
            __syncthreads();
            if (t == 0) s = (prod[0] + prod[1]) + (prod[2] + prod[3]);
            epi.row(d.r0, s, t == 0, pre);
        }
        __syncthreads();             // prod is rewritten by the next row block
        d = dn;
        if (b + 1 < b1) {
#pragma unroll
            for (int j = 0; j < J; ++j) cur[j] = nxt[j];
        }
    }
    epi.finish();
}

// The panel form with packed entries (Panels::fmt 12, FP16 or BF16 weights,
// W): the stored entries in octets of 8, their 12-bit local indices in 3
// words (entry e at bits 12 e .. 12 e + 11 of the 96) and their 8 weights in
// 16 bytes -- 28 bytes for 8 entries instead of 32, for panels of at most
// 4,096 sources (A's boxes of 16^3 voxels, A^T panels of at most 4,096
// rays).  Row blocks start and end 8-aligned; rows unpadded (G = 1) or
// padded to groups of 2 or 4 (an octet holds four or two groups, whose sums
// are formed as k_panel forms them, so both kernels give the same bits for
// the same rows and blocks).  Sources are contiguous (src null), staged
// through a tile map (src, gather = 0) or gathered from a ray list (src,
// gather = 1).
struct Oct {
    uint3 i;
    uint4 w;
};

__device__ __forceinline__ Oct load_oct(const uint32_t *pi, const uint4 *pw, int64_t o) {
    const uint32_t *q = pi + 3 * o;
    return Oct{make_uint3(q[0], q[1], q[2]), pw[o]};
}

// weight e of a word (low or high half): FP16, or BF16 (the top half of a float)
template <typename W> __device__ __forceinline__ float oct_w(uint32_t u, int hi);
template <> __device__ __forceinline__ float oct_w<__half>(uint32_t u, int hi) {
    return __half2float(__ushort_as_half((unsigned short)(hi ? u >> 16 : u & 0xffffu)));
}
template <> __device__ __forceinline__ float oct_w<__hip_bfloat16>(uint32_t u, int hi) {
    return __uint_as_float(hi ? (u & 0xffff0000u) : (u << 16));
}

// the 8 products of an octet with the staged sources
template <typename W>
__device__ __forceinline__ void oct_prod(const Oct &c, const float *xs, float4 &p0, float4 &p1) {
    const uint32_t a = c.i.x, b = c.i.y, d = c.i.z;
    const int i0 = a & 0xfff, i1 = (a >> 12) & 0xfff, i2 = (a >> 24) | ((b & 0xf) << 8),
              i3 = (b >> 4) & 0xfff, i4 = (b >> 16) & 0xfff, i5 = (b >> 28) | ((d & 0xff) << 4),
              i6 = (d >> 8) & 0xfff, i7 = d >> 20;
    p0 = make_float4(oct_w<W>(c.w.x, 0) * xs[i0], oct_w<W>(c.w.x, 1) * xs[i1],
                     oct_w<W>(c.w.y, 0) * xs[i2], oct_w<W>(c.w.y, 1) * xs[i3]);
    p1 = make_float4(oct_w<W>(c.w.z, 0) * xs[i4], oct_w<W>(c.w.z, 1) * xs[i5],
                     oct_w<W>(c.w.w, 0) * xs[i6], oct_w<W>(c.w.w, 1) * xs[i7]);
}

template <int NB, int G, typename W, typename Epi>
__global__ __launch_bounds__(256) void k_panel12(const int32_t *__restrict__ order,
                                                 const int32_t *__restrict__ src_off,
                                                 const int32_t *__restrict__ chunk,
                                                 const int32_t *__restrict__ src, int gather,
                                                 const BlkDesc *__restrict__ desc,
                                                 const int32_t *__restrict__ pptr,
                                                 const uint32_t *__restrict__ pi,
                                                 const uint4 *__restrict__ pw, int64_t olast,
                                                 const float *__restrict__ x, Epi epi) {
    constexpr int NO = NB / 8;                   // octets per block
    constexpr int J = NO / 256;
    static_assert(J >= 1, "blocks of at least 2048 entries");
    static_assert(G == 1 || G == 2 || G == 4, "rows in groups of 1, 2 or 4");
    extern __shared__ float xs[];                // the panel's source values
    __shared__ __align__(16) float prod[NB / G]; // the block's products (group sums)
    constexpr int NX = ExtraBlocks<Epi>::value;
    if constexpr (NX > 0) {
        if ((int)blockIdx.x < NX) {
            epi.extra();
            return;
        }
    }
    if (epi.skip()) return;
    const int t = threadIdx.x;
    const int item = order[blockIdx.x - NX];
    const int pan = chunk[3 * item], bb = chunk[3 * item + 1], b1 = chunk[3 * item + 2];
    const int s0 = src_off[pan], ns = src_off[pan + 1] - s0;
    BlkDesc d = desc[bb];
    if constexpr (HasSkipRows<Epi>::value) {
        if (bb < b1 && epi.skip_rows(d.r0)) return;
    }
    const int64_t ohi = min(olast, max((int64_t)desc[b1].p0 / 8 - 1, (int64_t)d.p0 / 8));
    Oct cur[J];
#pragma unroll
    for (int j = 0; j < J; ++j) cur[j] = load_oct(pi, pw, min((int64_t)d.p0 / 8 + j * 256 + t, ohi));
    for (int k0 = 0; k0 < ns; k0 += 8 * 256) {
        float xv[8];
        if (src == nullptr) {
#pragma unroll
            for (int i = 0; i < 8; ++i) xv[i] = x[s0 + min(k0 + i * 256 + t, ns - 1)];
        } else if (!gather) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = s0 + min(k0 + i * 256 + t, ns - 1);
                xv[i] = x[((int64_t)src[k >> 9] << 9) | (k & 511)];
            }
        } else {
            int rv[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) rv[i] = src[s0 + min(k0 + i * 256 + t, ns - 1)];
#pragma unroll
            for (int i = 0; i < 8; ++i) xv[i] = x[rv[i]];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (k0 + i * 256 + t < ns) xs[k0 + i * 256 + t] = xv[i];
    }
    __syncthreads();
    for (int b = bb; b < b1; ++b) {
        const BlkDesc dn = desc[b + 1];      // in flight while this block runs
        Oct nxt[J];
        if (b + 1 < b1) {
            const int64_t on = ((int64_t)d.p0 + d.len) / 8;
#pragma unroll
            for (int j = 0; j < J; ++j) nxt[j] = load_oct(pi, pw, min(on + j * 256 + t, ohi));
        }
        if (d.len <= NB) {
            const int row = block_row(d);
            const int a = (pptr[row] - d.p0) / G, e = (pptr[row + 1] - d.p0) / G;
            const auto pre = epi.load(row);
            if (d.len > 0) {
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    float4 v0, v1;
                    oct_prod<W>(cur[j], xs, v0, v1);
                    const int o = j * 256 + t;
                    if constexpr (G == 4) {
                        *reinterpret_cast<float2 *>(prod + 2 * o) =
                            make_float2((v0.x + v0.y) + (v0.z + v0.w),
                                        (v1.x + v1.y) + (v1.z + v1.w));
                    } else if constexpr (G == 2) {
                        *reinterpret_cast<float4 *>(prod + 4 * o) =
                            make_float4(v0.x + v0.y, v0.z + v0.w, v1.x + v1.y, v1.z + v1.w);
                    } else {
                        *reinterpret_cast<float4 *>(prod + 8 * o) = v0;
                        *reinterpret_cast<float4 *>(prod + 8 * o + 4) = v1;
                    }
                }
            }
            __syncthreads();
            block_rows<NB / G>(prod, d, a, e, epi, row, pre);
        } else {
            // one row longer than NB
            const auto pre = epi.load(d.r0);
            float s = 0;
            for (int g = t; g < (d.len >> 3); g += 256) {
                float4 v0, v1;
                oct_prod<W>(load_oct(pi, pw, (int64_t)d.p0 / 8 + g), xs, v0, v1);
                s += ((v0.x + v0.y) + (v0.z + v0.w)) + ((v1.x + v1.y) + (v1.z + v1.w));
            }
            s = lane_sum<64>(s);
            if ((t & 63) == 0) prod[t >> 6] = s;
            __syncthreads();
            if (t == 0) s = (prod[0] + prod[1]) + (prod[2] + prod[3]);
            epi.row(d.r0, s, t == 0, pre);
        }
        __syncthreads();             // prod is rewritten by the next row block
        d = dn;
        if (b + 1 < b1) {
#pragma unroll
            for (int j = 0; j < J; ++j) cur[j] = nxt[j];
        }
    }
    epi.finish();
}

// Pass 2 of A x: L lanes per ray sum its segment partials (batches of 4 per
// lane, loads clamped to the ray's last segment), then the epilogue.
template <int L, int NLB, typename Epi>
__global__ __launch_bounds__(256) void k_seg_sum(int64_t rows, const int32_t *__restrict__ ptr,
                                                 const int32_t *__restrict__ seg,
                                                 const float *__restrict__ part, Epi epi,
                                                 int64_t nlogical) {
    // A workgroup owns NLB consecutive logical blocks (256 / L rays each).
    // The three dependent loads of a ray -- its segment range, the segment
    // indices, the partial sums -- are issued for all NLB blocks together, so
    // the workgroup waits for three memory round trips instead of 3 NLB.
    if (epi.skip()) return;
    const int64_t w = xcd_remap(blockIdx.x, (int)gridDim.x);
    const int lane = threadIdx.x & (L - 1);
    int64_t row[NLB];
    bool has[NLB];
    int beg[NLB], end[NLB];
    typename Epi::Pre pre[NLB];
#pragma unroll
    for (int j = 0; j < NLB; ++j) {
        const int64_t lb = w * NLB + j;
        const int64_t row0 = (lb * 256 + threadIdx.x) / L;
        has[j] = lb < nlogical && row0 < rows;
        if constexpr (HasSkipRows<Epi>::value) {
            // a logical block's rays lie in one row range of the epilogue
            if (has[j] && epi.skip_rows((int)(lb * (256 / L)))) has[j] = false;
        }
        row[j] = has[j] ? row0 : rows - 1;
        beg[j] = ptr[row[j]];
        end[j] = has[j] ? ptr[row[j] + 1] : beg[j];
        pre[j] = epi.load((int)row[j]);
    }
    // first batch of 4 segments per lane of every block (loads clamped into
    // the ray's range, masked after the read)
    int c[NLB][4];
#pragma unroll
    for (int j = 0; j < NLB; ++j) {
        const int last = end[j] > 0 ? end[j] - 1 : 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) c[j][i] = seg[min(beg[j] + lane + i * L, last)];
    }
    float v[NLB][4];
#pragma unroll
    for (int j = 0; j < NLB; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) v[j][i] = part[c[j][i]];
#pragma unroll
    for (int j = 0; j < NLB; ++j) {
        const int k0 = beg[j] + lane;
        float s0 = (k0 < end[j] ? v[j][0] : 0.f) + (k0 + 2 * L < end[j] ? v[j][2] : 0.f);
        float s1 = (k0 + L < end[j] ? v[j][1] : 0.f) + (k0 + 3 * L < end[j] ? v[j][3] : 0.f);
        // rays with more than 4 L segments
        for (int k = k0 + 4 * L; k < end[j]; k += 4 * L) {
            int cc[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) cc[i] = seg[min(k + i * L, end[j] - 1)];
            float vv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) vv[i] = part[cc[i]];
            s0 += vv[0] + (k + 2 * L < end[j] ? vv[2] : 0.f);
            s1 += (k + L < end[j] ? vv[1] : 0.f) + (k + 3 * L < end[j] ? vv[3] : 0.f);
        }
        const float sum = lane_sum<L>(s0 + s1);
        epi.row((int)row[j], sum, has[j] && lane == 0, pre[j]);
    }
    epi.finish();
}

template <typename V, bool CONTIG, typename Epi>
void launch_panel_v(const Panels &P, const float *x, const Epi &epi, hipStream_t s) {
    const unsigned g = (unsigned)P.nitem + ExtraBlocks<Epi>::value;
    const size_t lds = sizeof(float) * (size_t)(P.max_src > 0 ? P.max_src : 1);
    if (g == 0) return;
#define TV3_PCASE(NBV, GV)                                                                    \
    case NBV * 8 + GV:                                                                        \
        hipLaunchKernelGGL((k_panel<NBV, GV, V, CONTIG, Epi>), dim3(g), dim3(256), lds, s,      \
                           P.order.p, P.src_off.p, P.chunk.p, P.src.p, P.desc.p, P.pptr.p,    \
                           P.pk.p, P.ps.p, P.pv.p, P.stored - 4, x, epi);                     \
        break;
    switch (P.blk_nnz * 8 + P.group) {
        TV3_PCASE(1024, 1) TV3_PCASE(2048, 1) TV3_PCASE(4096, 1)
        TV3_PCASE(1024, 2) TV3_PCASE(2048, 2) TV3_PCASE(4096, 2)
        TV3_PCASE(1024, 4) TV3_PCASE(2048, 4) TV3_PCASE(4096, 4)
    default:
        throw tv3_error(TV3_E_ARG, "panel block size must be 1024, 2048 or 4096 entries and "
                                   "the group 1, 2 or 4");
    }
#undef TV3_PCASE
}

// Panel product y = M x with epilogue epi (FP32 accumulation).
template <typename Epi>
void launch_panel(const Panels &P, const float *x, const Epi &epi, hipStream_t s) {
    if (P.fmt == 12) {
        const unsigned g = (unsigned)P.nitem + ExtraBlocks<Epi>::value;
        const size_t lds = sizeof(float) * (size_t)(P.max_src > 0 ? P.max_src : 1);
        if (g == 0) return;
        // sources: A's contiguous range or tile map, A^T's ray list
        const int gather = P.contiguous ? 0 : 1;
        const bool bf = P.vtype == TV3_BF16;
        TV3_REQUIRE(bf || P.vtype == TV3_F16, TV3_E_FORMAT, "packed entries hold FP16 or BF16");
#define TV3_P12CASE(NBV, GV)                                                                  \
    case NBV * 8 + GV:                                                                        \
        if (bf)                                                                               \
            hipLaunchKernelGGL((k_panel12<NBV, GV, __hip_bfloat16, Epi>), dim3(g), dim3(256),   \
                               lds, s, P.order.p, P.src_off.p, P.chunk.p, P.src.p, gather,    \
                               P.desc.p, P.pptr.p, P.pk.p,                                    \
                               reinterpret_cast<const uint4 *>(P.pw.p), P.stored / 8 - 1, x,  \
                               epi);                                                          \
        else                                                                                  \
            hipLaunchKernelGGL((k_panel12<NBV, GV, __half, Epi>), dim3(g), dim3(256), lds, s,   \
                               P.order.p, P.src_off.p, P.chunk.p, P.src.p, gather, P.desc.p,  \
                               P.pptr.p, P.pk.p, reinterpret_cast<const uint4 *>(P.pw.p),     \
                               P.stored / 8 - 1, x, epi);                                     \
        break;
        switch (P.blk_nnz * 8 + P.group) {
            TV3_P12CASE(2048, 1) TV3_P12CASE(4096, 1) TV3_P12CASE(2048, 2) TV3_P12CASE(4096, 2)
            TV3_P12CASE(2048, 4) TV3_P12CASE(4096, 4)
        default:
            throw tv3_error(TV3_E_ARG, "packed panel entries take blocks of 2048 or 4096 and "
                                       "groups of 1, 2 or 4");
        }
#undef TV3_P12CASE
        return;
    }
    switch (P.vtype) {
    case TV3_F32:
        P.contiguous ? launch_panel_v<float, true, Epi>(P, x, epi, s)
                     : launch_panel_v<float, false, Epi>(P, x, epi, s);
        break;
    case TV3_F16:
        P.contiguous ? launch_panel_v<__half, true, Epi>(P, x, epi, s)
                     : launch_panel_v<__half, false, Epi>(P, x, epi, s);
        break;
    case TV3_BF16:
        P.contiguous ? launch_panel_v<__hip_bfloat16, true, Epi>(P, x, epi, s)
                     : launch_panel_v<__hip_bfloat16, false, Epi>(P, x, epi, s);
        break;
    default: throw tv3_error(TV3_E_FORMAT, "panels need F32, F16 or BF16 weights");
    }
}

// ------------------------------------------------------------------------
// Matrix-free Siddon products in boxes (DynTiles, built by dyntile.hip).
struct DynView {
    const DynPan *pan;
    const int32_t *order, *seg_ray, *seg_eoff;
    const uint16_t *seg_c0;
    const float *ray;
    const uint32_t *pptr;
    const uint16_t *pos;
    int emax_cap;
    int tile_dx, tile_dz;   // internal offset of the next tile along x, z
};

inline DynView dyn_view(const DynTiles &D, const DynBoxes &X) {
    DynView v;
    v.pan = X.pan.p; v.order = X.order.p; v.seg_ray = X.seg_ray.p; v.seg_eoff = X.seg_eoff.p;
    v.seg_c0 = X.seg_c0.p;
    v.ray = D.ray.p; v.pptr = D.pptr.p; v.pos = D.pos.p;
    v.emax_cap = D.emax_cap;
    v.tile_dx = D.tile_dx;
    v.tile_dz = D.tile_dz;
    return v;
}

// Box-local coordinates (x, y, z) of the voxel at offset i of a box in the
// internal order of brick_index: bricks of 4^3 in (y, x, z) bit order, y
// fastest in a brick (a 4^3 box holds one brick); in a box of two tiles
// consecutive in y the second tile's offsets start at 512, y + 8.
__host__ __device__ __forceinline__ void dyn_local_xyz(int i, int &x, int &y, int &z) {
    y = (((i >> 9) & 1) << 3) | (((i >> 6) & 1) << 2) | (i & 3);
    x = (((i >> 7) & 1) << 2) | ((i >> 2) & 3);
    z = (((i >> 8) & 1) << 2) | ((i >> 4) & 3);
}

// The traversal of one segment (the part of an internal ray inside box P,
// starting in the voxel of box-local coordinates c0 = x | y << 4 | z << 8):
// emit(offset of the voxel in the box, length) in traversal order, the
// offset in linear box order y + ey x + ey bs z (edges bs, ey, bs along x,
// y, z).  These are the entries the
// walk of the whole ray (sid_walk, siddon.hpp) emits inside the box:
//   * the ray record holds sid_setup's A, B, L, amin, amax (12 words:
//     A[3], L, B[3], amin, amax, 3 unused);
//   * at its first voxel c the walk's state is the next plane c + 1 (c when
//     travelling down) per axis, and its parameter the voxel's entry event
//     max(amin, t of the entry plane per axis) -- the largest plane
//     parameter not above it, as in sid_box;
//   * it ends at min(amax, t of the box's exit plane per axis);
//   * each step is sid_walk's: the nearest next plane, the length
//     (t_next - t_cur) L emitted when positive, every axis whose next plane
//     is crossed at t_next advanced; the plane parameters are the same fused
//     operation on the same values.  A parallel axis has A = 0, B = +inf, so
//     its next plane never comes; no plane beyond the box's exit is ever
//     evaluated (the walk stops at its parameter first).
struct DynSeg {
    float4 ra, rb;   // A[3], L; B[3], amin
    float amax;
    int c0;          // box-local first voxel x | y << 4 | z << 8
};

// a segment's traversal constants: its ray's record and its first voxel
__host__ __device__ __forceinline__ DynSeg dyn_seg(const DynView &g, int ray, int c0) {
    const float *rec = g.ray + 12 * (int64_t)ray;
    return DynSeg{reinterpret_cast<const float4 *>(rec)[0],
                  reinterpret_cast<const float4 *>(rec)[1], rec[8], c0};
}

template <typename F>
__host__ __device__ __forceinline__ void dyn_seg_walk(const DynSeg &sr, const DynPan &P, F emit) {
#pragma clang fp contract(off)
    const float4 ra = sr.ra, rb = sr.rb;
    const float amax = sr.amax;
    const int c0 = sr.c0;
    const float A[3] = {ra.x, ra.y, ra.z}, B[3] = {rb.x, rb.y, rb.z};
    const int lc[3] = {c0 & 15, (c0 >> 4) & 15, (c0 >> 8) & 15};
    const int b0[3] = {P.bx, P.by, P.bz};
    const int bs = P.bs, ey = P.ey;
    const int ext[3] = {bs, ey, bs};
    const int stride[3] = {ey, 1, ey * bs};
    float kf[3], tk[3], st[3];
    int dstep[3], off = 0;
    float ts = rb.w, te = amax;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool up = A[a] > 0.f, par = A[a] == 0.f;
        const int c = b0[a] + lc[a];
        st[a] = up ? 1.f : (par ? 0.f : -1.f);
        kf[a] = (float)(up ? c + 1 : c);
        tk[a] = sid_fma(kf[a], A[a], B[a]);
        if (!par) {
            const float tin = sid_fma((float)(up ? c : c + 1), A[a], B[a]);
            const float tout = sid_fma((float)(up ? b0[a] + ext[a] : b0[a]), A[a], B[a]);
            ts = tin > ts ? tin : ts;
            te = tout < te ? tout : te;
        }
        off += lc[a] * stride[a];
        dstep[a] = up ? stride[a] : (par ? 0 : -stride[a]);
    }
    // the offset is masked all the same, so that no index leaves the box
    const int mask = ey * bs * bs - 1;
    float tcur = ts;
    for (int step = 0; step < 2 * bs + ey + 4; ++step) {
        float tnext = te;
#pragma unroll
        for (int a = 0; a < 3; ++a) tnext = tk[a] < tnext ? tk[a] : tnext;
        const float seg = (tnext - tcur) * ra.w;
        if (seg > 0) emit(off & mask, seg);
        if (!(tnext < te)) break;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const bool m = tk[a] == tnext;
            kf[a] += m ? st[a] : 0.f;
            off += m ? dstep[a] : 0;
            tk[a] = sid_fma(kf[a], A[a], B[a]);
        }
        tcur = tnext;
    }
}

template <typename F>
__host__ __device__ __forceinline__ void dyn_seg_walk(const DynView &g, const DynPan &P, int ray,
                                                      int c0, F emit) {
    dyn_seg_walk(dyn_seg(g, ray, c0), P, emit);
}

// A x, first pass: per box its voxels staged in LDS, each thread walks
// segments and stores their partial sums (walk order) in part[segment].
// `active` (may be NULL) predicates the launch like an epilogue's skip().
template <typename T>
__global__ __launch_bounds__(256) void k_dyn_tile_N(DynView g, const T *__restrict__ x,
                                                    T *__restrict__ part,
                                                    const int32_t *active) {
    if (active && !*active) return;
    __shared__ T xs[4096];
    const DynPan P = g.pan[g.order[blockIdx.x]];
    const int t = threadIdx.x, bs = P.bs, ey = P.ey, nv = ey * bs * bs;
    for (int i = t; i < nv; i += 256) {     // linear box order y + ey x + ey bs z
        // tile sb of the box (bits: y, x, z; boxes of one tile or brick: 0)
        const int j = i & 511, sb = i >> 9;
        int lx, ly, lz;
        dyn_local_xyz(j, lx, ly, lz);
        lx += (sb & 2) << 2;
        ly += (sb & 1) << 3;
        lz += (sb & 4) << 1;
        xs[ly + ey * (lx + bs * lz)] =
            x[P.v0 + (sb & 1) * 512 + ((sb >> 1) & 1) * g.tile_dx + (sb >> 2) * g.tile_dz + j];
    }
    __syncthreads();
    for (int sg = P.s0 + t; sg < P.s1; sg += 256) {
        T acc = 0;
        dyn_seg_walk(g, P, g.seg_ray[sg], g.seg_c0[sg],
                     [&](int li, float w) { acc += (T)w * xs[li]; });
        part[sg] = acc;
    }
}

// A^T r: per box each thread walks segments and writes w r_ray at the
// entry's position in the box's entry array in LDS (ray-major, emax_cap + 1
// slots); then each voxel row sums its entries in ray order through the
// stored positions (padding points at the zero slot ne) and the epilogue
// takes the sum.
template <typename T, int NT, typename Epi>
__global__ __launch_bounds__(NT) void k_dyn_tile_T(DynView g, const T *__restrict__ x, Epi epi) {
    static_assert(ExtraBlocks<Epi>::value == 0, "no extra workgroups in the tile product");
    extern __shared__ __align__(16) unsigned char dyn_smem[];
    if (epi.skip()) return;
    static_assert(NT >= 512, "one voxel row per thread (boxes of at most 8^3)");
    const DynPan P = g.pan[g.order[blockIdx.x]];
    T *pe = reinterpret_cast<T *>(dyn_smem);
    const int t = threadIdx.x, ns = P.s1 - P.s0, nv = P.ey * P.bs * P.bs;
    // phase 2's operands do not depend on phase 1: its row starts, first
    // positions and epilogue operands are requested before the walks
    const bool has_row = t < nv;
    const int row = P.v0 + (has_row ? t : 0);
    uint32_t q0 = 0, q1 = 0;
    if (has_row) {
        q0 = g.pptr[row];
        q1 = g.pptr[row + 1];
    }
    uint2 u0 = make_uint2(0, 0);
    if (q0 < q1) u0 = *reinterpret_cast<const uint2 *>(g.pos + q0);
    const auto pre = epi.load(row);
    // phase 1: the next segment's record, slots and residual are requested
    // while the current one is walked
    struct Cur {
        DynSeg sr;
        int e0, e1;
        T xv;
    };
    auto fetch = [&](int j) {
        const int sg = P.s0 + j, ray = g.seg_ray[sg];
        return Cur{dyn_seg(g, ray, g.seg_c0[sg]), g.seg_eoff[sg],
                   j + 1 < ns ? g.seg_eoff[sg + 1] : P.ne, x[ray]};
    };
    if (t < ns) {
        Cur cur = fetch(t);
        for (int j = t; j < ns; j += NT) {
            Cur nxt = cur;
            if (j + NT < ns) nxt = fetch(j + NT);
            int e = cur.e0;
            const int e1 = cur.e1;
            const T xv = cur.xv;
            dyn_seg_walk(cur.sr, P, [&](int, float w) {
                if (e < e1) pe[e] = (T)w * xv;
                ++e;
            });
            cur = nxt;
        }
    }
    if (t == 0) pe[P.ne] = 0;
    __syncthreads();
    if (has_row) {
        T acc = 0;
        if (q0 < q1)
            acc += (pe[u0.x & 0xffffu] + pe[u0.x >> 16]) + (pe[u0.y & 0xffffu] + pe[u0.y >> 16]);
        for (uint32_t q = q0 + 4; q < q1; q += 4) {
            const uint2 u = *reinterpret_cast<const uint2 *>(g.pos + q);
            acc += (pe[u.x & 0xffffu] + pe[u.x >> 16]) + (pe[u.y & 0xffffu] + pe[u.y >> 16]);
        }
        epi.row(row, acc, true, pre);
    }
    epi.finish();
}

template <typename T>
void launch_dyn_N(const DynTiles &D, const T *x, T *part, const int32_t *active, hipStream_t s) {
    if (D.fwd.npan == 0) return;
    hipLaunchKernelGGL((k_dyn_tile_N<T>), dim3((unsigned)D.fwd.npan), dim3(256), 0, s,
                       dyn_view(D, D.fwd), x, part, active);
}

template <typename T, typename Epi>
void launch_dyn_T(const DynTiles &D, const T *x, const Epi &epi, hipStream_t s) {
    if (D.bwd.npan == 0) return;
    const size_t lds = sizeof(T) * (size_t)(D.emax_cap + 1);
    switch (D.threads_T) {
    case 512:
        hipLaunchKernelGGL((k_dyn_tile_T<T, 512, Epi>), dim3((unsigned)D.bwd.npan), dim3(512), lds,
                           s, dyn_view(D, D.bwd), x, epi);
        break;
    case 1024:
        hipLaunchKernelGGL((k_dyn_tile_T<T, 1024, Epi>), dim3((unsigned)D.bwd.npan), dim3(1024),
                           lds, s, dyn_view(D, D.bwd), x, epi);
        break;
    default: throw tv3_error(TV3_E_ARG, "the A^T tile product runs 512 or 1024 threads");
    }
}

#ifndef TV3_SEGSUM_NLB
#define TV3_SEGSUM_NLB 2
#endif
constexpr int kSegSumNLB = TV3_SEGSUM_NLB;   // logical blocks (256 / L rays each) per pass-2 workgroup

// Logical blocks and workgroups of pass 2 of A x.  The workgroup count is
// also the number of block partials its epilogue stores: every reader of
// those partials takes it from here.
inline int64_t seg_sum_logical(const SegSum &R) { return (R.rows * R.lanes + 255) / 256; }
inline int seg_sum_blocks(const SegSum &R) {
    const int64_t ng = (seg_sum_logical(R) + kSegSumNLB - 1) / kSegSumNLB;
    TV3_REQUIRE(ng < ((int64_t)1 << 31), TV3_E_SHAPE, "too many rays for one launch");
    return (int)ng;
}

template <typename Epi>
void launch_seg_sum(const SegSum &R, const Epi &epi, hipStream_t s) {
    constexpr int NLB = kSegSumNLB;
    const int L = R.lanes;
    const int64_t nl = seg_sum_logical(R);
    const unsigned g = (unsigned)seg_sum_blocks(R);
    if (g == 0) return;
#define TV3_SCASE(LL)                                                                     \
    case LL:                                                                              \
        hipLaunchKernelGGL((k_seg_sum<LL, NLB, Epi>), dim3(g), dim3(256), 0, s, R.rows,     \
                           R.ptr.p, R.seg.p, R.part.p, epi, nl);                          \
        break;
    switch (L) {
        TV3_SCASE(2) TV3_SCASE(4) TV3_SCASE(8) TV3_SCASE(16) TV3_SCASE(32)
    default: throw tv3_error(TV3_E_ARG, "segment-sum lanes must be a power of two in [2, 32]");
    }
#undef TV3_SCASE
}

// A x of a fused layout (rows = rays) with epilogue epi; `active` (may be
// NULL) predicates the first pass of the panel form like epi.skip().
template <typename Acc, typename Epi>
void launch_forward(const FusedLayout &F, const float *x, const Epi &epi, hipStream_t s,
                    const int32_t *active = nullptr) {
    if (F.dyn) {
        launch_dyn_N<float>(*F.dyn, x, F.RS.part.p, active, s);
        launch_seg_sum(F.RS, epi, s);
    } else if (F.panels_A) {
        launch_panel(F.PA, x, EpiStoreIf{F.RS.part.p, active}, s);
        launch_seg_sum(F.RS, epi, s);
    } else {
        launch_product<Acc, Epi>(F.A, x, epi, s);
    }
}

// A^T r of a fused layout (rows = voxels) with epilogue epi.
template <typename Acc, typename Epi>
void launch_transposed(const FusedLayout &F, const float *x, const Epi &epi, hipStream_t s) {
    if (F.dyn) launch_dyn_T<float>(*F.dyn, x, epi, s);
    else if (F.panels_T) launch_panel(F.PT, x, epi, s);
    else launch_product<Acc, Epi>(F.AT, x, epi, s);
}

}  // namespace tv3
