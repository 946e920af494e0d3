// Fused 2D route for gfx950: X = M_v Y M_h^T on MFMA, with M = G^-1 A^T the
// explicit operator of each axis (operator_plan.hpp) and no horizontal result
// in memory.
//
// A workgroup (four wavefronts) owns kOpTileCols destination columns of one
// frame and a run of units of kOpUnit destination rows.  It walks down the
// source rows in slabs of kOpSlab:
//   1. the slab's source columns [kstart_h[first block], ... ) are staged in
//      LDS, already split into 16-bit parts;
//   2. wavefront w multiplies the slab by horizontal block w of the tile
//      (16x16x32 MFMA tiles, the operator held in registers) and writes the
//      F32 results, split again, into a ring of kOpRing source rows x 64
//      columns in LDS;
//   3. once the ring holds every source row a unit's two vertical blocks
//      read, wavefront w multiplies the vertical operator by the ring's
//      columns of block w and stores the unit's outputs.
//
// Precision.  Each operand is a sum of 16-bit parts and each product the sum
// of the part products whose order is at most the operand's own rounding:
//   - integer samples (u8/u16, values y): y / 256 = (y >> 8) + (y & 255) / 256
//     exactly as two FP16 parts, operators and the F32 ring values as FP16 +
//     FP16 remainder; all four part products in the horizontal product, three
//     in the vertical one;
//   - F32 samples, first launch: FP16 + FP16 remainder like the integer
//     route (22 bits), all four part products horizontally.  FP16 has a
//     narrow range, so every staged sample is checked: a non-finite sample,
//     or one large enough that the horizontal results could leave FP16's
//     range (job.range, from the operator's largest row L1 norm), sets the
//     call's flag;
//   - F32 samples, second launch (Exact): returns at once unless the flag is
//     set, else recomputes the call with every F32 value split exactly into
//     three bfloat16 parts by truncation (F32's own range), operators as
//     three rounded bfloat16 parts, six products.
// MFMA products of 16-bit parts are exact in F32 and accumulate in F32.  The
// integer route applies the input offset and scale to the result (the
// operator maps constants to themselves), then the output conversion.
#include <algorithm>
#include <type_traits>

#include "kernels_common.hpp"

namespace dsmvc::hip {
namespace dev {
namespace {

using u32x4 = std::uint32_t __attribute__((ext_vector_type(4)));
using u32x2 = std::uint32_t __attribute__((ext_vector_type(2)));
using h8 = _Float16 __attribute__((ext_vector_type(8)));
using h2 = _Float16 __attribute__((ext_vector_type(2)));
using f2 = float __attribute__((ext_vector_type(2)));
using bf8 = __bf16 __attribute__((ext_vector_type(8)));
using f32x4 = float __attribute__((ext_vector_type(4)));

constexpr int kThreads = 512;   // eight wavefronts: one (16-row, 16-column) MFMA block each
constexpr int kYPitch = kOpStageCols + 8;      // 16-bit values per staged source row
constexpr int kRPitch = kOpRing + 8;           // 16-bit values per ring column
constexpr int kChunksPerRow = kOpStageCols / 8;
constexpr int kChunks = (kOpSlab * kChunksPerRow + kThreads - 1) / kThreads;
// Pitches of 4 x odd dwords: 16 lanes reading 16 bytes each from 16 rows (or
// ring columns) hit 16 distinct groups of 4 banks.
static_assert((kYPitch / 2) % 8 == 4 && (kRPitch / 2) % 8 == 4, "conflict-free LDS pitches");
static_assert(kOpRing % 8 == 0 && kOpSlab % 8 == 0, "ring rows in groups of 8");

template <bool Bf>
__device__ __forceinline__ f32x4 mma(u32x4 a, u32x4 b, f32x4 c) {
    if constexpr (Bf)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                       __builtin_bit_cast(bf8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                      __builtin_bit_cast(h8, b), c, 0, 0, 0);
}

// c + sum of a[i] b[j] over the kept part pairs, smallest first.  `All`:
// every pair of two-part operands (exact products of the parts).
template <bool Bf, bool All = false>
__device__ __forceinline__ f32x4 products(const u32x4 *a, const u32x4 *b, f32x4 c) {
    if constexpr (All && !Bf) c = mma<false>(a[1], b[1], c);
    if constexpr (Bf) {
        c = mma<true>(a[2], b[0], c);
        c = mma<true>(a[1], b[1], c);
        c = mma<true>(a[0], b[2], c);
        c = mma<true>(a[1], b[0], c);
        c = mma<true>(a[0], b[1], c);
        c = mma<true>(a[0], b[0], c);
    } else {
        c = mma<false>(a[1], b[0], c);
        c = mma<false>(a[0], b[1], c);
        c = mma<false>(a[0], b[0], c);
    }
    return c;
}

__device__ __forceinline__ std::uint32_t fbits(float x) { return __builtin_bit_cast(std::uint32_t, x); }
__device__ __forceinline__ float bitsf(std::uint32_t u) { return __builtin_bit_cast(float, u); }

// Two u16 samples (low and high half of d) -> FP16 pairs of y >> 8 and
// (y & 255) / 256, exactly: a byte permute places the high bytes in the
// mantissa of 1024.0 (0x64xx) and the low bytes in that of 4.0 (0x44xx,
// spacing 1/256), and the bias is subtracted.
__device__ __forceinline__ void split_u16(std::uint32_t d, std::uint32_t &p0, std::uint32_t &p1) {
    const std::uint32_t hb = __builtin_amdgcn_perm(0x64646464U, d, 0x04030401U);
    const std::uint32_t lb = __builtin_amdgcn_perm(0x44444444U, d, 0x04020400U);
    const h2 hv = __builtin_bit_cast(h2, hb) - h2{static_cast<_Float16>(1024.0F), static_cast<_Float16>(1024.0F)};
    const h2 lv = __builtin_bit_cast(h2, lb) - h2{static_cast<_Float16>(4.0F), static_cast<_Float16>(4.0F)};
    p0 = __builtin_bit_cast(std::uint32_t, hv);
    p1 = __builtin_bit_cast(std::uint32_t, lv);
}

// Two F32 values as three packed pairs of bfloat16 parts (low half from
// x0) by truncation: the high 16 bits, then those of the remainder, then the
// remainder's remainder.  Exact: each remainder has at most the bits the
// next part keeps.  One byte permute packs each pair's high halves.
__device__ __forceinline__ void split_f32_pair(float x0, float x1, std::uint32_t &p0, std::uint32_t &p1,
                                               std::uint32_t &p2) {
    const std::uint32_t u0 = fbits(x0), u1 = fbits(x1);
    p0 = __builtin_amdgcn_perm(u1, u0, 0x07060302U);
    const float r0 = __fsub_rn(x0, bitsf(u0 & 0xFFFF0000U)), r1 = __fsub_rn(x1, bitsf(u1 & 0xFFFF0000U));
    const std::uint32_t v0 = fbits(r0), v1 = fbits(r1);
    p1 = __builtin_amdgcn_perm(v1, v0, 0x07060302U);
    const float t0 = __fsub_rn(r0, bitsf(v0 & 0xFFFF0000U)), t1 = __fsub_rn(r1, bitsf(v1 & 0xFFFF0000U));
    p2 = __builtin_amdgcn_perm(fbits(t1), fbits(t0), 0x07060302U);
}

// Two F32 values as FP16 + FP16 remainder pairs (round to nearest; packed,
// low half from x0): one packed conversion per part, the high part widened
// back from its packed form.
__device__ __forceinline__ void split_half_pair(float x0, float x1, std::uint32_t &hi, std::uint32_t &lo) {
    const h2 h = __builtin_convertvector(f2{x0, x1}, h2);
    const f2 w = __builtin_convertvector(h, f2);
    const h2 r = __builtin_convertvector(f2{__fsub_rn(x0, w[0]), __fsub_rn(x1, w[1])}, h2);
    hi = __builtin_bit_cast(std::uint32_t, h);
    lo = __builtin_bit_cast(std::uint32_t, r);
}

// Staged samples of one chunk (8 consecutive source columns of one row).
template <class In> struct Chunk;
template <> struct Chunk<std::uint16_t> { u32x4 v; };
template <> struct Chunk<std::uint8_t> { u32x2 v; };
template <> struct Chunk<float> { u32x4 v[2]; };

template <class In> __device__ __forceinline__ std::uint32_t chunk_word(const Chunk<In> &k) {
    if constexpr (std::is_same_v<In, float>) return k.v[0][0];
    else return k.v[0];
}

// Byte offset that every buffer load treats as out of range (reads 0).
constexpr int kOutside = 0x7FFFFFF0;

// One chunk from the plane's buffer resource at byte offset `off` (kOutside
// for a row past the plane).  `Vec`: the offset is aligned for one wide
// load; otherwise one load per sample.  Columns at or past the plane's width
// read whatever follows in memory (the row's padding or the next row): the
// operator's entries there are zero, and an integer sample times zero is
// zero; F32 samples there are cleared when staged (clear_past), since
// padding may hold non-finite bits.  Buffer loads never touch the LDS
// counter, so the chunk stays in flight across the wavefront's LDS work.
template <class In, bool Vec>
__device__ __forceinline__ void fetch_chunk(Chunk<In> &k, __amdgpu_buffer_rsrc_t rs, int off) {
    if constexpr (std::is_same_v<In, std::uint16_t>) {
        if constexpr (Vec) {
            k.v = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const std::uint32_t a = __builtin_amdgcn_raw_buffer_load_b16(rs, off + 4 * q, 0, 0);
                const std::uint32_t b = __builtin_amdgcn_raw_buffer_load_b16(rs, off + 4 * q + 2, 0, 0);
                k.v[q] = a | (b << 16);
            }
        }
    } else if constexpr (std::is_same_v<In, std::uint8_t>) {
        if constexpr (Vec) {
            k.v = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
        } else {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                std::uint32_t w = 0;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    w |= static_cast<std::uint32_t>(__builtin_amdgcn_raw_buffer_load_b8(rs, off + 4 * q + e, 0, 0)) << (8 * e);
                k.v[q] = w;
            }
        }
    } else {
        if constexpr (Vec) {
            k.v[0] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
            k.v[1] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16, 0, 0));
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                k.v[0][q] = __builtin_amdgcn_raw_buffer_load_b32(rs, off + 4 * q, 0, 0);
                k.v[1][q] = __builtin_amdgcn_raw_buffer_load_b32(rs, off + 16 + 4 * q, 0, 0);
            }
        }
    }
}

// F32 samples of a chunk at columns c.. that lie at or past `cols`, cleared.
__device__ __forceinline__ void clear_past(Chunk<float> &k, int c, int cols) {
    if (c + 8 > cols) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (c + q >= cols) k.v[0][q] = 0U;
            if (c + 4 + q >= cols) k.v[1][q] = 0U;
        }
    }
}

// Writes one chunk's parts to the staged rows: part p at
// ylds[p][row][col .. col + 8).  F32 samples on the FP16 scheme: returns
// true when a sample is non-finite or above `range` in magnitude.
template <class In, bool Exact>
__device__ __forceinline__ bool stage_chunk(const Chunk<In> &k, std::uint16_t *ylds, int row, int col,
                                            float range) {
    constexpr int plane = kOpSlab * kYPitch;
    u32x4 p0, p1;
    bool bad = false;
    if constexpr (std::is_same_v<In, float> && !Exact) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float x0 = bitsf(q < 2 ? k.v[0][2 * q] : k.v[1][2 * q - 4]);
            const float x1 = bitsf(q < 2 ? k.v[0][2 * q + 1] : k.v[1][2 * q - 3]);
            bad |= !(__builtin_fabsf(x0) <= range) || !(__builtin_fabsf(x1) <= range);
            std::uint32_t a, b;
            split_half_pair(x0, x1, a, b);
            p0[q] = a;
            p1[q] = b;
        }
    } else if constexpr (std::is_same_v<In, float>) {
        u32x4 p2;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            std::uint32_t a, b, c;
            const float x0 = bitsf(q < 2 ? k.v[0][2 * q] : k.v[1][2 * q - 4]);
            const float x1 = bitsf(q < 2 ? k.v[0][2 * q + 1] : k.v[1][2 * q - 3]);
            split_f32_pair(x0, x1, a, b, c);
            p0[q] = a;
            p1[q] = b;
            p2[q] = c;
        }
        *reinterpret_cast<u32x4 *>(ylds + 2 * plane + row * kYPitch + col) = p2;
    } else {
        u32x4 d;
        if constexpr (std::is_same_v<In, std::uint16_t>) {
            d = k.v;
        } else {
            // Bytes to u16 pairs.
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                d[2 * q] = (k.v[q] & 0xFFU) | ((k.v[q] & 0xFF00U) << 8);
                d[2 * q + 1] = ((k.v[q] >> 16) & 0xFFU) | ((k.v[q] >> 8) & 0x00FF0000U);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            std::uint32_t a, b;
            split_u16(d[q], a, b);
            p0[q] = a;
            p1[q] = b;
        }
    }
    *reinterpret_cast<u32x4 *>(ylds + row * kYPitch + col) = p0;
    *reinterpret_cast<u32x4 *>(ylds + plane + row * kYPitch + col) = p1;
    return bad;
}

// Workgroup b runs on XCD b % 8 (each XCD has its own L2), as the b / 8-th
// workgroup there.  With few jobs (the planes of a frame batch), every job's
// items are split into 8 contiguous shares and share x goes to XCD x, one job
// after another: each XCD then holds the same mix of long (luma) and short
// (chroma) items, and neighbouring tiles, whose source columns overlap,
// share an L2.  With many jobs (batches, served rounds) the launch's items
// are split into 8 contiguous shares, and an item's job is found by binary
// search over the jobs' first items.  The launch has 8 * (ceil(items / 8)
// + jobs) workgroups, enough for the largest share sum; workgroups past
// their XCD's shares return.  Returns the job, or -1.
constexpr long kXcds = 8;
constexpr int kBalancedJobs = 16;
__device__ __forceinline__ int find_item(const OpJob *table, int count, long items, long b, long &item) {
    const long x = b % kXcds, slot = b / kXcds;
    if (count > kBalancedJobs) {
        const long share = (items + kXcds - 1) / kXcds;
        item = x * share + slot;
        if (slot >= share || item >= items) return -1;
        int lo = 0, hi = count - 1;   // the last job whose first item is at most `item`
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (table[mid].first_item <= item) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    }
    long before = 0;
    for (int j = 0; j < count; ++j) {
        const long n = operator_items(table[j]);
        const long lo = n * x / kXcds, share = n * (x + 1) / kXcds - lo;
        if (slot < before + share) {
            item = table[j].first_item + lo + (slot - before);
            return j;
        }
        before += share;
    }
    return -1;
}

// Output sample from the result x: F32 as is; integer samples
// clamp(x * mul + add, 0, max) rounded to nearest even, mul and add folding
// the input and output conversions.  A store at or past the resource's size
// (a row past the plane) is discarded by the buffer's range check.
template <class Out> __device__ __forceinline__ void store_sample(__amdgpu_buffer_rsrc_t rs, int off, float x,
                                                                float mul, float add, float max) {
    if constexpr (std::is_same_v<Out, float>) {
        __builtin_amdgcn_raw_buffer_store_b32(fbits(x), rs, off, 0, 0);
    } else {
        // v + 2^23 rounds v in [0, 2^23) to the nearest integer, ties to
        // even, and leaves it in the low mantissa bits.
        const float v = __builtin_amdgcn_fmed3f(__builtin_fmaf(x, mul, add), 0.0F, max);
        const std::uint32_t q = fbits(v + 8388608.0F);
        if constexpr (std::is_same_v<Out, std::uint16_t>)
            __builtin_amdgcn_raw_buffer_store_b16(static_cast<std::uint16_t>(q), rs, off, 0, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b8(static_cast<std::uint8_t>(q), rs, off, 0, 0);
    }
}

// Workgroup slot b of a launch (b % 8 is its XCD).  Vec: every job of the
// launch allows 16-byte source loads (8 for u8).
template <class In, int S, bool Exact, bool Vec>
__device__ __forceinline__ void operator2d_item(const OpJob *__restrict__ table, int count, long items, long b) {
    static_assert(!Exact || std::is_same_v<In, float>, "the exact scheme is for F32 samples");
    constexpr bool Bf = Exact;                              // three bfloat16 parts
    constexpr bool Half32 = std::is_same_v<In, float> && !Exact;   // F32 samples, FP16 scheme
    constexpr int NP = Bf ? 3 : 2;
    constexpr int kSize = static_cast<int>(sizeof(In));
    constexpr int kYPlane = kOpSlab * kYPitch;          // 16-bit values per staged part
    constexpr int kRPlane = kOpTileCols * kRPitch;      // 16-bit values per ring part
    __shared__ __attribute__((aligned(16))) std::uint16_t ylds[NP * kYPlane];
    __shared__ __attribute__((aligned(16))) std::uint16_t ring[NP * kRPlane];

    long item = 0;
    const int job = find_item(table, count, items, b, item);
    if (job < 0) return;
    // The job's fields, read once (scalar registers).
    const OpJob jb = table[job];
    // The exact pass runs only when the FP16 pass found a sample it could
    // not take.
    if constexpr (Exact)
        if (uniform(__hip_atomic_load(jb.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) return;
    bool bad = false;
    const long local = item - jb.first_item;
    const int tile = static_cast<int>(local % jb.tiles);
    const long rest = local / jb.tiles;
    const int seg = static_cast<int>(rest % jb.segments);
    const long frame = rest / jb.segments;
    const int u0 = seg * jb.units_per_item;
    const int u1 = min(jb.units, u0 + jb.units_per_item);

    // Wavefront w owns horizontal block cb = w % 4 of the tile and the
    // 16-row half hb = w / 4 of every slab (stage 1) and of every unit
    // (stage 2).
    const int tid = threadIdx.x, l = tid & 63, lr = l & 15, lk = l >> 4;
    const int w = uniform(tid >> 6), cb = w & 3, hb = w >> 2;
    const ConstPtr<std::int32_t> kh = constant(jb.h.kstart), kv = constant(jb.v.kstart);
    const int cs0 = kh[4 * tile];
    const int wc = kh[4 * tile + 3] - cs0 + kOpStep * S;
    const int mh = jb.h.m, mv = jb.v.m, nh = jb.h.n, nv = jb.v.n;

    // Horizontal block 4 * tile + cb, in registers for the whole item.
    u32x4 bh[S][NP];
    {
        const GlobalPtr<u32x4> d = global(reinterpret_cast<const u32x4 *>(jb.h.data))
                                   + static_cast<long>(4 * tile + cb) * S * NP * 64 + l;
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int p = 0; p < NP; ++p) bh[s][p] = d[(s * NP + p) * 64];
    }

    // Slabs start on multiples of kOpSlab, so a slab's rows never wrap in
    // the ring; from the first unit's first source row to the last unit's last.
    const int first = kv[2 * u0] / kOpSlab * kOpSlab;
    const int slabs = (kv[2 * (u1 - 1) + 1] + kOpStep * S - first + kOpSlab - 1) / kOpSlab;

    // This thread's staging chunks (fixed for the item): slab row, column,
    // and byte offset at the first slab.
    int crow[kChunks], ccol[kChunks], coff[kChunks];
    bool cact[kChunks];
    const long s_row = jb.s_row;
#pragma unroll
    for (int i = 0; i < kChunks; ++i) {
        const int q = tid + kThreads * i;
        crow[i] = q / kChunksPerRow;
        ccol[i] = 8 * (q % kChunksPerRow);
        cact[i] = crow[i] < kOpSlab && ccol[i] < wc;
        coff[i] = static_cast<int>(((first + crow[i]) * s_row + cs0 + ccol[i]) * kSize);
    }
    const auto rs = plane_rsrc(static_cast<const In *>(jb.src) + frame * jb.s_frame,
                               (static_cast<long>(mv - 1) * s_row + mh) * kSize);
    const int slab_bytes = static_cast<int>(kOpSlab * s_row * kSize);
    // A slab at or past `slabs` is fetched with every load out of range.
    auto fetch = [&](Chunk<In> (&buf)[kChunks], int k) {
        const int rows_left = k < slabs ? mv - first - kOpSlab * k : 0;   // slab rows inside the plane
#pragma unroll
        for (int i = 0; i < kChunks; ++i) {
            const int off = cact[i] && crow[i] < rows_left ? coff[i] + k * slab_bytes : kOutside;
            fetch_chunk<In, Vec>(buf[i], rs, off);
        }
    };

    // LDS addresses (16-bit units): the A operand of stage 1, the ring
    // column this wavefront writes and reads.
    const std::uint16_t *a_at = ylds + (kOpBlock * hb + lr) * kYPitch + (kh[4 * tile + cb] - cs0) + 8 * lk;
    std::uint16_t *r_col = ring + (kOpBlock * cb + lr) * kRPitch;
    const int r_lane = kOpBlock * hb + 4 * lk;          // row of the wavefront's results in a slab

    const int col0 = kOpTileCols * tile + kOpBlock * cb + lr;
    const int d_row_bytes = static_cast<int>(jb.d_row * kSize);
    const int lane_out = 4 * lk * d_row_bytes + col0 * kSize;   // this lane's first output, unit row 0
    const float out_mul = jb.out_mul, out_add = jb.out_add, out_max = jb.conv.out_max;
    const auto ws = plane_rsrc(static_cast<In *>(jb.dst) + frame * jb.d_frame,
                               (static_cast<long>(nv - 1) * jb.d_row + nh) * kSize);
    const GlobalPtr<u32x4> vdata = global(reinterpret_cast<const u32x4 *>(jb.v.data)) + l;

    // Vertical block 2u + hb of unit u.
    u32x4 av[S][NP];
    auto load_unit = [&](int u) {
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int p = 0; p < NP; ++p)
                av[s][p] = vdata[((static_cast<long>(2 * u + hb) * S + s) * NP + p) * 64];
    };
    // Slab k x horizontal block cb, into the ring (split into parts).
    auto horizontal = [&](int k) {
        f32x4 acc = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < S; ++s) {
            u32x4 a[NP];
#pragma unroll
            for (int p = 0; p < NP; ++p) a[p] = *reinterpret_cast<const u32x4 *>(a_at + p * kYPlane + kOpStep * s);
            acc = products<Bf, true>(a, bh[s], acc);
        }
        // Results (source rows r0 + r_lane + v, column 16 cb + lr) into the
        // ring, split into parts; r0 % kOpRing is a multiple of kOpSlab.
        const int rr = (first + kOpSlab * k) % kOpRing + r_lane;
        if constexpr (Bf) {
            std::uint32_t a0, b0, c0, a1, b1, c1;
            split_f32_pair(acc[0], acc[1], a0, b0, c0);
            split_f32_pair(acc[2], acc[3], a1, b1, c1);
            *reinterpret_cast<u32x2 *>(r_col + rr) = u32x2{a0, a1};
            *reinterpret_cast<u32x2 *>(r_col + kRPlane + rr) = u32x2{b0, b1};
            *reinterpret_cast<u32x2 *>(r_col + 2 * kRPlane + rr) = u32x2{c0, c1};
        } else {
            std::uint32_t h0, l0, h1, l1;
            split_half_pair(acc[0], acc[1], h0, l0);
            split_half_pair(acc[2], acc[3], h1, l1);
            *reinterpret_cast<u32x2 *>(r_col + rr) = u32x2{h0, h1};
            *reinterpret_cast<u32x2 *>(r_col + kRPlane + rr) = u32x2{l0, l1};
        }
    };
    // Unit u: vertical block 2u + hb times the ring's column, and the lane's
    // four outputs.  Rows at or past nv fall outside `ws` (the host keeps the
    // offsets of a unit's rows below 2^31).
    auto unit = [&](int u) {
        f32x4 out = f32x4{0, 0, 0, 0};
        const int k0 = kv[2 * u + hb];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            // Wave-uniform ring row of the step's K range, then the lane's
            // 8-row group, which wraps at most once.
            const int rs = uniform((k0 + kOpStep * s) % kOpRing);
            int rr = rs + 8 * lk;
            rr = rr >= kOpRing ? rr - kOpRing : rr;
            u32x4 b[NP];
#pragma unroll
            for (int p = 0; p < NP; ++p) b[p] = *reinterpret_cast<const u32x4 *>(r_col + p * kRPlane + rr);
            out = products<Bf>(av[s], b, out);
        }
        if (col0 < nh) {
            const int base = (kOpUnit * u + kOpBlock * hb) * d_row_bytes + lane_out;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                store_sample<In>(ws, base + v * d_row_bytes, out[v], out_mul, out_add, out_max);
        }
    };
    // Whether the next unit's source rows all lie below `produced`.
    int u = u0;
    auto ready = [&](int produced) { return u < u1 && kv[2 * u + 1] + kOpStep * S <= produced; };
    // The units whose source rows all lie below `produced`: at most
    // kOpMaxUnitsPerSlab (the host checks the vertical operator).  The first
    // one's operator blocks are loaded already; further ones (near the
    // plane's edges, near-identity plans) load their own.  The loop is
    // unrolled whole: a back edge would merge its pending loads and stores
    // into the first unit's waits.  Returns whether any ran.
    auto units = [&](int produced) {
        if (!ready(produced)) return false;
        unit(u);
        ++u;
#pragma unroll
        for (int t = 1; t < kOpMaxUnitsPerSlab; ++t) {
            if (!ready(produced)) break;
            load_unit(u);
            unit(u);
            ++u;
        }
        return true;
    };

    // Step k: stage slab k (its loads were issued a step earlier), fetch
    // slab k + 1, run the unit that the slabs below k completed (its operator
    // blocks were loaded a step earlier), load the next unit's blocks, then
    // multiply slab k into the ring.  The compiler waits for every
    // outstanding vector-memory operation when stores are among them, so
    // each step has one such wait, at the staging, and everything issued
    // after it has a whole step to complete.  A unit run in step k may read
    // ring rows that slab k overwrites (jobs whose units span many rows), and
    // then a barrier follows it.  Every chunk's first word goes into
    // `sink`, which keeps the loads of inactive chunks issued (a fixed set of
    // loads per step) and makes the staging's wait one point on every path.
    Chunk<In> buf[kChunks];
    std::uint32_t sink = 0;
    load_unit(u0);
    fetch(buf, 0);
    for (int k = 0; k < slabs; ++k) {
        __syncthreads();   // every wavefront is done with the staged rows of slab k - 1
#pragma unroll
        for (int i = 0; i < kChunks; ++i) sink ^= chunk_word(buf[i]);
        asm volatile("" : "+v"(sink));
#pragma unroll
        for (int i = 0; i < kChunks; ++i) {
            if constexpr (std::is_same_v<In, float>) clear_past(buf[i], cs0 + ccol[i], mh);
            if (cact[i]) bad |= stage_chunk<In, Exact>(buf[i], ylds, crow[i], ccol[i], jb.range);
        }
        fetch(buf, k + 1);
        __syncthreads();   // slab k staged; the ring rows of slabs below k written
        if (units(first + kOpSlab * k)) {
            if (u < u1) load_unit(u);
            if (jb.unit_barrier) __syncthreads();
        }
        horizontal(k);
    }
    __syncthreads();
    units(first + kOpSlab * slabs);
    asm volatile("" ::"v"(sink));
    if constexpr (Half32)
        if (bad) __hip_atomic_fetch_or(jb.flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The first pass takes one workgroup slot per workgroup.  The exact pass,
// which usually returns at once (its job's flag is clear), runs on a fixed
// grid of kExactGrid workgroups that walk the slots (a multiple of 8, so a
// workgroup stays on its XCD): a mostly empty launch costs a few hundred
// workgroups instead of one per slot.  `shared_flag`: the one flag every
// job of the launch uses (nullptr when they differ), read before any job
// lookup.
constexpr long kExactGrid = 256;
template <class In, int S, bool Exact, bool Vec>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(Exact ? 2 : 4)))
void operator2d(const OpJob *__restrict__ table, int count, long items, long slots, const int *shared_flag) {
    if constexpr (Exact) {
        if (shared_flag && uniform(__hip_atomic_load(shared_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0)
            return;
        for (long b = blockIdx.x; b < slots; b += gridDim.x) {
            operator2d_item<In, S, Exact, Vec>(table, count, items, b);
            __syncthreads();   // the next slot reuses the LDS
        }
    } else {
        operator2d_item<In, S, Exact, Vec>(table, count, items, blockIdx.x);
    }
}

__global__ __launch_bounds__(256) void fill_i32(int *__restrict__ p, int n, int value) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i < n) p[i] = value;
}

} // namespace
} // namespace dev

hipError_t launch_fill_i32(int *p, int n, int value, hipStream_t stream) {
    if (n < 1) return hipSuccess;
    hipLaunchKernelGGL(dev::fill_i32, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, p, n,
                       value);
    return hipGetLastError();
}

template <class In, int S, bool Exact, bool Vec>
static hipError_t launch_op(const OpJob *table, int count, long items, const int *shared_flag,
                            hipStream_t stream) {
    const long slots = dev::kXcds * ((items + dev::kXcds - 1) / dev::kXcds + count);
    const long blocks = Exact ? std::min(slots, dev::kExactGrid) : slots;
    hipLaunchKernelGGL((dev::operator2d<In, S, Exact, Vec>), dim3(static_cast<unsigned>(blocks)),
                       dim3(dev::kThreads), 0, stream, table, count, items, slots, shared_flag);
    return hipGetLastError();
}

template <class In, bool Exact, bool Vec>
static hipError_t launch_op_steps(const OpJob *table, int count, long items, int steps, const int *shared_flag,
                                  hipStream_t stream) {
    switch (steps) {
    case 2: return launch_op<In, 2, Exact, Vec>(table, count, items, shared_flag, stream);
    case 3: return launch_op<In, 3, Exact, Vec>(table, count, items, shared_flag, stream);
    case 4: return launch_op<In, 4, Exact, Vec>(table, count, items, shared_flag, stream);
    default: return hipErrorInvalidValue;
    }
}

template <bool Vec>
static hipError_t launch_op_kind(const OpJob *table, int count, long items, int steps, Sample kind, bool exact,
                                 const int *shared_flag, hipStream_t stream) {
    if (exact)
        return kind == Sample::f32
                   ? launch_op_steps<float, true, Vec>(table, count, items, steps, shared_flag, stream)
                   : hipErrorInvalidValue;
    switch (kind) {
    case Sample::f32: return launch_op_steps<float, false, Vec>(table, count, items, steps, nullptr, stream);
    case Sample::u8: return launch_op_steps<std::uint8_t, false, Vec>(table, count, items, steps, nullptr, stream);
    case Sample::u16: return launch_op_steps<std::uint16_t, false, Vec>(table, count, items, steps, nullptr, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_operator_table(const OpJob *table, int count, long items, int steps, Sample kind,
                                 bool exact, bool vector, const int *shared_flag, hipStream_t stream) {
    if (count < 1 || items < 1) return hipSuccess;
    if (items + 8L * count > 0x7FFFFFF0L) return hipErrorInvalidValue;
    return vector ? launch_op_kind<true>(table, count, items, steps, kind, exact, shared_flag, stream)
                  : launch_op_kind<false>(table, count, items, steps, kind, exact, shared_flag, stream);
}

} // namespace dsmvc::hip
