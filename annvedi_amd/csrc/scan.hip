// Candidate-generation scan: approximate squared L2 of every (query, row) pair
// in a row range, computed as ||x||^2 + ||q||^2 - 2 x.q with MFMA, filtered
// against a per-query threshold and appended to per-query candidate lists.
//
// No reference counterpart beyond the role: the reference scores every row
// with a scalar/AVX loop (src/linear_index.cpp:38-59, gpu_filter.cu:52-63).
//
// Work decomposition (one workgroup = 4 waves, 256 lanes):
//   grid.x  row blocks, each a contiguous run of 16-row tiles;
//   grid.y  query tiles of 16*QW queries, staged once in LDS.
// A wave owns a 16-row tile at a time: every lane loads its row fragment
// straight from HBM in the MFMA B-operand layout (16 rows x 16 bytes per
// load instruction, no LDS round trip), and multiplies it against QW
// query fragments read from LDS.  The 16x16 fp32 accumulator holds, per lane,
// one row (n = lane % 16) and four queries (4*(lane/16) + i).
//
// Element policies fix, per MFMA family, how many dims one 16-byte lane load
// carries and how it feeds the MFMA:
//   F32  : v_mfma_f32_16x16x4_f32     16B = 4 dims  -> 4 MFMAs (K=4 each)
//   BF16 : v_mfma_f32_16x16x32_bf16   16B = 8 dims  -> 1 MFMA  (K=32 over 4 lane groups)
//   FP8  : v_mfma_f32_16x16x32_fp8    16B = 16 dims -> 2 MFMAs (K=32 each)
//   SQ8  : v_mfma_i32_16x16x64_i8     16B = 16 dims -> 1 MFMA  (K=64)
//   SQ4  : v_mfma_i32_16x16x64_i8     8B packed = 16 nibbles -> 1 MFMA (K=64)
// Lane group g = lane/16 always covers the same dims for the query (A) and
// the row (B) operand, so the contraction pairs the right elements whatever
// the hardware's internal k order.
#include "common.h"

#include <cstdio>
#include <cstdlib>
#include <type_traits>

namespace {

constexpr int kThreads = 256;

// ----------------------------------------------------------------- policies
struct PolF32 {
    static constexpr int kDimsPerLoad = 4;     // per lane per 16B load
    static constexpr int kChunk = 16;          // dims per wave-wide load step (4 lane groups)
    static constexpr int kElemBytes = 4;
    static constexpr int kRing = 2;       // chunks of row fragments in flight per wave (wide scan)
    typedef v4f Acc;
    typedef float El;  // accumulator element
    __device__ static inline void mma(v4f& acc, const i32x4& a, const i32x4& b) {
        f32x4 af = __builtin_bit_cast(f32x4, a);
        f32x4 bf = __builtin_bit_cast(f32x4, b);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[0], bf[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[1], bf[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[2], bf[2], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[3], bf[3], acc, 0, 0, 0);
    }
    // acc = a.b (zero C operand)
    __device__ static inline void mma0(v4f& acc, const i32x4& a, const i32x4& b) {
        acc = (v4f){0.f, 0.f, 0.f, 0.f};
        mma(acc, a, b);
    }
};

struct PolBF16 {
    static constexpr int kDimsPerLoad = 8;
    static constexpr int kChunk = 32;
    static constexpr int kElemBytes = 2;
    static constexpr int kRing = 4;       // chunks of row fragments in flight per wave (wide scan)
    typedef v4f Acc;
    typedef float El;  // accumulator element
    __device__ static inline void mma(v4f& acc, const i32x4& a, const i32x4& b) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                      __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
    }
    __device__ static inline void mma0(v4f& acc, const i32x4& a, const i32x4& b) {
        acc = (v4f){0.f, 0.f, 0.f, 0.f};
        mma(acc, a, b);
    }
};

struct PolFP8 {
    static constexpr int kDimsPerLoad = 16;
    static constexpr int kChunk = 64;
    static constexpr int kElemBytes = 1;
    static constexpr int kRing = 4;       // chunks of row fragments in flight per wave (wide scan)
    typedef v4f Acc;
    typedef float El;  // accumulator element
    __device__ static inline void mma(v4f& acc, const i32x4& a, const i32x4& b) {
        long a0 = ((long)(uint32_t)a[1] << 32) | (uint32_t)a[0];
        long a1 = ((long)(uint32_t)a[3] << 32) | (uint32_t)a[2];
        long b0 = ((long)(uint32_t)b[1] << 32) | (uint32_t)b[0];
        long b1 = ((long)(uint32_t)b[3] << 32) | (uint32_t)b[2];
        acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a1, b1, acc, 0, 0, 0);
    }
    __device__ static inline void mma0(v4f& acc, const i32x4& a, const i32x4& b) {
        acc = (v4f){0.f, 0.f, 0.f, 0.f};
        mma(acc, a, b);
    }
};

// FP8 on the gfx950 16x16x128 f8f6f4 MFMA (e4m3 operands, unit scales),
// which issues at twice the rate of 16x16x32 fp8 (MI355X, register
// operands: 4.98 vs 2.32 PFLOP/s): the wide scan pairs consecutive 64-dim
// chunks, a lane's 16 bytes of each making its 32 bytes of one 128-deep K
// step.  Queries and rows take the same order, which is all a dot product
// needs.  The tiled layout and the fragments are those of PolFP8.
struct PolFP8P : PolFP8 {
    __device__ static inline void mma8(v4f& acc, const i32x8& a, const i32x8& b) {
        // formats 0 / 0: fp8 e4m3 for both operands; scale exponents 127: 2^0
        acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc, 0, 0, 0, 127, 0, 127);
    }
};
// half h of an 8-register operand (the pair's chunk c - 1 or c), written in
// place so the MFMA takes the tuple without copies
__device__ __forceinline__ void set_half(i32x8& v, int h, const i32x4& x) {
    if (h) {
        v[4] = x[0], v[5] = x[1], v[6] = x[2], v[7] = x[3];
    } else {
        v[0] = x[0], v[1] = x[1], v[2] = x[2], v[3] = x[3];
    }
}
template <class P>
struct IsPair : std::false_type {};
template <>
struct IsPair<PolFP8P> : std::true_type {};

// int8 MFMA with an int32 accumulator.
struct PolSQ8 {
    static constexpr int kDimsPerLoad = 16;
    static constexpr int kChunk = 64;
    static constexpr int kElemBytes = 1;
    static constexpr int kRing = 4;       // chunks of row fragments in flight per wave (wide scan)
    typedef i32x4 Acc;
    typedef int El;  // accumulator element
    __device__ static inline void mma(i32x4& acc, const i32x4& a, const i32x4& b) {
        // codes c in [0,255] are stored as c ^ 0x80, which the MFMA reads as
        // the signed int8 c - 128.  The query side adds 128*sum(w) back.
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc, 0, 0, 0);
    }
    __device__ static inline void mma0(i32x4& acc, const i32x4& a, const i32x4& b) {
        acc = (i32x4){0, 0, 0, 0};
        mma(acc, a, b);
    }
};

// SQ4: 8 packed bytes per lane per chunk of 64 dims.  Unpacked as
// [lo nibbles of bytes 0..7, hi nibbles of bytes 0..7] (values 0..15, exact
// int8); the query operand is stored in the same permuted order.
struct PolSQ4 {
    static constexpr int kDimsPerLoad = 16;
    static constexpr int kChunk = 64;
    static constexpr int kElemBytes = 0;  // packed: 2 dims per byte
    static constexpr int kRing = 4;       // chunks of row fragments in flight per wave (wide scan)
    typedef i32x4 Acc;
    typedef int El;  // accumulator element
    __device__ static inline void mma(i32x4& acc, const i32x4& a, const i32x4& b) {
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc, 0, 0, 0);
    }
    __device__ static inline void mma0(i32x4& acc, const i32x4& a, const i32x4& b) {
        acc = (i32x4){0, 0, 0, 0};
        mma(acc, a, b);
    }
};


// Row fragments.  A lane's bytes of one row chunk are loaded raw (16 bytes;
// SQ4: 8 packed bytes) and turned into the MFMA B operand where they are
// used, so a prefetched load is not waited for before its MFMA.
template <class P>
struct RawOf {
    typedef i32x4 T;
};
template <>
struct RawOf<PolSQ4> {
    typedef uint2 T;
};

// bytes of one row chunk, and of one lane group's share of it
template <class P>
__host__ __device__ constexpr int chunk_bytes() {
    return P::kElemBytes == 0 ? P::kChunk / 2 : P::kChunk * P::kElemBytes;
}
template <class P>
__host__ __device__ constexpr int lane_bytes() {
    return P::kElemBytes == 0 ? P::kDimsPerLoad / 2 : P::kDimsPerLoad * P::kElemBytes;
}

template <class P>
__device__ inline typename RawOf<P>::T load_raw(const uint8_t* p) {
    return *reinterpret_cast<const typename RawOf<P>::T*>(p);
}

template <class P>
__device__ inline i32x4 to_frag(const typename RawOf<P>::T& v) {
    if constexpr (P::kElemBytes == 0) {
        // SQ4: chunk c covers 64 dims = 32 packed bytes; lane group g holds
        // bytes [32c + 8g, +8), unpacked to 16 int8: low nibbles, then high
        uint32_t lo0 = v.x & 0x0F0F0F0Fu, lo1 = v.y & 0x0F0F0F0Fu;
        uint32_t hi0 = (v.x >> 4) & 0x0F0F0F0Fu, hi1 = (v.y >> 4) & 0x0F0F0F0Fu;
        return (i32x4){(int)lo0, (int)lo1, (int)hi0, (int)hi1};
    } else {
        return v;
    }
}

// Fragment of row `row` covering dims [dim0, dim0 + kDimsPerLoad), in the
// row-major or the tiled layout (tiled_off).
template <class P>
__device__ inline i32x4 load_row_frag(const uint8_t* X, int64_t xstride, int64_t row, int dim0, int tiled) {
    const int64_t off = P::kElemBytes == 0 ? (dim0 >> 1) : (int64_t)dim0 * P::kElemBytes;
    const int64_t at = tiled ? tiled_off(row, off, xstride, chunk_bytes<P>()) : row * xstride + off;
    return to_frag<P>(load_raw<P>(X + at));
}

// Query tile in LDS: 16*QW queries x kdim elements, stored in the scan
// element type, row stride padded by 16 bytes against bank conflicts.
template <class P>
__host__ __device__ constexpr int qrow_bytes(int kdim) {
    int b = P::kElemBytes == 0 ? kdim : kdim * P::kElemBytes;  // SQ4 query weights are int8 per dim
    return b + 16;
}

// Reduce the per-lane running (dist, row) minima over the 16 lanes that hold
// the same queries (different rows) and publish them with a 64-bit atomicMin.
template <int QS>
__device__ inline void argmin_flush(unsigned long long (&best)[QS][4], const ScanArgs& a, int q0, int g, int n) {
#pragma unroll
    for (int w = 0; w < QS; ++w)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            unsigned long long v = best[w][i];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                unsigned long long u = __shfl_xor(v, o, 16);
                v = u < v ? u : v;
            }
            const int qi = q0 + 16 * w + 4 * g + i;
            if (n == 0 && qi < a.B && v != ~0ull) atomicMin(&a.argmin[qi], v);
        }
}

template <class P, int QW>
__global__ __launch_bounds__(kThreads) void scan_kernel(ScanArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;   // lane group
    const int n = lane & 15;   // row within tile / query within subtile (A operand)
    const int qtile = 16 * QW;
    const int q0 = blockIdx.y * qtile;
    const int qbytes = P::kElemBytes == 0 ? a.kdim : a.kdim * P::kElemBytes;
    const int ldq = qrow_bytes<P>(a.kdim);

    // stage the query tile (zero rows past B)
    for (int idx = threadIdx.x; idx < qtile * (qbytes / 16); idx += kThreads) {
        int r = idx / (qbytes / 16);
        int c = idx - r * (qbytes / 16);
        i32x4 v = (i32x4){0, 0, 0, 0};
        if (q0 + r < a.B) v = reinterpret_cast<const i32x4*>(static_cast<const uint8_t*>(a.Q) + (int64_t)(q0 + r) * qbytes)[c];
        *reinterpret_cast<i32x4*>(smem + r * ldq + c * 16) = v;
    }
    // per-lane thresholds / norms for the 4*QW queries this lane scores
    float thr[QW][4], qn[QW][4], qs[QW][4];
#pragma unroll
    for (int w = 0; w < QW; ++w)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int qi = q0 + 16 * w + 4 * g + i;
            bool ok = qi < a.B;
            thr[w][i] = ok ? a.thr[qi] : -__builtin_inff();
            qn[w][i] = ok ? a.qn[qi] : 0.0f;
            qs[w][i] = (ok && a.qscale) ? a.qscale[qi] : 1.0f;
        }
    __syncthreads();

    const int64_t rows = a.row_end - a.row_begin;
    const int64_t tiles = (rows + 15) >> 4;
    const int64_t tiles_per_block = (tiles + gridDim.x - 1) / gridDim.x;
    const int64_t t_begin = (int64_t)blockIdx.x * tiles_per_block;
    const int64_t t_end = t_begin + tiles_per_block < tiles ? t_begin + tiles_per_block : tiles;
    const uint8_t* X = static_cast<const uint8_t*>(a.X);
    const int nchunk = a.kdim / P::kChunk;
    const int lane_dim = P::kDimsPerLoad * g;  // dims offset of this lane group inside a chunk
    unsigned long long best[QW][4];
#pragma unroll
    for (int w = 0; w < QW; ++w)
#pragma unroll
        for (int i = 0; i < 4; ++i) best[w][i] = ~0ull;

    // Distance, filter and append for the 16-row tile t (acc complete).
    auto epilogue = [&](int64_t t, typename P::Acc (&acc)[QW]) {
        const int64_t row = a.row_begin + t * 16 + n;
        const bool row_ok = row < a.row_end;
        const int64_t lrow = row_ok ? row : a.row_end - 1;
        const float xn = a.xn[lrow];
        if (a.argmin) {
#pragma unroll
            for (int w = 0; w < QW; ++w)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float dist = xn + qn[w][i] - 2.0f * qs[w][i] * (float)acc[w][i];
                    unsigned long long kv = ((unsigned long long)f2key(dist) << 32) | (uint32_t)row;
                    if (row_ok && kv < best[w][i]) best[w][i] = kv;
                }
            return;
        }
        // Two phases: the tile's list reservations (one returning atomic per
        // query and 16-row group with passing rows) are all issued before
        // any result is used, so their round trips overlap instead of
        // chaining (early ranges, where most pairs pass, waited on up to
        // 4 * QW of them in a row per tile).
        float dist[QW][4];
        uint32_t gm[QW][4], base[QW][4];
        uint32_t live = 0;  // bit 4w+i: some lane of the wave passes (wave-uniform)
#pragma unroll
        for (int w = 0; w < QW; ++w) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float dot = (float)acc[w][i];
                // dist = xn + qn - 2 * qscale * dot  (qscale folds the query/base scaling)
                dist[w][i] = xn + qn[w][i] - 2.0f * qs[w][i] * dot;
                const bool pass = row_ok && dist[w][i] <= thr[w][i];
                const uint64_t m = __ballot(pass);
                gm[w][i] = (uint32_t)(m >> (16 * g)) & 0xFFFFu;
                base[w][i] = 0;
                if (m != 0) {
                    live |= 1u << (4 * w + i);
                    if (n == 0 && gm[w][i])
                        base[w][i] = atomicAdd(&a.cand.cnt[q0 + 16 * w + 4 * g + i], (uint32_t)__builtin_popcount(gm[w][i]));
                }
            }
        }
        if (live == 0) return;
#pragma unroll
        for (int w = 0; w < QW; ++w) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (!((live >> (4 * w + i)) & 1u)) continue;
                const uint32_t b = __shfl(base[w][i], lane & ~15);
                if ((gm[w][i] >> n) & 1u) {
                    const uint32_t off = b + __builtin_popcount(gm[w][i] & ((1u << n) - 1u));
                    if (off < (uint32_t)a.cand.cap) {
                        const int64_t slot = (int64_t)(q0 + 16 * w + 4 * g + i) * a.cand.cap + off;
                        a.cand.key[slot] = dist[w][i];
                        a.cand.rid[slot] = a.rid_map ? a.rid_map[row] : (uint32_t)row;
                    }
                }
            }
        }
    };
    // acc[w] += query fragments of chunk c (LDS) x row fragment b
    auto mma_chunk = [&](typename P::Acc (&acc)[QW], int c, const i32x4& b) {
#pragma unroll
        for (int w = 0; w < QW; ++w) {
            const int qoff = (16 * w + n) * ldq;
            int eoff = P::kElemBytes == 0 ? (c * P::kChunk + lane_dim) : (c * P::kChunk + lane_dim) * P::kElemBytes;
            i32x4 qa = *reinterpret_cast<const i32x4*>(smem + qoff + eoff);
            P::mma(acc[w], qa, b);
        }
    };
    const int64_t wt0 = t_begin + wave;  // the wave's tiles: wt0, wt0 + 4, ...
    // Double-buffered stream (chunk count a multiple of U): the wave's tiles
    // form one sequence of batches of U chunks, and batch j + 1 (of this tile
    // or the next) is requested before the MFMAs of batch j, so loads stay in
    // flight while the wave computes.
    auto stream = [&](auto uc) {
        constexpr int U = decltype(uc)::value;
        const int nb = nchunk / U;
        const int64_t ntw = wt0 < t_end ? (t_end - wt0 + 3) / 4 : 0;
        const int64_t npos = ntw * nb;
        typename P::Acc acc[QW];
        i32x4 b0[U], b1[U];
        auto load = [&](i32x4 (&b)[U], int64_t p) {
            const int64_t t = wt0 + 4 * (p / nb);
            const int c0 = (int)(p % nb) * U;
            const int64_t row = a.row_begin + t * 16 + n;
            const int64_t lrow = row < a.row_end ? row : a.row_end - 1;
#pragma unroll
            for (int u = 0; u < U; ++u) b[u] = load_row_frag<P>(X, a.xstride, lrow, (c0 + u) * P::kChunk + lane_dim, a.tiled);
        };
        auto step = [&](const i32x4 (&b)[U], int64_t p) {
            const int cb = (int)(p % nb);
            if (cb == 0) {
#pragma unroll
                for (int w = 0; w < QW; ++w) acc[w] = typename P::Acc{};
            }
#pragma unroll
            for (int u = 0; u < U; ++u) mma_chunk(acc, cb * U + u, b[u]);
            if (cb == nb - 1) epilogue(wt0 + 4 * (p / nb), acc);
        };
        if (npos > 0) load(b0, 0);
        for (int64_t p = 0; p < npos; p += 2) {
            if (p + 1 < npos) load(b1, p + 1);
            step(b0, p);
            if (p + 1 < npos) {
                if (p + 2 < npos) load(b0, p + 2);
                step(b1, p + 1);
            }
        }
    };
    // (small query tiles only: with 64 queries the second buffer would cost
    // the kernel half its waves)
    if (QW <= 2 && nchunk % 8 == 0) {
        stream(std::integral_constant<int, 8>{});
    } else if (QW <= 2 && nchunk % 6 == 0) {  // d = 384 in 64-dim chunks (byte codes), 32-dim (bf16)
        stream(std::integral_constant<int, 6>{});
    } else if (QW <= 2 && nchunk % 4 == 0) {
        stream(std::integral_constant<int, 4>{});
    } else {
        for (int64_t t = wt0; t < t_end; t += 4) {
            const int64_t row = a.row_begin + t * 16 + n;
            const int64_t lrow = row < a.row_end ? row : a.row_end - 1;
            typename P::Acc acc[QW];
#pragma unroll
            for (int w = 0; w < QW; ++w) acc[w] = typename P::Acc{};
            constexpr int U = 8;
            i32x4 buf[U];
            int c = 0;
            for (; c + U <= nchunk; c += U) {
#pragma unroll
                for (int u = 0; u < U; ++u)
                    buf[u] = load_row_frag<P>(X, a.xstride, lrow, (c + u) * P::kChunk + lane_dim, a.tiled);
#pragma unroll
                for (int u = 0; u < U; ++u) mma_chunk(acc, c + u, buf[u]);
            }
            for (; c < nchunk; ++c) mma_chunk(acc, c, load_row_frag<P>(X, a.xstride, lrow, c * P::kChunk + lane_dim, a.tiled));
            epilogue(t, acc);
        }
    }
    if (a.argmin) argmin_flush<QW>(best, a, q0, g, n);
}

template <class P>
const char* pol_name();
template <> const char* pol_name<PolF32>() { return "f32"; }
template <> const char* pol_name<PolBF16>() { return "bf16"; }
template <> const char* pol_name<PolFP8>() { return "fp8"; }
template <> const char* pol_name<PolFP8P>() { return "fp8x128"; }
template <> const char* pol_name<PolSQ8>() { return "sq8"; }
template <> const char* pol_name<PolSQ4>() { return "sq4"; }

thread_local const char* t_variant = "";

template <class P, int QW>
hipError_t launch_t(const ScanArgs& a, hipStream_t st) {
    static const char* const name = [] {  // built once per instantiation (thread-safe static init)
        static char buf[64];
        snprintf(buf, sizeof(buf), "scan_kernel<%s,QW=%d>", pol_name<P>(), QW);
        return buf;
    }();
    t_variant = name;
    const int qtile = 16 * QW;
    const int qtiles = (a.B + qtile - 1) / qtile;
    const int64_t rows = a.row_end - a.row_begin;
    const int64_t tiles = (rows + 15) / 16;
    // aim for ~8 workgroups per CU across all query tiles, 4 tiles a wave
    // (its double-buffered stream runs across them), but at least 64
    // workgroups a query tile down to one tile a wave: a small range (the
    // first ones of a step) otherwise runs on a few workgroups whose tiles
    // follow one another at load and list-reservation latency (1024 rows x
    // 128 queries: 58 us on 8 workgroups, 21 us on 32)
    int64_t want = (2048 + qtiles - 1) / qtiles;
    int64_t maxb = (tiles + 15) / 16;
    const int64_t spread = (tiles + 3) / 4 < 64 ? (tiles + 3) / 4 : 64;
    if (maxb < spread) maxb = spread;
    int64_t gx = want < maxb ? want : maxb;
    if (gx < 1) gx = 1;
    size_t lds = (size_t)qtile * qrow_bytes<P>(a.kdim);
    if (lds > 65536) {  // long rows (one 16-query subtile): up to the 160 KB of LDS
        static const bool attr = [] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&scan_kernel<P, QW>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            return true;
        }();
        (void)attr;
    }
    ANNMI_LAUNCH((scan_kernel<P, QW>), dim3((unsigned)gx, (unsigned)qtiles), dim3(kThreads), lds, st, a);
    return hipGetLastError();
}


// ------------------------------------------------------------------ wide --
// Large batches: query-stationary kernel over short row chunks.  A workgroup
// (4-8 waves) stages a tile of 16*QS queries in LDS and scores it against
// one chunk of rows; every wave takes 16*RS rows at a time and multiplies
// them, fragment by fragment, against all 16*QS queries.  Row fragments go
// from L2/HBM straight into MFMA B registers through a ring of kRing chunks
// in flight; query fragments come from LDS one chunk ahead of their MFMAs.
//
// Workgroup -> work mapping is XCD-aware: dispatch places workgroup b on XCD
// b % 8, and each XCD walks its own contiguous run of row chunks, row chunk
// major and query tile minor.  The workgroups resident on an XCD at any time
// therefore cover one or a few row chunks for all query tiles and share
// those rows through the XCD's L2: HBM sees each row about once per launch.
//
// Passing (query, row) pairs are queued per wave in LDS (slot from an LDS
// atomic) and moved to the per-query lists in bulk, so the scan never waits
// on a returning global atomic, which would also wait for the row loads in
// flight.  A full queue spills single entries straight to the lists.

struct WideQueue {
    uint32_t* row;  // (row - chunk begin) << 7 | query slot in the tile
    float* key;
    uint32_t* count;  // LDS counter of this wave
    uint32_t* hist;   // [128] per-slot counts, then next list slots, of a flush
    int cap;
};

// LDS words a wave's flush histogram takes (one per query slot of a tile)
constexpr int kFlushSlots = 128;

// Append one pair (lanes diverge: only passing lanes get here).
__device__ inline void wide_append(const ScanArgs& a, const WideQueue& wq, int qi, int m, uint32_t roff,
                                   uint32_t row, float dist) {
    const uint32_t pos = atomicAdd(wq.count, 1u);
    if (pos < (uint32_t)wq.cap) {
        wq.row[pos] = (roff << 7) | (uint32_t)m;
        wq.key[pos] = dist;
        return;
    }
    // the list address is formed here, on this rare path: the query index
    // is opaque to the optimiser, which would otherwise hoist one 64-bit
    // list offset per (query, lane) out of the tile loop and hold (or
    // spill) them all for the whole kernel
    asm volatile("" : "+v"(qi));
    const uint32_t off = atomicAdd(&a.cand.cnt[qi], 1u);
    if (off < (uint32_t)a.cand.cap) {
        const int64_t sl = (int64_t)qi * a.cand.cap + off;
        a.cand.key[sl] = dist;
        a.cand.rid[sl] = a.rid_map ? a.rid_map[row] : row;
    }
}

// Move the wave's queued pairs to the per-query lists, aggregated per query:
// count the entries of each query slot in LDS, reserve each query's space
// with one global atomic (at most 128 a flush, not one per entry: a range
// appending ~R * rho pairs a query spent about half of its first wide launch
// on those atomics), then scatter with LDS offsets.  A list's entries are a
// set; the order in which a flush writes them does not matter.
__device__ inline void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline void wide_flush(const ScanArgs& a, const WideQueue& wq, int q0, int64_t r_begin, int lane) {
    wave_sync_lds();
    uint32_t cnt = __builtin_amdgcn_readfirstlane(*wq.count);
    const int count = cnt < (uint32_t)wq.cap ? (int)cnt : wq.cap;
    if (count > 0) {
        for (int m = lane; m < kFlushSlots; m += 64) wq.hist[m] = 0u;
        wave_sync_lds();
        for (int e = lane; e < count; e += 64) atomicAdd(&wq.hist[wq.row[e] & 127u], 1u);
        wave_sync_lds();
        for (int m = lane; m < kFlushSlots; m += 64) {
            const uint32_t c = wq.hist[m];
            wq.hist[m] = c ? atomicAdd(&a.cand.cnt[q0 + m], c) : 0u;
        }
        wave_sync_lds();
        constexpr int U = 4;
        for (int e0 = 0; e0 < count; e0 += 64 * U) {
            uint32_t v[U], off[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int e = e0 + 64 * j + lane;
                v[j] = e < count ? wq.row[e] : 0u;
                off[j] = e < count ? atomicAdd(&wq.hist[v[j] & 127u], 1u) : ~0u;
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int e = e0 + 64 * j + lane;
                if (off[j] < (uint32_t)a.cand.cap) {
                    const int qi = q0 + (int)(v[j] & 127u);
                    const uint32_t row = (uint32_t)(r_begin + (v[j] >> 7));
                    const int64_t sl = (int64_t)qi * a.cand.cap + off[j];
                    a.cand.key[sl] = wq.key[e];
                    a.cand.rid[sl] = a.rid_map ? a.rid_map[row] : row;
                }
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) *wq.count = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// LDS of a wide-scan workgroup: the query tile (QT rows of ldq bytes), five
// words per query (threshold, norm, scale, float bound, global-form bound),
// one queue counter per wave, then the waves' candidate queues.
struct WideLds {
    float* thr;
    float* qn;
    float* qsc;
    float* bnd;
    void* bndg;  // El[QT]
    uint32_t* cnt;
};

// Stage a workgroup's query tile and per-query bounds (every thread takes
// part; ends with a barrier).  Queries past B are zero rows with threshold
// -inf, which no pair passes.
template <class El>
__device__ inline WideLds wide_stage(const ScanArgs& a, uint8_t* smem, int QT, int ldq, int qbytes, int q0) {
    WideLds L;
    L.thr = reinterpret_cast<float*>(smem + QT * ldq);
    L.qn = L.thr + QT;
    L.qsc = L.qn + QT;
    L.bnd = L.qsc + QT;
    El* bndg_s = reinterpret_cast<El*>(L.bnd + QT);
    L.bndg = bndg_s;
    L.cnt = reinterpret_cast<uint32_t*>(bndg_s + QT);
    const int nthr = (int)blockDim.x;
    // the query tile: eight 16-byte loads per lane in flight
    {
        const int per_row = qbytes / 16;
        const int total = QT * per_row;
        constexpr int U = 8;
        for (int i0 = 0; i0 < total; i0 += nthr * U) {
            i32x4 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = i0 + u * nthr + (int)threadIdx.x;
                const int r = idx / per_row;
                v[u] = (i32x4){0, 0, 0, 0};
                if (idx < total && q0 + r < a.B)
                    v[u] = reinterpret_cast<const i32x4*>(static_cast<const uint8_t*>(a.Q) +
                                                          (int64_t)(q0 + r) * qbytes)[idx - r * per_row];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = i0 + u * nthr + (int)threadIdx.x;
                const int r = idx / per_row;
                if (idx < total) *reinterpret_cast<i32x4*>(smem + r * ldq + (idx - r * per_row) * 16) = v[u];
            }
        }
    }
    for (int m = threadIdx.x; m < QT; m += nthr) {
        int qi = q0 + m;
        bool ok = qi < a.B;
        const float th = ok ? a.thr[qi] : -__builtin_inff();
        const float qv = ok ? a.qn[qi] : 0.f;
        const float qs = (ok && a.qscale) ? a.qscale[qi] : 1.f;
        L.thr[m] = th;
        L.qn[m] = qv;
        L.qsc[m] = qs;
        // Prefilter bound: a pair passes dist = (xn + qv) - 2*qs*dot <= th
        // only if 2*qs*dot - xn >= qv - th - (rounding), so that value
        // reaching bnd[m] is necessary.  The 2^-14 relative margin is
        // ~1000x the fp32 rounding of the test; the exact test decides.
        // th = -inf (padding) gives NaN: never passes; th = +inf: -inf.
        const float M = a.xn_amax + fabsf(qv) + fabsf(th);
        L.bnd[m] = 0.5f * ((qv - th) - M * 0x1p-14f);  // halved: the test scales by 1/2
        // global form (prefilter 2): xn >= xn_lo for every row, so
        // dot >= ((xn_lo + qv) - th - margin) / (2 qs) is necessary
        const float sc = 2.0f * qs;
        const float num = ((a.xn_lo + qv) - th) - M * 0x1p-14f;
        const float bg = sc > 0.f ? num / sc : (num <= 0.f ? -__builtin_inff() : __builtin_inff());
        if constexpr (std::is_same<El, float>::value) {
            bndg_s[m] = bg;
        } else {  // integer dot: the smallest integer >= bg; NaN (padding) never passes
            bndg_s[m] = bg != bg ? 0x7fffffff
                                 : (bg <= -2147483648.0f ? (int)0x80000000
                                                         : (bg >= 2147483647.0f ? 0x7fffffff : (int)ceilf(bg)));
        }
    }
    if ((int)threadIdx.x < nthr / 64) L.cnt[threadIdx.x] = 0u;
    __syncthreads();
    return L;
}

// Prefilter 3 (integer dots): the accumulators of row r start at -D_r,
// D_r = floor((xn_r - xn_lo) * kap) with kap = 1 / (2 * max qs over the
// tile's queries) <= 1 / (2 qs) for each of them.  A pass needs
// dot >= ((xn_lo + qv) - th - margin) / (2 qs) + (xn_r - xn_lo) / (2 qs),
// so dot - D_r >= the global form's integer bound is necessary, and one
// integer compare per (query, lane) is as tight as the per-row bound when
// the tile's query scales are equal.  The exact test adds D_r back, so it
// sees the same integer dot.  Returns 0 (no offsets) in the other modes.
__device__ inline float wide_kap(const ScanArgs& a, const WideLds& L, int QT, int lane) {
    if (a.prefilter != 3) return 0.f;
    float mq = 0.f;
    for (int m = lane; m < QT; m += 64)
        if (L.thr[m] != -__builtin_inff()) mq = fmaxf(mq, L.qsc[m]);
#pragma unroll
    for (int o = 32; o; o >>= 1) mq = fmaxf(mq, __shfl_xor(mq, o));
    // the factor keeps the rounding of kap and of the product below the
    // exact quotient
    return (mq > 0.f && mq < __builtin_inff()) ? (0.5f / mq) * (1.0f - 0x1p-20f) : 0.f;
}

// Row offset D = floor((xn - xn_lo) * kap) of prefilter 3, clamped to
// [0, 2^24] (a smaller offset keeps the bound necessary); 0 when kap is 0.
__device__ inline int wide_row_offset(const ScanArgs& a, float xn, float kap) {
    const float v = (xn - a.xn_lo) * kap;  // NaN (kap 0, xn_lo -inf) gives 0
    return v > 0.f ? (v < 16777216.f ? (int)v : 16777216) : 0;
}

// KD > 0: the padded dim, fixed at compile time (query-tile offsets become
// instruction offsets and the chunk loop unrolls whole); 0: a.kdim.
template <class P, int QS, int RS, int kWideWaves, int QH, bool kQPre, int KD>
__global__ __launch_bounds__(64 * kWideWaves) __attribute__((amdgpu_waves_per_eu(kWideWaves >= 12 ? 3 : 1)))
void scan_wide_kernel(ScanArgs a, int qtiles, int64_t chunk_rows, int64_t nrc, int qcap) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int n = lane & 15;
    // work item of this workgroup: XCD x owns row chunks [x*per, (x+1)*per)
    const int xcd = blockIdx.x & 7;
    const int64_t s = blockIdx.x >> 3;
    const int64_t per = (nrc + 7) / 8;
    if (s / qtiles >= per) return;
    const int64_t rc = xcd * per + s / qtiles;
    if (rc >= nrc) return;
    const int qtile = (int)(s % qtiles);
    const int64_t r_begin = a.row_begin + rc * chunk_rows;
    const int64_t r_end = r_begin + chunk_rows < a.row_end ? r_begin + chunk_rows : a.row_end;

    // QH > 1: the waves split the query tile QH ways and share row tiles
    constexpr int QT = 16 * QS * QH;
    constexpr int NS = kWideWaves / QH;  // row streams
    static_assert(QT <= 128, "queue entries hold the query slot in 7 bits");
    const int qbase = 16 * QS * (wave % QH);
    const int stream = wave / QH;
    const int q0 = qtile * QT;
    const int kdim = KD > 0 ? KD : a.kdim;
    const int qbytes = P::kElemBytes == 0 ? kdim : kdim * P::kElemBytes;
    const int ldq = qrow_bytes<P>(kdim);
    typedef typename P::El El;
    const WideLds L = wide_stage<El>(a, smem, QT, ldq, qbytes, q0);
    const float* thr_s = L.thr;
    const float* qn_s = L.qn;
    const float* qsc_s = L.qsc;
    const float* bnd_s = L.bnd;
    const El* bndg_s = static_cast<const El*>(L.bndg);
    WideQueue wq;
    wq.row = L.cnt + kWideWaves + (int64_t)wave * 2 * qcap;
    wq.key = reinterpret_cast<float*>(wq.row + qcap);
    wq.count = L.cnt + wave;
    wq.hist = L.cnt + kWideWaves + (int64_t)kWideWaves * 2 * qcap + wave * kFlushSlots;
    wq.cap = qcap;
    constexpr bool kIntDot = std::is_same<El, int>::value;
    constexpr bool kPair = IsPair<P>::value;
    static_assert(!kPair || (!kQPre && P::kRing % 2 == 0), "pairs: in-slot query reads, an even ring");
    const float kap = kIntDot ? wide_kap(a, L, QT, lane) : 0.f;

    const uint8_t* X = static_cast<const uint8_t*>(a.X);
    const int nchunk = kdim / P::kChunk;
    const int lane_dim = P::kDimsPerLoad * g;
    constexpr int TR = 16 * RS;  // rows per wave step
    constexpr int D = P::kRing;
    static_assert(D % 2 == 0, "query fragments alternate between two buffers per ring slot");
    constexpr int CB = chunk_bytes<P>();
    const int npad = (nchunk + D - 1) / D * D;  // chunks per tile, rounded up to whole ring turns
    const int64_t ntiles = (r_end - r_begin + TR - 1) / TR;
    if (stream >= ntiles) return;

    // Row fragments stream through a ring of D chunks: the chunk sequence of
    // the wave's tiles is one stream (tile t chunks 0..npad-1, then tile
    // t + NS), and the load of stream position j + D is issued right after
    // the MFMAs of position j, so D chunks are in flight across tile
    // boundaries too.  Positions past the last chunk of a tile reload its
    // last chunk and skip their MFMAs.  With kQPre, query fragments
    // alternate between two register sets: slot s multiplies qa[s & 1] and
    // fills qa[(s+1) & 1] with the next chunk's (chunk 0 after the last);
    // otherwise they are read where they are used and a second wave on the
    // SIMD hides the LDS latency.
    // Loads address rows by 32-bit byte offsets from the chunk's first row
    // (a uniform base, so the offset is the only per-lane operand).
    typedef typename RawOf<P>::T Raw;
    Raw ring[D][RS];
    i32x8 rp[D / 2][RS];  // pairs: the ring as 8-register operands (slot s in half s % 2 of rp[s / 2])
    i32x8 qa8[QS];        // pairs: query fragments of chunks c - 1 and c
    uint32_t cur[RS], nxt[RS];
    // Tiled layout: chunk c of a 16-row tile lies 16 * CB bytes after chunk
    // c - 1, and lane n's row is the tile's row n (row offsets from the tile
    // holding the chunk's first row).
    const int64_t cstep = a.tiled ? 16 * CB : CB;
    const uint8_t* xbase = a.tiled ? X + (r_begin >> 4) * 16 * a.xstride : X + r_begin * a.xstride;
    // KD > 0 (tiled rows only): rows in 32-bit arithmetic relative to the
    // chunk's first row; u = (r_begin % 16) + row counts rows from the tile
    // holding r_begin
    const uint32_t nrows32 = (uint32_t)(r_end - r_begin);
    const uint32_t rb15 = (uint32_t)(r_begin & 15);
    const uint32_t t16 = (uint32_t)(16 * a.xstride);
    auto tile_offs = [&](int64_t t, uint32_t (&o)[RS]) {
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            if constexpr (KD > 0) {
                uint32_t row = (uint32_t)t * TR + 16 * r + n;
                row = row < nrows32 ? row : nrows32 - 1;
                const uint32_t u = rb15 + row;
                o[r] = (u >> 4) * t16 + (u & 15) * CB + g * lane_bytes<P>();
                continue;
            }
            int64_t row = t * TR + 16 * r + n;
            if (row >= r_end - r_begin) row = r_end - r_begin - 1;
            if (a.tiled) {
                const int64_t R = r_begin + row;
                o[r] = (uint32_t)(((R >> 4) - (r_begin >> 4)) * 16 * a.xstride + (R & 15) * CB) +
                       (uint32_t)(g * lane_bytes<P>());
            } else {
                o[r] = (uint32_t)(row * a.xstride) + (uint32_t)(g * lane_bytes<P>());
            }
        }
    };
    auto qfrag_off = [&](int c) {
        return P::kElemBytes == 0 ? (c * P::kChunk + lane_dim) : (c * P::kChunk + lane_dim) * P::kElemBytes;
    };
    // row norms, loaded a tile ahead: the first MFMAs of a tile take them
    // (prefilter 3) without waiting for that tile's chunk prefetches.  The
    // first tile's are requested before its chunks, as every later tile's
    // are (a tile ahead): the wait for them at a tile's start then leaves
    // the chunk loads in flight on every pass of the loop, the first one
    // included (the compiler's wait covers all passes alike).
    float xnn[RS];
    auto load_xn = [&](int64_t t) {
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            if constexpr (KD > 0) {
                const uint32_t row = (uint32_t)t * TR + 16 * r + n;
                xnn[r] = (a.xn + r_begin)[row < nrows32 ? row : nrows32 - 1];
            } else {
                const int64_t row = r_begin + t * TR + 16 * r + n;
                xnn[r] = a.xn[row < r_end ? row : r_end - 1];
            }
        }
    };
    load_xn(stream);
    tile_offs(stream, cur);
    // issued in slot order (the scheduler may not reorder them: the loop's
    // waits assume slot s was requested before slot s + 1)
#pragma unroll
    for (int s2 = 0; s2 < D; ++s2) {
        const uint8_t* xb = xbase + (int64_t)(s2 < nchunk ? s2 : nchunk - 1) * cstep;
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            if constexpr (kPair)
                set_half(rp[s2 >> 1][r], s2 & 1, load_raw<P>(xb + cur[r]));
            else
                ring[s2][r] = load_raw<P>(xb + cur[r]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    i32x4 qa[2][QS];
    if constexpr (kQPre) {
#pragma unroll
        for (int q = 0; q < QS; ++q)
            qa[0][q] = *reinterpret_cast<const i32x4*>(smem + (qbase + 16 * q + n) * ldq + qfrag_off(0));
    }


    for (int64_t t = stream; t < ntiles; t += NS) {
        const int64_t r0 = r_begin + t * TR;
        if constexpr (KD == 0) tile_offs(t, cur);  // KD > 0: the previous tile's nxt
        tile_offs(t + NS < ntiles ? t + NS : t, nxt);
        bool rok[RS];
        float xn[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            rok[r] = KD > 0 ? (uint32_t)t * TR + 16 * r + n < nrows32 : r0 + 16 * r + n < r_end;
            xn[r] = xnn[r];
        }
        load_xn(t + NS < ntiles ? t + NS : t);
        int dofs[RS];
        typename P::Acc cinit[RS];
        if constexpr (kIntDot) {
#pragma unroll
            for (int r = 0; r < RS; ++r) {
                dofs[r] = wide_row_offset(a, xn[r], kap);
                cinit[r] = (i32x4){-dofs[r], -dofs[r], -dofs[r], -dofs[r]};
            }
        }
        typename P::Acc acc[QS][RS];
        // One ring turn (D chunks).  The first turn of a tile starts the
        // accumulators from a zero C operand instead of clearing them.
        auto turn = [&](int c0, auto first) {
            constexpr bool kFirst = decltype(first)::value;
            // all slots of a turn prefetch from the same tile
            const bool same = c0 + D < npad;
#pragma unroll
            for (int s2 = 0; s2 < D; ++s2) {
                const int c = c0 + s2;
                if constexpr (kQPre) {
                    const int cq = c + 1 < nchunk ? c + 1 : 0;
#pragma unroll
                    for (int q = 0; q < QS; ++q)
                        qa[(s2 + 1) & 1][q] =
                            *reinterpret_cast<const i32x4*>(smem + (qbase + 16 * q + n) * ldq + qfrag_off(cq));
                }
                if (c < nchunk) {
                    if constexpr (kPair) {
#pragma unroll
                        for (int q = 0; q < QS; ++q)
                            set_half(qa8[q], s2 & 1,
                                     *reinterpret_cast<const i32x4*>(smem + (qbase + 16 * q + n) * ldq + qfrag_off(c)));
                    } else if constexpr (!kQPre) {
#pragma unroll
                        for (int q = 0; q < QS; ++q)
                            qa[s2 & 1][q] =
                                *reinterpret_cast<const i32x4*>(smem + (qbase + 16 * q + n) * ldq + qfrag_off(c));
                    }
                    if constexpr (kPair) {
                        // chunks (c - 1, c) in one 128-deep MFMA at odd slots
                        // (D is even and turns start at multiples of D, so a
                        // pair never straddles a turn; nchunk is even)
                        if (s2 & 1) {
#pragma unroll
                            for (int q = 0; q < QS; ++q)
#pragma unroll
                                for (int r = 0; r < RS; ++r) {
                                    if (kFirst && s2 == 1) acc[q][r] = (typename P::Acc){0.f, 0.f, 0.f, 0.f};
                                    P::mma8(acc[q][r], qa8[q], rp[s2 >> 1][r]);
                                }
                        }
                    } else {
                    i32x4 bb[RS];
#pragma unroll
                    for (int r = 0; r < RS; ++r) bb[r] = to_frag<P>(ring[s2][r]);
                    // the other wave of the SIMD yields while these MFMAs
                    // issue (10M x 384 SQ8, 1024 queries: 1760 -> 1846 TOP/s)
                    if constexpr (KD > 0) __builtin_amdgcn_s_setprio(1);
#pragma unroll
                    for (int q = 0; q < QS; ++q)
#pragma unroll
                        for (int r = 0; r < RS; ++r) {
                            if (kFirst && s2 == 0) {
                                if constexpr (kIntDot) {
                                    acc[q][r] = cinit[r];
                                    P::mma(acc[q][r], qa[s2 & 1][q], bb[r]);
                                } else {
                                    P::mma0(acc[q][r], qa[s2 & 1][q], bb[r]);
                                }
                            } else
                                P::mma(acc[q][r], qa[s2 & 1][q], bb[r]);
                        }
                    if constexpr (KD > 0) __builtin_amdgcn_s_setprio(0);
                    }
                    // the slot's query fragments are read together, ahead of
                    // its MFMAs: left to itself the scheduler reuses one
                    // register set and waits out each read's latency before
                    // every second MFMA (10M x 384 SQ8, 4096 queries: 1804 ->
                    // 1943 TOP/s; with the dim fixed at compile time, 1962)
                    if constexpr (!kQPre) {
                        __builtin_amdgcn_sched_group_barrier(0x100, QS, 0);
                        if (!kPair || (s2 & 1)) __builtin_amdgcn_sched_group_barrier(0x008, QS * RS, 0);
                    }
                }
                // refill: a pair's first slot waits for the pair's MFMAs
#pragma unroll
                for (int s3 = kPair ? (s2 & 1 ? s2 - 1 : D) : s2; s3 <= s2; ++s3) {
                    const int cn = c0 + s3 + D;
                    int cc = same ? cn : cn - npad;
                    cc = cc < nchunk ? cc : nchunk - 1;
                    const uint8_t* xb = xbase + (int64_t)cc * cstep;
#pragma unroll
                    for (int r = 0; r < RS; ++r) {
                        if constexpr (kPair)
                            set_half(rp[s3 >> 1][r], s3 & 1, load_raw<P>(xb + (same ? cur[r] : nxt[r])));
                        else
                            ring[s3][r] = load_raw<P>(xb + (same ? cur[r] : nxt[r]));
                    }
                }
            }
        };
        turn(0, std::true_type{});
        for (int c0 = D; c0 < npad; c0 += D) turn(c0, std::false_type{});

        // epilogue.  Prefilter per (query, lane): a pair can pass
        // dist = xn + qv - 2*qs*dot <= th only if qs*dot - xn/2 reaches the
        // query's bound (half of the margin-widened qv - th).  Over the lane's
        // RS rows, qs*max(dot) - min(xn)/2 bounds every row's value from above
        // (qs > 0, fma rounding is monotone), so testing it per (query, lane)
        // is still necessary for every row; when all row norms are nearly
        // equal, max(dot) against one bound per query instead.  Only lanes
        // and query groups that may pass evaluate the exact test and append,
        // and a tile where no lane of the wave may pass skips that part.
        // Prefilters 2 and 3 compare max(dot) (max(dot - D_r) under 3) with
        // one bound per query.
        uint32_t pend = 0;  // bit q: query group q may pass in this lane
        if (a.prefilter >= 2) {
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                const int m0 = qbase + 16 * q + 4 * g;
                El bg4[4];
                if constexpr (kIntDot) {
                    const i32x4 v = *reinterpret_cast<const i32x4*>(bndg_s + m0);
#pragma unroll
                    for (int i = 0; i < 4; ++i) bg4[i] = v[i];
                } else {
                    const f32x4 v = *reinterpret_cast<const f32x4*>(bndg_s + m0);
#pragma unroll
                    for (int i = 0; i < 4; ++i) bg4[i] = v[i];
                }
                bool maybe = false;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    El mx = acc[q][0][i];
#pragma unroll
                    for (int r = 1; r < RS; ++r) mx = acc[q][r][i] > mx ? acc[q][r][i] : mx;
                    maybe |= mx >= bg4[i];
                }
                pend |= maybe ? 1u << q : 0u;
            }
        } else {
            float hx = xn[0];
#pragma unroll
            for (int r = 1; r < RS; ++r) hx = fminf(hx, xn[r]);
            hx *= 0.5f;
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                const int m0 = qbase + 16 * q + 4 * g;
                const f32x4 qs4 = *reinterpret_cast<const f32x4*>(qsc_s + m0);
                const f32x4 bn4 = *reinterpret_cast<const f32x4*>(bnd_s + m0);
                bool maybe = false;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    El mx = acc[q][0][i];
#pragma unroll
                    for (int r = 1; r < RS; ++r) mx = acc[q][r][i] > mx ? acc[q][r][i] : mx;
                    maybe |= fmaf(qs4[i], (float)mx, -hx) >= bn4[i];
                }
                pend |= maybe ? 1u << q : 0u;
            }
        }
        if (__ballot(pend != 0u) != 0) {
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                if (!((pend >> q) & 1u)) continue;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int m = qbase + 16 * q + 4 * g + i;
                    // skip a query of the group that no lane's one-bound
                    // prefilter passed (the bound is necessary for every row;
                    // 10M x 384 SQ8, 4096 queries: 2080 -> 2175 TOP/s)
                    if (a.prefilter >= 2) {
                        El mx = acc[q][0][i];
#pragma unroll
                        for (int r = 1; r < RS; ++r) mx = acc[q][r][i] > mx ? acc[q][r][i] : mx;
                        if (__ballot(mx >= bndg_s[m]) == 0) continue;
                    }
                    const float th = thr_s[m], qv = qn_s[m], sc = 2.0f * qsc_s[m];
#pragma unroll
                    for (int r = 0; r < RS; ++r) {
                        float dot;
                        if constexpr (kIntDot)
                            dot = (float)(acc[q][r][i] + dofs[r]);
                        else
                            dot = (float)acc[q][r][i];
                        const float dist = xn[r] + qv - sc * dot;
                        if (rok[r] && dist <= th) {
                            const uint32_t roff = (uint32_t)(r0 - r_begin + 16 * r + n);
                            wide_append(a, wq, q0 + m, m, roff, (uint32_t)(r_begin + roff), dist);
                        }
                    }
                }
            }
        }
        // keep room for a tile's typical appends (half of a short queue)
        if (__builtin_amdgcn_readfirstlane(*wq.count) + (qcap >= 1024 ? 512u : (uint32_t)qcap / 2) > (uint32_t)qcap)
            wide_flush(a, wq, q0, r_begin, lane);
        if constexpr (KD > 0) {
#pragma unroll
            for (int r = 0; r < RS; ++r) cur[r] = nxt[r];
        }
    }
    wide_flush(a, wq, q0, r_begin, lane);
}

template <class P, int QS, int RS, int NW, int QH, bool QPRE, int KD = 0>
hipError_t launch_wide_t(const ScanArgs& a, hipStream_t st) {
    static const char* const name = [] {
        static char buf[80];
        snprintf(buf, sizeof(buf), "scan_wide_kernel<%s,QS=%d,RS=%d,NW=%d,QH=%d%s>", pol_name<P>(), QS, RS, NW, QH,
                 QPRE ? ",QP" : "");
        return buf;
    }();
    t_variant = name;
    const int ncu = device_cu_count();
    constexpr int QT = 16 * QS * QH;
    const size_t lds = (size_t)QT * qrow_bytes<P>(a.kdim) + 5 * QT * sizeof(float) + NW * sizeof(uint32_t);
    static const bool attr = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&scan_wide_kernel<P, QS, RS, NW, QH, QPRE, KD>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        return true;
    }();
    (void)attr;
    const int qtiles = (a.B + QT - 1) / QT;
    const int64_t rows = a.row_end - a.row_begin;
    // row chunk: a few workgroups per CU over the launch, one wave step
    // (16 * RS rows per wave) to 32K rows, whole wave steps
    constexpr int64_t step = 16 * RS * (NW / QH);
    static const int64_t chunk_env = [] {  // tuning: ANNMI_WIDE_CHUNK_MAX rows
        const char* e = getenv("ANNMI_WIDE_CHUNK_MAX");
        const long v = e ? atol(e) : 0;
        return v >= 2048 ? (int64_t)v : (int64_t)32768;
    }();
    const int64_t chunk_max = a.chunk_max >= 2048 ? a.chunk_max : chunk_env;
    // workgroups per CU over the launch: up to 8, but no more than the query
    // tiles, since each workgroup stages its tile and with few tiles the
    // staging is not shared by many workgroups of a row chunk (128 queries
    // over 10M x 384 SQ8: 1.166 -> 1.084 ms of scan a step with one per CU;
    // 512 queries best at 4, 4096 at 4-8)
    const int64_t wg_per_cu = qtiles < 8 ? qtiles : 8;
    // the rest of the LDS holds the waves' flush histograms and candidate
    // queues (8 bytes an entry); a short queue flushes at half full, none
    // sends every pair straight to the lists
    const size_t hist_bytes = (size_t)NW * kFlushSlots * sizeof(uint32_t);
    int qcap = (int)((160 * 1024 - lds - hist_bytes) / (NW * 8)) / 64 * 64;
    if (qcap > 4096) qcap = 4096;
    if (const char* e = getenv("ANNMI_WIDE_QCAP")) {  // lower cap: tuning and tests
        const int v = atoi(e);
        if (v >= 0 && v < qcap) qcap = v / 64 * 64;
    }
    if (qcap < 64) qcap = 0;
    const size_t lds_total = lds + (size_t)NW * qcap * 8 + hist_bytes;
    int64_t chunk = rows * qtiles / ((int64_t)ncu * wg_per_cu);
    chunk = chunk < step ? step : (chunk > chunk_max ? chunk_max : chunk);
    chunk = (chunk + step - 1) / step * step;
    const int64_t nrc = (rows + chunk - 1) / chunk;
    const int64_t per = (nrc + 7) / 8;
    if (KD > 0 && (a.kdim != KD || !a.tiled)) return hipErrorInvalidValue;
    ANNMI_LAUNCH((scan_wide_kernel<P, QS, RS, NW, QH, QPRE, KD>), dim3((unsigned)(8 * per * qtiles)), dim3(64 * NW),
                       lds_total, st, a, qtiles, chunk, nrc, qcap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Wide scan on 32 x 32 MFMA tiles (SQ8: v_mfma_i32_32x32x32_i8, BF16:
// v_mfma_f32_32x32x16_bf16), an autotuner candidate.  One 32 x 32 MFMA does
// twice the work of a 16 x 16 one from the same operand bytes, so a query
// fragment read from LDS feeds twice the MFMA work.  (With register
// operands and eight independent accumulators a wave, the 16 x 16 forms
// issue at the full rate on MI355X: int8 16x16x64 4.46 POP/s, bf16 16x16x32
// 2.41 PFLOP/s; an earlier probe with fewer accumulators had shown 2.8 POP/s
// and misread it as a ceiling.)  This kernel measured no faster than the
// 16 x 16 one: bound by its LDS and load stream (SQ8 10M x 384 at 4096
// queries: 1.8 POP/s with everything, 2.7 without the epilogue and the row
// loads), so it is not the default.  Same workgroup plan, query staging,
// prefilters and queues; per wave step a 32 * RS row tile against the
// 32 * QS query tile.
//
// Operands: lane l holds index l % 32 (query of A, row of B) over bytes
// [16 * (l / 32), +16) of a 32-byte K step (the same K order for A and B,
// which is all the dot needs).  Accumulator element e of lane l is row
// (query) 8 * (e / 4) + 4 * (l / 32) + e % 4, column (row) l % 32.
//
// Rows in the tiled layout (16-row tiles of 64-byte chunks): K step k of
// row R lies at tile(R) + (k / 2) * 1024 + (R % 16) * 64 + (k % 2) * 32.
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct W32SQ8 {
    static constexpr int kElemBytes = 1;
    typedef int El;
    typedef i32x16 Acc;
    __device__ static inline void mma(Acc& acc, const i32x4& a, const i32x4& b) {
        // codes are stored as c ^ 0x80, the int8 c - 128 (the query side adds
        // 128 * sum(w) back)
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc, 0, 0, 0);
    }
    static const char* name() { return "sq8"; }
};
struct W32BF16 {
    static constexpr int kElemBytes = 2;
    typedef float El;
    typedef f32x16 Acc;
    __device__ static inline void mma(Acc& acc, const i32x4& a, const i32x4& b) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                      0, 0, 0);
    }
    static const char* name() { return "bf16"; }
};

template <class W, int QS, int RS, int kWaves, int D, bool kQPre>
__global__ __launch_bounds__(64 * kWaves) void scan_wide32_kernel(ScanArgs a, int qtiles, int64_t chunk_rows,
                                                                   int64_t nrc, int qcap) {
    extern __shared__ __align__(16) uint8_t smem[];
    typedef typename W::El El;
    typedef typename W::Acc Acc;
    constexpr bool kIntDot = std::is_same<El, int>::value;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = lane >> 5;  // K half of the operands; row quad of the outputs
    const int j = lane & 31;  // operand index; output column
    // work item of this workgroup: XCD x owns row chunks [x*per, (x+1)*per)
    const int xcd = blockIdx.x & 7;
    const int64_t s = blockIdx.x >> 3;
    const int64_t per = (nrc + 7) / 8;
    if (s / qtiles >= per) return;
    const int64_t rc = xcd * per + s / qtiles;
    if (rc >= nrc) return;
    const int qtile = (int)(s % qtiles);
    const int64_t r_begin = a.row_begin + rc * chunk_rows;
    const int64_t r_end = r_begin + chunk_rows < a.row_end ? r_begin + chunk_rows : a.row_end;

    constexpr int QT = 32 * QS;
    static_assert(QT <= 128, "queue entries hold the query slot in 7 bits");
    const int q0 = qtile * QT;
    const int qbytes = a.kdim * W::kElemBytes;
    const int ldq = qbytes + 16;
    const WideLds L = wide_stage<El>(a, smem, QT, ldq, qbytes, q0);
    const El* bndg_s = static_cast<const El*>(L.bndg);
    WideQueue wq;
    wq.row = L.cnt + kWaves + (int64_t)wave * 2 * qcap;
    wq.key = reinterpret_cast<float*>(wq.row + qcap);
    wq.count = L.cnt + wave;
    wq.hist = L.cnt + kWaves + (int64_t)kWaves * 2 * qcap + wave * kFlushSlots;
    wq.cap = qcap;
    const float kap = kIntDot ? wide_kap(a, L, QT, lane) : 0.f;

    const uint8_t* X = static_cast<const uint8_t*>(a.X);
    // K steps of a row; D (K steps of row fragments in flight per wave)
    // divides it, so every ring slot of a turn multiplies
    const int nks = qbytes / 32;
    constexpr int TR = 32 * RS;  // rows per wave step
    const int npad = nks;
    const int64_t ntiles = (r_end - r_begin + TR - 1) / TR;
    if (wave >= ntiles) return;
    const bool tiled = a.tiled != 0;
    const uint8_t* xbase = tiled ? X + (r_begin >> 4) * 16 * a.xstride : X + r_begin * a.xstride;
    // uniform byte offset of K step k from a row's offset
    auto koff = [&](int k) -> int64_t { return tiled ? (int64_t)(k >> 1) * 1024 + (k & 1) * 32 : (int64_t)k * 32; };
    // per-lane row offsets of wave step t (rows past the range repeat the last)
    auto tile_offs = [&](int64_t t, uint32_t (&o)[RS]) {
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            int64_t row = t * TR + 32 * r + j;
            if (row >= r_end - r_begin) row = r_end - r_begin - 1;
            if (tiled) {
                const int64_t R = r_begin + row;
                o[r] = (uint32_t)(((R >> 4) - (r_begin >> 4)) * 16 * a.xstride + (R & 15) * 64) + (uint32_t)(16 * h);
            } else {
                o[r] = (uint32_t)(row * a.xstride) + (uint32_t)(16 * h);
            }
        }
    };
    const uint8_t* qlane = smem + j * ldq + 16 * h;  // this lane's A bytes of query tile 0, K step 0

    i32x4 ring[D][RS];
    uint32_t cur[RS], nxt[RS];
    tile_offs(wave, cur);
    // issued in slot order: the loop's waits assume slot s was requested
    // before slot s + 1
#pragma unroll
    for (int s2 = 0; s2 < D; ++s2) {
        const uint8_t* xb = xbase + koff(s2);
#pragma unroll
        for (int r = 0; r < RS; ++r) ring[s2][r] = *reinterpret_cast<const i32x4*>(xb + cur[r]);
        __builtin_amdgcn_sched_barrier(0);
    }
    // row norms a wave step ahead
    float xnn[RS];
    auto load_xn = [&](int64_t t) {
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int64_t row = r_begin + t * TR + 32 * r + j;
            xnn[r] = a.xn[row < r_end ? row : r_end - 1];
        }
    };
    load_xn(wave);
    // kQPre: query fragments alternate between two register sets, K step c
    // multiplying qa[c & 1] while the next step's are read (D and the K step
    // count are even, so the parity carries across turns and wave steps)
    i32x4 qa[2][QS];
    if constexpr (kQPre) {
#pragma unroll
        for (int q = 0; q < QS; ++q) qa[0][q] = *reinterpret_cast<const i32x4*>(qlane + 32 * q * ldq);
    }

    for (int64_t t = wave; t < ntiles; t += kWaves) {
        const int64_t r0 = r_begin + t * TR;
        tile_offs(t, cur);
        tile_offs(t + kWaves < ntiles ? t + kWaves : t, nxt);
        bool rok[RS];
        float xn[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            rok[r] = r0 + 32 * r + j < r_end;
            xn[r] = xnn[r];
        }
        load_xn(t + kWaves < ntiles ? t + kWaves : t);
        int dofs[RS];
        Acc cinit[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            if constexpr (kIntDot) {
                dofs[r] = wide_row_offset(a, xn[r], kap);
#pragma unroll
                for (int e = 0; e < 16; ++e) cinit[r][e] = -dofs[r];
            } else {
                dofs[r] = 0;
#pragma unroll
                for (int e = 0; e < 16; ++e) cinit[r][e] = 0.f;
            }
        }
        Acc acc[QS][RS];
        // One ring turn (D K steps); the first turn of a wave step starts
        // the accumulators from cinit.
        auto turn = [&](int c0, auto first) {
            constexpr bool kFirst = decltype(first)::value;
            const bool same = c0 + D < npad;  // all slots prefetch from this wave step
#pragma unroll
            for (int s2 = 0; s2 < D; ++s2) {
                const int c = c0 + s2;
                if constexpr (kQPre) {
                    const int c1 = c + 1 < nks ? c + 1 : 0;
#pragma unroll
                    for (int q = 0; q < QS; ++q)
                        qa[(s2 + 1) & 1][q] = *reinterpret_cast<const i32x4*>(qlane + 32 * q * ldq + 32 * c1);
                } else {
#pragma unroll
                    for (int q = 0; q < QS; ++q)
                        qa[s2 & 1][q] = *reinterpret_cast<const i32x4*>(qlane + 32 * q * ldq + 32 * c);
                }
#pragma unroll
                for (int q = 0; q < QS; ++q)
#pragma unroll
                    for (int r = 0; r < RS; ++r) {
                        if (kFirst && s2 == 0) acc[q][r] = cinit[r];
                        W::mma(acc[q][r], qa[s2 & 1][q], ring[s2][r]);
                    }
                const int cn = c + D;
                const int cc = same ? cn : cn - npad;
                const uint8_t* xb = xbase + koff(cc);
#pragma unroll
                for (int r = 0; r < RS; ++r) ring[s2][r] = *reinterpret_cast<const i32x4*>(xb + (same ? cur[r] : nxt[r]));
            }
        };
        turn(0, std::true_type{});
        for (int c0 = D; c0 < npad; c0 += D) turn(c0, std::false_type{});

        // epilogue: prefilter per (query group q, lane), as in the 16 x 16
        // kernel (modes 2 and 3: max over the lane's rows of dot (- D_r)
        // against one bound per query; mode 1: qs * max(dot) - min(xn) / 2
        // against the per-query bound), then the exact test where it may pass
        uint32_t pend = 0;
        if (a.prefilter >= 2) {
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                bool maybe = false;
#pragma unroll
                for (int e4 = 0; e4 < 4; ++e4) {
                    const int m0 = 32 * q + 8 * e4 + 4 * h;
                    El bg4[4];
                    if constexpr (kIntDot) {
                        const i32x4 v = *reinterpret_cast<const i32x4*>(bndg_s + m0);
#pragma unroll
                        for (int i = 0; i < 4; ++i) bg4[i] = v[i];
                    } else {
                        const f32x4 v = *reinterpret_cast<const f32x4*>(bndg_s + m0);
#pragma unroll
                        for (int i = 0; i < 4; ++i) bg4[i] = v[i];
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        El mx = acc[q][0][4 * e4 + i];
#pragma unroll
                        for (int r = 1; r < RS; ++r) mx = acc[q][r][4 * e4 + i] > mx ? acc[q][r][4 * e4 + i] : mx;
                        maybe |= mx >= bg4[i];
                    }
                }
                pend |= maybe ? 1u << q : 0u;
            }
        } else {
            float hx = xn[0];
#pragma unroll
            for (int r = 1; r < RS; ++r) hx = fminf(hx, xn[r]);
            hx *= 0.5f;
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                bool maybe = false;
#pragma unroll
                for (int e4 = 0; e4 < 4; ++e4) {
                    const int m0 = 32 * q + 8 * e4 + 4 * h;
                    const f32x4 qs4 = *reinterpret_cast<const f32x4*>(L.qsc + m0);
                    const f32x4 bn4 = *reinterpret_cast<const f32x4*>(L.bnd + m0);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        El mx = acc[q][0][4 * e4 + i];
#pragma unroll
                        for (int r = 1; r < RS; ++r) mx = acc[q][r][4 * e4 + i] > mx ? acc[q][r][4 * e4 + i] : mx;
                        maybe |= fmaf(qs4[i], (float)mx, -hx) >= bn4[i];
                    }
                }
                pend |= maybe ? 1u << q : 0u;
            }
        }
        if (__ballot(pend != 0u) != 0) {
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                if (!((pend >> q) & 1u)) continue;
#pragma unroll
                for (int e4 = 0; e4 < 4; ++e4) {
                    const int m0 = 32 * q + 8 * e4 + 4 * h;
                    const f32x4 th4 = *reinterpret_cast<const f32x4*>(L.thr + m0);
                    const f32x4 qv4 = *reinterpret_cast<const f32x4*>(L.qn + m0);
                    const f32x4 qs4 = *reinterpret_cast<const f32x4*>(L.qsc + m0);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float sc = 2.0f * qs4[i];
#pragma unroll
                        for (int r = 0; r < RS; ++r) {
                            float dot;
                            if constexpr (kIntDot)
                                dot = (float)(acc[q][r][4 * e4 + i] + dofs[r]);
                            else
                                dot = (float)acc[q][r][4 * e4 + i];
                            const float dist = xn[r] + qv4[i] - sc * dot;
                            if (rok[r] && dist <= th4[i]) {
                                const uint32_t roff = (uint32_t)(r0 - r_begin + 32 * r + j);
                                wide_append(a, wq, q0 + m0 + i, m0 + i, roff, (uint32_t)(r_begin + roff), dist);
                            }
                        }
                    }
                }
            }
        }
        // keep room for a wave step's typical appends
        if (__builtin_amdgcn_readfirstlane(*wq.count) + 512u > (uint32_t)qcap) wide_flush(a, wq, q0, r_begin, lane);
    }
    wide_flush(a, wq, q0, r_begin, lane);
}

template <class W, int QS, int RS, int NW, int D, bool QP>
hipError_t launch_wide32_t(const ScanArgs& a, hipStream_t st) {
    static const char* const name = [] {
        static char buf[80];
        snprintf(buf, sizeof(buf), "scan_wide32_kernel<%s,QS=%d,RS=%d,NW=%d,D=%d%s>", W::name(), QS, RS, NW, D,
                 QP ? ",QP" : "");
        return buf;
    }();
    t_variant = name;
    const int ncu = device_cu_count();
    constexpr int QT = 32 * QS;
    const int qbytes = a.kdim * W::kElemBytes;
    if (qbytes % (32 * D) != 0) return hipErrorInvalidValue;  // whole ring turns of 32-byte K steps
    const size_t lds = (size_t)QT * (qbytes + 16) + 5 * QT * sizeof(float) + NW * sizeof(uint32_t);
    if (lds + 8 * 1024 > 160 * 1024) return hipErrorInvalidValue;
    static const bool attr = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&scan_wide32_kernel<W, QS, RS, NW, D, QP>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        return true;
    }();
    (void)attr;
    const int qtiles = (a.B + QT - 1) / QT;
    const int64_t rows = a.row_end - a.row_begin;
    // row chunk as for the 16 x 16 kernel: up to 8 workgroups per CU over
    // the launch (no more than the query tiles), whole wave steps, at most
    // 32K rows (or the tuned maximum)
    constexpr int64_t step = 32 * RS * NW;
    static const int64_t chunk_env = [] {
        const char* e = getenv("ANNMI_WIDE_CHUNK_MAX");
        const long v = e ? atol(e) : 0;
        return v >= 2048 ? (int64_t)v : (int64_t)32768;
    }();
    const int64_t chunk_max = a.chunk_max >= 2048 ? a.chunk_max : chunk_env;
    // workgroups per CU over the launch: up to 8, but no more than the query
    // tiles, since each workgroup stages its tile and with few tiles the
    // staging is not shared by many workgroups of a row chunk (128 queries
    // over 10M x 384 SQ8: 1.166 -> 1.084 ms of scan a step with one per CU;
    // 512 queries best at 4, 4096 at 4-8)
    const int64_t wg_per_cu = qtiles < 8 ? qtiles : 8;
    int64_t chunk = rows * qtiles / ((int64_t)ncu * wg_per_cu);
    chunk = chunk < step ? step : (chunk > chunk_max ? chunk_max : chunk);
    chunk = (chunk + step - 1) / step * step;
    const int64_t nrc = (rows + chunk - 1) / chunk;
    const int64_t per = (nrc + 7) / 8;
    // the rest of the LDS holds the waves' flush histograms and candidate
    // queues (8 bytes an entry)
    const size_t hist_bytes = (size_t)NW * kFlushSlots * sizeof(uint32_t);
    int qcap = (int)((160 * 1024 - lds - hist_bytes) / (NW * 8)) / 64 * 64;
    if (qcap > 4096) qcap = 4096;
    if (const char* e = getenv("ANNMI_WIDE_QCAP")) {  // lower cap: tuning and tests
        const int v = atoi(e);
        if (v >= 0 && v < qcap) qcap = v / 64 * 64;
    }
    if (qcap < 1024) qcap = 0;
    ANNMI_LAUNCH((scan_wide32_kernel<W, QS, RS, NW, D, QP>), dim3((unsigned)(8 * per * qtiles)), dim3(64 * NW),
                       lds + (size_t)NW * qcap * 8 + hist_bytes, st, a, qtiles, chunk, nrc, qcap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SQ8 wide scan on 32 x 32 x 32 int8 MFMA tiles with a lean epilogue (an
// autotuner shape, kWideCfg32R4W4 / kWideCfg32R2W4 for SQ8).
//
// A wave tile is 128 queries x 32 * RS rows: each query fragment read from
// LDS feeds RS MFMAs and each row fragment four.  The epilogue keeps the
// register footprint small: the integer-dot prefilter (the accumulators
// start at -D_r, so max(dot - D_r) against one bound per query is necessary
// for a pass) marks query tiles that may pass in some lane; only those tiles
// are written, one 32 x 32 tile at a time, to a 4 KB LDS slot of the wave and
// tested exactly in a loop that is not unrolled, so the rare appends do not
// multiply the kernel's code and live ranges (the other 32 x 32 kernel,
// with its epilogue unrolled over every accumulator, spills 30-200 VGPRs
// in every SQ8 shape; this one spills none at RS <= 2).  The exact test is
// the one of the other wide kernels (dist = xn + qv - 2 qs (dot), the same
// integer dot), so every shape appends the same candidates.  Prefilter mode
// 3 only (the default for integer codes); the tiled row layout only.
//
// Measured (MI355X, 10M x 384, 4096 queries, largest launch of a step):
// 1913 TOP/s at RS = 2 on 8 waves against 1877 for the default 16 x 16
// kernel, with the MFMA pipe busy for the same cycles (SQ_VALU_MFMA_BUSY_CYCLES
// 8.9e9 for both: the 32 x 32 form halves the MFMA instructions, not the
// MFMA time) and no LDS bank conflicts (1.1e9 conflict cycles in the
// default kernel); its smaller launches are slower, so a whole step is not
// (20.2 vs 18.0 ms of scan).  One wave a SIMD (RS = 2 or 4 on 4 waves, the
// accumulators in AGPRs) measured 1470 / 855 TOP/s: a single wave does not
// hide the row-load latency with a ring of 4 K steps.
// LDS the 32 x 32 lean-epilogue kernel needs at the least: the 128-query
// tile, five words a query, the queue counters and flush histograms, a 4 KB
// exact-test slot a wave and a queue of 1024 entries a wave
constexpr size_t i8x32_lds_need(int qbytes, int nw) {
    return (size_t)128 * (qbytes + 16) + 5 * 128 * sizeof(float) + nw * sizeof(uint32_t) +
           (size_t)nw * kFlushSlots * sizeof(uint32_t) + (size_t)nw * 4096 + (size_t)nw * 1024 * 8;
}

template <int RS, int D, int NW>
__global__ __launch_bounds__(64 * NW) void scan_i8x32_kernel(ScanArgs a, int qtiles, int64_t chunk_rows, int64_t nrc,
                                                              int qcap) {
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr int QS = 4, QT = 128;
    typedef W32SQ8 W;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = lane >> 5;  // K half of the operands; row quad of the outputs
    const int j = lane & 31;  // operand index; output column (row of the tile)
    const int xcd = blockIdx.x & 7;
    const int64_t s = blockIdx.x >> 3;
    const int64_t per = (nrc + 7) / 8;
    if (s / qtiles >= per) return;
    const int64_t rc = xcd * per + s / qtiles;
    if (rc >= nrc) return;
    const int qtile = (int)(s % qtiles);
    const int64_t r_begin = a.row_begin + rc * chunk_rows;
    const int64_t r_end = r_begin + chunk_rows < a.row_end ? r_begin + chunk_rows : a.row_end;
    const int q0 = qtile * QT;
    const int qbytes = a.kdim;
    const int ldq = qbytes + 16;
    const WideLds L = wide_stage<int>(a, smem, QT, ldq, qbytes, q0);
    const int* bndg_s = static_cast<const int*>(L.bndg);
    WideQueue wq;
    wq.row = L.cnt + NW + (int64_t)wave * 2 * qcap;
    wq.key = reinterpret_cast<float*>(wq.row + qcap);
    wq.count = L.cnt + wave;
    wq.hist = L.cnt + NW + (int64_t)NW * 2 * qcap + wave * kFlushSlots;
    wq.cap = qcap;
    int* stage = reinterpret_cast<int*>(L.cnt + NW + (int64_t)NW * 2 * qcap + NW * kFlushSlots) + wave * 1024;
    const float kap = wide_kap(a, L, QT, lane);

    const uint8_t* X = static_cast<const uint8_t*>(a.X);
    const int nks = qbytes / 32;  // a multiple of D (launcher)
    constexpr int TR = 32 * RS;
    const int64_t ntiles = (r_end - r_begin + TR - 1) / TR;
    if (wave >= ntiles) return;
    const uint8_t* xbase = X + (r_begin >> 4) * 16 * a.xstride;
    // K step k of a row: 64-byte chunk k / 2 of its 16-row tile, half k % 2
    auto koff = [&](int k) -> int64_t { return (int64_t)(k >> 1) * 1024 + (k & 1) * 32; };
    auto tile_offs = [&](int64_t t, uint32_t (&o)[RS]) {
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            int64_t row = t * TR + 32 * r + j;
            if (row >= r_end - r_begin) row = r_end - r_begin - 1;
            const int64_t R = r_begin + row;
            o[r] = (uint32_t)(((R >> 4) - (r_begin >> 4)) * 16 * a.xstride + (R & 15) * 64) + (uint32_t)(16 * h);
        }
    };
    const uint8_t* qlane = smem + j * ldq + 16 * h;

    i32x4 ring[D][RS];
    uint32_t cur[RS], nxt[RS];
    tile_offs(wave, cur);
#pragma unroll
    for (int s2 = 0; s2 < D; ++s2) {
        const uint8_t* xb = xbase + koff(s2);
#pragma unroll
        for (int r = 0; r < RS; ++r) ring[s2][r] = *reinterpret_cast<const i32x4*>(xb + cur[r]);
        __builtin_amdgcn_sched_barrier(0);
    }
    float xnn[RS];
    auto load_xn = [&](int64_t t) {
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int64_t row = r_begin + t * TR + 32 * r + j;
            xnn[r] = a.xn[row < r_end ? row : r_end - 1];
        }
    };
    load_xn(wave);
    // query fragments alternate between two register sets: K step c
    // multiplies qa[c & 1] while the next step's are read (D and the K step
    // count are even, so the parity carries across turns and wave steps)
    i32x4 qa[2][QS];
#pragma unroll
    for (int q = 0; q < QS; ++q) qa[0][q] = *reinterpret_cast<const i32x4*>(qlane + 32 * q * ldq);

    for (int64_t t = wave; t < ntiles; t += NW) {
        const int64_t r0 = r_begin + t * TR;
        tile_offs(t, cur);
        tile_offs(t + NW < ntiles ? t + NW : t, nxt);
        float xn[RS];
        int dofs[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            xn[r] = xnn[r];
            dofs[r] = wide_row_offset(a, xn[r], kap);
        }
        load_xn(t + NW < ntiles ? t + NW : t);
        i32x16 acc[QS][RS];
        auto turn = [&](int c0, auto first) {
            constexpr bool kFirst = decltype(first)::value;
            const bool same = c0 + D < nks;  // all slots prefetch from this wave step
#pragma unroll
            for (int s2 = 0; s2 < D; ++s2) {
                const int c = c0 + s2;
                const int c1 = c + 1 < nks ? c + 1 : 0;
#pragma unroll
                for (int q = 0; q < QS; ++q)
                    qa[(s2 + 1) & 1][q] = *reinterpret_cast<const i32x4*>(qlane + 32 * q * ldq + 32 * c1);
#pragma unroll
                for (int q = 0; q < QS; ++q)
#pragma unroll
                    for (int r = 0; r < RS; ++r) {
                        if (kFirst && s2 == 0) {
#pragma unroll
                            for (int e = 0; e < 16; ++e) acc[q][r][e] = -dofs[r];
                        }
                        W::mma(acc[q][r], qa[s2 & 1][q], ring[s2][r]);
                    }
                const int cn = c + D;
                const int cc = same ? cn : cn - nks;
                const uint8_t* xb = xbase + koff(cc);
#pragma unroll
                for (int r = 0; r < RS; ++r)
                    ring[s2][r] = *reinterpret_cast<const i32x4*>(xb + (same ? cur[r] : nxt[r]));
            }
        };
        turn(0, std::true_type{});
        for (int c0 = D; c0 < nks; c0 += D) turn(c0, std::false_type{});

        // prefilter: query tile q may pass in this lane when some element's
        // max over the lane's rows of dot - D_r reaches its query's bound
        uint32_t pend = 0;
#pragma unroll
        for (int q = 0; q < QS; ++q) {
            bool maybe = false;
#pragma unroll
            for (int e4 = 0; e4 < 4; ++e4) {
                const i32x4 b = *reinterpret_cast<const i32x4*>(bndg_s + 32 * q + 8 * e4 + 4 * h);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    int mx = acc[q][0][4 * e4 + i];
#pragma unroll
                    for (int r = 1; r < RS; ++r) mx = acc[q][r][4 * e4 + i] > mx ? acc[q][r][4 * e4 + i] : mx;
                    maybe |= mx >= b[i];
                }
            }
            pend |= maybe ? 1u << q : 0u;
        }
        if (__ballot(pend != 0u) != 0) {
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                if (__ballot((pend >> q) & 1u) == 0) continue;
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    // this lane's 16 values of tile (q, r), read back below
                    // by the same lane only (LDS accesses of a lane stay in
                    // order)
#pragma unroll
                    for (int e4 = 0; e4 < 4; ++e4)
                        *reinterpret_cast<i32x4*>(stage + (e4 * 64 + lane) * 4) =
                            (i32x4){acc[q][r][4 * e4], acc[q][r][4 * e4 + 1], acc[q][r][4 * e4 + 2],
                                    acc[q][r][4 * e4 + 3]};
                    const int64_t row = r0 + 32 * r + j;
                    if (((pend >> q) & 1u) && row < r_end) {
                        const float xr = xn[r];
                        const int dr = dofs[r];
#pragma unroll 1
                        for (int e = 0; e < 16; ++e) {
                            const int m = 32 * q + 8 * (e >> 2) + 4 * h + (e & 3);
                            const float th = L.thr[m], qv = L.qn[m], sc = 2.0f * L.qsc[m];
                            const float dot = (float)(stage[((e >> 2) * 64 + lane) * 4 + (e & 3)] + dr);
                            const float dist = xr + qv - sc * dot;
                            if (dist <= th)
                                wide_append(a, wq, q0 + m, m, (uint32_t)(row - r_begin), (uint32_t)row, dist);
                        }
                    }
                }
            }
        }
        // keep room for a wave step's typical appends
        if (__builtin_amdgcn_readfirstlane(*wq.count) + 512u > (uint32_t)qcap) wide_flush(a, wq, q0, r_begin, lane);
    }
    wide_flush(a, wq, q0, r_begin, lane);
}

template <int RS, int D, int NW>
hipError_t launch_i8x32_t(const ScanArgs& a, hipStream_t st) {
    static const char* const name = [] {
        static char buf[80];
        snprintf(buf, sizeof(buf), "scan_i8x32_kernel<sq8,RS=%d,D=%d,NW=%d>", RS, D, NW);
        return buf;
    }();
    t_variant = name;
    constexpr int QT = 128;
    const int qbytes = a.kdim;
    if (!a.tiled || a.prefilter != 3 || qbytes % (32 * D) != 0) return hipErrorInvalidValue;
    const size_t lds = (size_t)QT * (qbytes + 16) + 5 * QT * sizeof(float) + NW * sizeof(uint32_t);
    const size_t hist_bytes = (size_t)NW * kFlushSlots * sizeof(uint32_t);
    const size_t stage_bytes = (size_t)NW * 4096;
    if (i8x32_lds_need(qbytes, NW) > 160 * 1024) return hipErrorInvalidValue;
    static const bool attr = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&scan_i8x32_kernel<RS, D, NW>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        return true;
    }();
    (void)attr;
    const int ncu = device_cu_count();
    const int qtiles = (a.B + QT - 1) / QT;
    const int64_t rows = a.row_end - a.row_begin;
    constexpr int64_t step = 32 * RS * NW;
    static const int64_t chunk_env = [] {
        const char* e = getenv("ANNMI_WIDE_CHUNK_MAX");
        const long v = e ? atol(e) : 0;
        return v >= 2048 ? (int64_t)v : (int64_t)32768;
    }();
    const int64_t chunk_max = a.chunk_max >= 2048 ? a.chunk_max : chunk_env;
    // a few workgroups a CU over the launch, no more than the query tiles
    const int64_t wg_per_cu = qtiles < 8 ? qtiles : 8;
    int64_t chunk = rows * qtiles / ((int64_t)ncu * wg_per_cu);
    chunk = chunk < step ? step : (chunk > chunk_max ? chunk_max : chunk);
    chunk = (chunk + step - 1) / step * step;
    const int64_t nrc = (rows + chunk - 1) / chunk;
    const int64_t per = (nrc + 7) / 8;
    int qcap = (int)((160 * 1024 - lds - hist_bytes - stage_bytes) / (NW * 8)) / 64 * 64;
    if (qcap > 4096) qcap = 4096;
    if (const char* e = getenv("ANNMI_WIDE_QCAP")) {  // lower cap: tuning and tests
        const int v = atoi(e);
        if (v >= 0 && v < qcap) qcap = v / 64 * 64;
    }
    if (qcap < 1024) qcap = 0;
    ANNMI_LAUNCH((scan_i8x32_kernel<RS, D, NW>), dim3((unsigned)(8 * per * qtiles)), dim3(64 * NW),
                 lds + (size_t)NW * qcap * 8 + hist_bytes + stage_bytes, st, a, qtiles, chunk, nrc, qcap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SQ8 wide scan with the rows staged in LDS and the queries held in
// registers (kWideCfgRL; 32 x 32 x 32 int8 MFMA, tiled rows, prefilter 3).
//
// The kernels above keep a row tile per wave in registers and read the
// query fragments from LDS for every MFMA: each 1 KB query fragment feeds
// two 16 x 16 MFMAs, so at the MFMA rate the LDS would have to deliver
// 128 B/clk per CU, and each wave streams its own rows from L2.  Here a
// workgroup of 8 waves owns 512 queries, 64 a wave, as register-resident A
// operands (2 fragments of 32 queries x NKS K steps), and the rows of its
// chunk pass through a ring of three 64-row stages in LDS, written by
// direct global-to-LDS loads (global_load_lds_dwordx4, no register round
// trip) two stages ahead and read by all 8 waves: each 1 KB row fragment read
// feeds two 32 x 32 MFMAs (64 B/clk per CU at the MFMA rate), and L2 serves
// each row once per 512 queries instead of once per 128.
//
// A stage is 4 consecutive 16-row tiles of the tiled layout, i.e. one
// contiguous run of 2 * NKS blocks of 1 KB (16 rows x one 64-byte chunk).
// Inside a block, 16-byte unit u of row r is stored at unit u ^ swz(r) of the
// row, so the 32 rows x 16 bytes a B-operand read takes from a block hit
// distinct banks.  Synchronisation per stage: each wave waits for its own
// loads of the stage (vmcnt), then one barrier publishes the stage and
// frees the buffer that every wave finished in the previous stage, which
// then receives the stage two ahead.
//
// The accumulators hold the plain integer dots.  Prefilter: for each of the
// lane's 32 (query, row pair) maxima of dot - D_r (D_r = floor((xn_r - xn_lo)
// * kap), kap <= 1 / (2 qs) over the wave's queries) against the query's
// integer bound, as in the other wide kernels; only the elements that pass
// it take the exact test, dist = xn + qv - 2 qs * dot <= thr, the same
// expression as the other kernels, so every shape appends the same
// candidates.  Appends go to the wave's LDS queue (query slot = query within
// the wave's 64) and are flushed per query.
//
// Measured (MI355X, SQ8, 10M rows, the largest launches of a step): d = 256,
// 4096 queries 1652 vs 1638 TOP/s for the built-in kernel, 1024 queries 1568
// vs 1558; the B-operand reads are free of bank conflicts (SQ_LDS_BANK_CONFLICT
// 2.5% of the built-in kernel's).  At d = 384 the 96 query-fragment registers
// and 64 accumulators leave too few of the 256 registers two waves a SIMD
// allow (the compiler spills, 1752 vs 1814 TOP/s), so this shape takes padded
// d = 128 and 256 only.  It is an autotuner candidate, not a default.
__device__ __forceinline__ int rl_swz(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }

template <int N>
__device__ __forceinline__ void rl_wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt field");
    // s_waitcnt: vmcnt N (bits 3:0 and 15:14), expcnt and lgkmcnt not waited
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

template <int NKS>
__global__ __launch_bounds__(512) void scan_rl_kernel(ScanArgs a, int qtiles, int64_t chunk_rows, int64_t nrc,
                                                      int qcap) {
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr int NW = 8, QF = 2, RS = 2, NST = 3, QT = 64 * NW, SR = 64;
    constexpr int SB = SR * 32 * NKS;  // stage bytes: 4 tiles x NKS / 2 blocks of 1 KB
    constexpr int NBW = NKS / 4;       // 1 KB blocks a wave loads per stage (2 * NKS over 8 waves)
    static_assert(NKS % 4 == 0, "whole blocks per wave");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // wave-uniform: scalar address math
    const int h = lane >> 5;  // K half of the operands; query quad of the outputs
    const int j = lane & 31;  // operand index (query of A, row of B); output column (row)
    const int xcd = blockIdx.x & 7;
    const int64_t s = blockIdx.x >> 3;
    const int64_t per = (nrc + 7) / 8;
    if (s / qtiles >= per) return;
    const int64_t rc = xcd * per + s / qtiles;
    if (rc >= nrc) return;
    const int qtile = (int)(s % qtiles);
    const int64_t r_begin = a.row_begin + rc * chunk_rows;
    const int64_t r_end = r_begin + chunk_rows < a.row_end ? r_begin + chunk_rows : a.row_end;
    const int q0 = qtile * QT;
    const int q0w = q0 + 64 * wave;  // this wave's queries

    // LDS: stages | row norms of the stages [NST][SR] | thr, qn, qsc [QT] |
    // integer bounds [QT] | queue counters [NW] (+pad) | flush histograms |
    // exact-test slots (4 KB a wave) | queues
    float* xn_s = reinterpret_cast<float*>(smem + NST * SB);
    float* thr_s = xn_s + NST * SR;
    float* qn_s = thr_s + QT;
    float* qsc_s = qn_s + QT;
    int* bg_s = reinterpret_cast<int*>(qsc_s + QT);
    uint32_t* cnt_s = reinterpret_cast<uint32_t*>(bg_s + QT);
    uint32_t* hist_s = cnt_s + 16;
    int* slot = reinterpret_cast<int*>(hist_s + NW * kFlushSlots) + wave * 1024;
    uint32_t* queues = reinterpret_cast<uint32_t*>(slot - wave * 1024 + NW * 1024);
    WideQueue wq;
    wq.row = queues + (int64_t)wave * 2 * qcap;
    wq.key = reinterpret_cast<float*>(wq.row + qcap);
    wq.count = cnt_s + wave;
    wq.hist = hist_s + wave * kFlushSlots;
    wq.cap = qcap;

    const int64_t nrows = r_end - r_begin;
    const int nst = (int)((nrows + SR - 1) / SR);
    const uint8_t* X = static_cast<const uint8_t*>(a.X);
    const int64_t tile_bytes = 16 * (int64_t)a.xstride;
    const int64_t t_first = r_begin >> 4;
    const int64_t t_last = (r_end - 1) >> 4;
    // direct loads of stage st into its ring buffer: block b = wave + 8 i
    // (tile b / (NKS/2) of the stage, chunk b % (NKS/2)); lane l writes unit l
    // of the block, i.e. row l / 4, unit (l % 4) ^ swz(row) of the source;
    // tiles past the chunk repeat its last tile (their rows are masked)
    const int dr = lane >> 2;
    const uint32_t loff = (uint32_t)(dr * 64 + 16 * ((lane & 3) ^ rl_swz(dr)));
    // The stage's row norms travel the same way (each wave loads 8 of the
    // 64 with one masked dword load), so the loop has no loads into
    // registers and every wave issues NBW + 1 loads per stage.
    auto issue = [&](int st) {
        const int sb = st % NST;
        uint8_t* dst = smem + sb * SB;
#pragma unroll
        for (int i = 0; i < NBW; ++i) {
            const int b = wave + 8 * i;
            int64_t T = t_first + 4 * (int64_t)st + b / (NKS / 2);
            T = T < t_last ? T : t_last;
            // the block's address is wave-uniform: kept in scalar registers
            // (read back through readfirstlane, so it is not folded with the
            // lane offset into per-lane pointers held across the loop)
            const uint64_t ub = (uint64_t)(X + T * tile_bytes + (b % (NKS / 2)) * 1024);
            const uint64_t us = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(ub >> 32)) << 32) |
                                (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)ub);
            __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(us + loff),
                                             (__attribute__((address_space(3))) void*)(dst + b * 1024), 16, 0, 0);
        }
        if ((lane >> 3) == wave) {
            // row r0 + lane, clamped to the chunk; the base is scalar
            const int64_t r0s = r_begin + (int64_t)st * SR;
            const int64_t lim = r_end - 1 - r0s;  // >= lane for all but the last stage
            const uint64_t ub = (uint64_t)(a.xn + r0s);
            const uint64_t us = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(ub >> 32)) << 32) |
                                (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)ub);
            const int64_t li = (int64_t)lane < lim ? (int64_t)lane : lim;
            __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(us + 4 * (uint64_t)li),
                                             (__attribute__((address_space(3))) void*)(xn_s + sb * SR), 4, 0, 0);
        }
    };
    // start the first two stages before the query set-up.  Every iteration
    // issues one stage (past the last stage: a repeat of the last tile into
    // the free buffer), so the loop's wait counts are the same every time.
    issue(0);
    issue(1);

    // per-query words of the workgroup's 512 queries (one a thread)
    {
        const int m = threadIdx.x;
        const int qi = q0 + m;
        const bool ok = qi < a.B;
        const float th = ok ? a.thr[qi] : -__builtin_inff();
        const float qv = ok ? a.qn[qi] : 0.f;
        const float qs = ok ? a.qscale[qi] : 0.f;
        thr_s[m] = th;
        qn_s[m] = qv;
        qsc_s[m] = qs;
        // integer bound of the prefilter: dot >= ((xn_lo + qv) - th - margin)
        // / (2 qs) is necessary for dist <= th (the wide_stage form)
        const float M = a.xn_amax + fabsf(qv) + fabsf(th);
        const float sc = 2.0f * qs;
        const float num = ((a.xn_lo + qv) - th) - M * 0x1p-14f;
        const float bg = sc > 0.f ? num / sc : (num <= 0.f ? -__builtin_inff() : __builtin_inff());
        bg_s[m] = bg != bg ? 0x7fffffff
                           : (bg <= -2147483648.0f ? (int)0x80000000
                                                   : (bg >= 2147483647.0f ? 0x7fffffff : (int)ceilf(bg)));
        if (m < NW) cnt_s[m] = 0u;
    }
    // query operands: lane (j, h) of fragment f holds query q0w + 32 f + j,
    // bytes [32 k + 16 h, +16) of K step k
    i32x4 qreg[QF][NKS];
#pragma unroll
    for (int f = 0; f < QF; ++f) {
        const int qi = q0w + 32 * f + j;
        const uint8_t* qp = static_cast<const uint8_t*>(a.Q) + (int64_t)(qi < a.B ? qi : 0) * a.kdim + 16 * h;
#pragma unroll
        for (int k = 0; k < NKS; ++k)
            qreg[f][k] = qi < a.B ? *reinterpret_cast<const i32x4*>(qp + 32 * k) : (i32x4){0, 0, 0, 0};
    }
    __syncthreads();
    // kap over the wave's queries (padding queries have threshold -inf)
    float kap;
    {
        float mq = thr_s[64 * wave + lane] != -__builtin_inff() ? qsc_s[64 * wave + lane] : 0.f;
#pragma unroll
        for (int o = 32; o; o >>= 1) mq = fmaxf(mq, __shfl_xor(mq, o));
        kap = (mq > 0.f && mq < __builtin_inff()) ? (0.5f / mq) * (1.0f - 0x1p-20f) : 0.f;
    }
    // bounds of the lane's outputs (read from LDS where used): element e of
    // fragment f is query 32 f + 8 (e / 4) + 4 h + e % 4 of the wave
    const int* bgl = bg_s + 64 * wave + 4 * h;
    // B-operand offsets in a stage: row fragment rf holds rows 32 rf + j,
    // i.e. tile 2 rf + j / 16, row j % 16; K step k reads chunk k / 2, unit
    // 2 (k % 2) + h (stored at that unit ^ swz(row))
    // Two lane offsets cover every read: even and odd K steps (unit h or
    // 2 + h, i.e. the even one ^ 32 after the swizzle); the row fragment
    // (NKS * 1024 bytes apart) and the chunk are immediate offsets.
    const int rr = j & 15;
    const uint32_t bl0 = (uint32_t)((j >> 4) * (NKS / 2) * 1024 + rr * 64 + 16 * (h ^ rl_swz(rr)));
    const uint32_t bl1 = bl0 ^ 32u;
    // every load so far (query operands, the first two stages) complete, so
    // the loop's own waits count only the loop's loads
    rl_wait_vm<0>();

    for (int st = 0; st < nst; ++st) {
        // own loads of stage st done (those of stage st + 1 may be in
        // flight), then the barrier: every wave's are, and every wave is
        // past stage st - 1, whose buffer takes stage st + 2
        rl_wait_vm<NBW + 1>();
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        const int64_t r0 = r_begin + (int64_t)st * SR;
        __builtin_amdgcn_sched_barrier(0);
        issue(st + 2);
        __builtin_amdgcn_sched_barrier(0);
        const uint8_t* buf0 = smem + (st % NST) * SB + bl0;
        const uint8_t* buf1 = smem + (st % NST) * SB + bl1;
        i32x16 acc[QF][RS];
        // row fragments one K step ahead of their MFMAs; the scheduling
        // barrier per step keeps the compiler from reading further ahead,
        // which would take registers the query fragments need
        i32x4 b[2][RS];
#pragma unroll
        for (int rf = 0; rf < RS; ++rf) b[0][rf] = *reinterpret_cast<const i32x4*>(buf0 + rf * NKS * 1024);
#pragma unroll
        for (int k = 0; k < NKS; ++k) {
            if (k + 1 < NKS) {
#pragma unroll
                for (int rf = 0; rf < RS; ++rf)
                    b[(k + 1) & 1][rf] = *reinterpret_cast<const i32x4*>(((k + 1) & 1 ? buf1 : buf0) + rf * NKS * 1024 +
                                                                         ((k + 1) >> 1) * 1024);
            }
#pragma unroll
            for (int f = 0; f < QF; ++f)
#pragma unroll
                for (int rf = 0; rf < RS; ++rf)
                    acc[f][rf] = __builtin_amdgcn_mfma_i32_32x32x32_i8(qreg[f][k], b[k & 1][rf],
                                                                       k == 0 ? (i32x16){} : acc[f][rf], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        // prefilter (dot - D_r against the query's bound, max over the
        // lane's two rows)
        float xn[RS];
#pragma unroll
        for (int rf = 0; rf < RS; ++rf) xn[rf] = xn_s[(st % NST) * SR + 32 * rf + j];
        int dofs[RS];
#pragma unroll
        for (int rf = 0; rf < RS; ++rf) dofs[rf] = wide_row_offset(a, xn[rf], kap);
        uint32_t pend = 0;
#pragma unroll
        for (int f = 0; f < QF; ++f) {
            bool maybe = false;
#pragma unroll
            for (int e4 = 0; e4 < 4; ++e4) {
                const i32x4 bv = *reinterpret_cast<const i32x4*>(bgl + 32 * f + 8 * e4);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int v0 = acc[f][0][4 * e4 + i] - dofs[0];
                    const int v1 = acc[f][1][4 * e4 + i] - dofs[1];
                    maybe |= (v0 > v1 ? v0 : v1) >= bv[i];
                }
            }
            pend |= maybe ? 1u << f : 0u;
        }
        // exact test of the pairs that pass the prefilter: in a fragment
        // where some lane may pass, each lane marks its elements that pass,
        // parks the fragment's 16 dots in its LDS slot and tests the marked
        // ones one at a time (a loop of as many turns as the most any lane
        // marked, usually one or two; not unrolled, so the rare appends do
        // not multiply the code and its live registers)
        if (__ballot(pend != 0u) != 0) {
#pragma unroll
            for (int f = 0; f < QF; ++f) {
                if (__ballot((pend >> f) & 1u) == 0) continue;
#pragma unroll
                for (int rf = 0; rf < RS; ++rf) {
                    const int64_t row = r0 + 32 * rf + j;
                    uint32_t mk = 0;
#pragma unroll
                    for (int e4 = 0; e4 < 4; ++e4) {
                        const i32x4 bv = *reinterpret_cast<const i32x4*>(bgl + 32 * f + 8 * e4);
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            mk |= (acc[f][rf][4 * e4 + i] - dofs[rf] >= bv[i]) ? 1u << (4 * e4 + i) : 0u;
                    }
                    if (row >= r_end) mk = 0;
                    if (__ballot(mk != 0u) == 0) continue;
#pragma unroll
                    for (int e4 = 0; e4 < 4; ++e4)
                        *reinterpret_cast<i32x4*>(slot + (e4 * 64 + lane) * 4) =
                            (i32x4){acc[f][rf][4 * e4], acc[f][rf][4 * e4 + 1], acc[f][rf][4 * e4 + 2],
                                    acc[f][rf][4 * e4 + 3]};
                    const float xr = xn[rf];
#pragma unroll 1
                    while (__ballot(mk != 0u) != 0) {
                        if (mk != 0u) {
                            const int e = __builtin_ctz(mk);
                            mk &= mk - 1u;
                            const int m = 32 * f + 8 * (e >> 2) + 4 * h + (e & 3);
                            const int mw = 64 * wave + m;
                            const float th = thr_s[mw], qv = qn_s[mw], sc = 2.0f * qsc_s[mw];
                            const float dot = (float)slot[((e >> 2) * 64 + lane) * 4 + (e & 3)];
                            const float dist = xr + qv - sc * dot;
                            if (dist <= th)
                                wide_append(a, wq, q0w + m, m, (uint32_t)(row - r_begin), (uint32_t)row, dist);
                        }
                    }
                }
            }
        }
        if (__builtin_amdgcn_readfirstlane(*wq.count) + 256u > (uint32_t)qcap) wide_flush(a, wq, q0w, r_begin, lane);
    }
    // the repeat loads of the last two iterations land before the
    // workgroup's LDS is released
    rl_wait_vm<0>();
    wide_flush(a, wq, q0w, r_begin, lane);
}

template <int NKS>
hipError_t launch_rl_t(const ScanArgs& a, hipStream_t st) {
    static const char* const name = [] {
        static char buf[80];
        snprintf(buf, sizeof(buf), "scan_rl_kernel<sq8,NKS=%d>", NKS);
        return buf;
    }();
    t_variant = name;
    constexpr int NW = 8, NST = 3, QT = 512, SR = 64;
    if (!a.tiled || a.prefilter != 3 || a.kdim != 32 * NKS || a.xstride != a.kdim || !a.qscale)
        return hipErrorInvalidValue;
    const size_t fixed = (size_t)NST * SR * 32 * NKS + (size_t)NST * SR * 4 + (size_t)QT * 16 + 64 +
                         (size_t)NW * kFlushSlots * 4 + (size_t)NW * 4096;
    if (fixed + (size_t)NW * 256 * 8 > 160 * 1024) return hipErrorInvalidValue;
    int qcap = (int)((160 * 1024 - fixed) / (NW * 8)) / 64 * 64;
    if (qcap > 2048) qcap = 2048;
    if (const char* e = getenv("ANNMI_WIDE_QCAP")) {  // lower cap: tuning and tests of the spill path
        const int v = atoi(e);
        if (v >= 0 && v < qcap) qcap = v / 64 * 64;
    }
    static const bool attr = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&scan_rl_kernel<NKS>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        return true;
    }();
    (void)attr;
    const int ncu = device_cu_count();
    const int qtiles = (a.B + QT - 1) / QT;
    const int64_t rows = a.row_end - a.row_begin;
    static const int64_t chunk_env = [] {
        const char* e = getenv("ANNMI_WIDE_CHUNK_MAX");
        const long v = e ? atol(e) : 0;
        return v >= 2048 ? (int64_t)v : (int64_t)32768;
    }();
    const int64_t chunk_max = a.chunk_max >= 2048 ? a.chunk_max : chunk_env;
    // one workgroup a CU at a time; about 8 a CU over the launch
    int64_t chunk = rows * qtiles / ((int64_t)ncu * 8);
    chunk = chunk < 1024 ? 1024 : (chunk > chunk_max ? chunk_max : chunk);
    chunk = (chunk + SR - 1) / SR * SR;
    const int64_t nrc = (rows + chunk - 1) / chunk;
    const int64_t per = (nrc + 7) / 8;
    ANNMI_LAUNCH((scan_rl_kernel<NKS>), dim3((unsigned)(8 * per * qtiles)), dim3(64 * NW),
                 fixed + (size_t)NW * qcap * 8, st, a, qtiles, chunk, nrc, qcap);
    return hipGetLastError();
}

// LDS of a wide configuration's query tile: 16*QS queries of qrow bytes plus
// five per-query floats (the queues take what is left of the 160 KB).
template <class P>
bool wide_fits(int qs, int kdim) {
    return (size_t)16 * qs * ((size_t)qrow_bytes<P>(kdim) + 20) + 8 * 1024 <= 160 * 1024;
}

// Query tile sizes (16*QS queries) per element type and dim, LDS <= 160 KB.
// cfg selects one of the kWideCfg* shapes (the autotuner's choice); the
// default picks per element type: 128 x 64 tiles on 4 waves for fp32, 128 x
// 32 tiles on 8 waves for bf16 and the byte-sized codes, smaller query tiles
// where 128 queries do not fit the LDS.
// the 32 x 32 MFMA form of a scan type (void: none)
template <class P>
struct W32Of {
    typedef void T;
};
template <>
struct W32Of<PolSQ8> {
    typedef W32SQ8 T;
};
template <>
struct W32Of<PolBF16> {
    typedef W32BF16 T;
};

// 32-byte K steps of a row of the 32 x 32 kernel's element type
template <class W>
int nks_of(const ScanArgs& a) {
    return a.kdim * W::kElemBytes / 32;
}

template <class P>
hipError_t launch_wide(const ScanArgs& a, hipStream_t st) {
    typedef typename W32Of<P>::T W;
    if constexpr (!std::is_void<W>::value) {
        if constexpr (std::is_same<P, PolSQ8>::value) {
            if (a.wide_cfg == kWideCfgRL) {
                switch (a.kdim) {
                    case 128: return launch_rl_t<4>(a, st);
                    case 256: return launch_rl_t<8>(a, st);
                    default: return hipErrorInvalidValue;
                }
            }
            if ((a.wide_cfg == kWideCfg32R4W4 || a.wide_cfg == kWideCfg32R2W4) && a.tiled && a.prefilter == 3) {
                const bool d4 = nks_of<W>(a) % 4 == 0;
                if (a.wide_cfg == kWideCfg32R4W4)
                    return d4 ? launch_i8x32_t<1, 4, 8>(a, st) : launch_i8x32_t<1, 2, 8>(a, st);
                return d4 ? launch_i8x32_t<2, 4, 8>(a, st) : launch_i8x32_t<2, 2, 8>(a, st);
            }
        }
        if (a.wide_cfg == kWideCfg32R4W4)
            return nks_of<W>(a) % 4 == 0 ? launch_wide32_t<W, 4, 4, 4, 4, true>(a, st)
                                         : launch_wide32_t<W, 4, 4, 4, 2, true>(a, st);
        if (a.wide_cfg == kWideCfg32R2W4)
            return nks_of<W>(a) % 4 == 0 ? launch_wide32_t<W, 4, 2, 4, 4, true>(a, st)
                                         : launch_wide32_t<W, 4, 2, 4, 2, true>(a, st);
        if (a.wide_cfg == kWideCfg32R1 || a.wide_cfg == kWideCfg32R2) {
            // ring depth: the deepest of 8, 6, 4, 2 K steps dividing the row's
            // (rows are whole 64-byte chunks: an even count); two row tiles a
            // wave step keep 4 at most for registers
            // (one row tile a wave step: query fragments read a K step ahead,
            // 3% faster on SQ8 10M x 384; two: no registers left for that)
            const int nks = a.kdim * W::kElemBytes / 32;
            if (a.wide_cfg == kWideCfg32R2)
                return nks % 4 == 0 ? launch_wide32_t<W, 4, 2, 8, 4, false>(a, st)
                                    : launch_wide32_t<W, 4, 2, 8, 2, false>(a, st);
            if (nks % 8 == 0) return launch_wide32_t<W, 4, 1, 8, 8, true>(a, st);
            if (nks % 6 == 0) return launch_wide32_t<W, 4, 1, 8, 6, true>(a, st);
            if (nks % 4 == 0) return launch_wide32_t<W, 4, 1, 8, 4, true>(a, st);
            return launch_wide32_t<W, 4, 1, 8, 2, true>(a, st);
        }
    }
    switch (a.wide_cfg) {
        case kWideCfgQ8R4W4:
            if (wide_fits<P>(8, a.kdim)) return launch_wide_t<P, 8, 4, 4, 1, true>(a, st);
            break;
        case kWideCfgQ8R2W8:
            if (wide_fits<P>(8, a.kdim)) return launch_wide_t<P, 8, 2, 8, 1, false>(a, st);
            break;
        case kWideCfgQ4R4W8:
            if (wide_fits<P>(4, a.kdim)) return launch_wide_t<P, 4, 4, 8, 1, false>(a, st);
            break;
        default: break;
    }
    const size_t row = (size_t)qrow_bytes<P>(a.kdim) + 12;
    if (128 * row <= 160 * 1024) {
        // fp32 (small d): 128 x 64 tiles on 4 waves; bf16 and the byte codes:
        // 128 x 32 tiles on 8 waves (with the tiled row layout bf16 measured
        // 1053 vs 952 TFLOP/s at 4096 queries, 10M x 384)
        if constexpr (P::kElemBytes == 4) {
            return launch_wide_t<P, 8, 4, 4, 1, true>(a, st);
        } else {
            // 384-dim SQ8 with the dim fixed at compile time; from 256
            // queries (MFMA-bound launches) on 12 waves, three a SIMD in 168
            // VGPRs (10M x 384, 4096 queries: 2240 -> 2390 TOP/s; the
            // HBM-bound launches of fewer queries stream faster on 8 waves,
            // one query 0.868 -> 0.804 ms a step)
            if constexpr (std::is_same<P, PolSQ8>::value)
                if (a.kdim == 384 && a.tiled)
                    return a.B >= 256 ? launch_wide_t<P, 8, 2, 12, 1, false, 384>(a, st)
                                      : launch_wide_t<P, 8, 2, 8, 1, false, 384>(a, st);
            // FP8 on the 128-deep MFMA where the chunks pair up
            if constexpr (std::is_same<P, PolFP8>::value) {
                if (a.kdim == 384 && a.tiled) return launch_wide_t<PolFP8P, 8, 2, 8, 1, false, 384>(a, st);
                if (a.kdim % 128 == 0) return launch_wide_t<PolFP8P, 8, 2, 8, 1, false>(a, st);
            }
            return launch_wide_t<P, 8, 2, 8, 1, false>(a, st);
        }
    }
    if (64 * row <= 160 * 1024) return launch_wide_t<P, 4, 4, 8, 1, false>(a, st);
    if (32 * row <= 160 * 1024) return launch_wide_t<P, 2, 4, 8, 1, false>(a, st);
    return hipErrorInvalidValue;
}

// Largest query tile (16, 32 or 64 queries) that the batch can use and
// whose LDS staging stays within 64 KB (several workgroups a CU); rows too
// long for that take 16-query tiles of up to the 160 KB of LDS (fp32 to
// d = 2544).
template <class P>
hipError_t launch_p(const ScanArgs& a, hipStream_t st) {
    const size_t row = (size_t)qrow_bytes<P>(a.kdim);
    if (a.B > 32 && 64 * row <= 65536) return launch_t<P, 4>(a, st);
    if (a.B > 16 && 32 * row <= 65536) return launch_t<P, 2>(a, st);
    if (16 * row <= 160 * 1024) return launch_t<P, 1>(a, st);
    return hipErrorInvalidValue;
}

}  // namespace

// Batches from this size up take the wide kernel: the crossover measured on
// 10M x 384 (the streaming kernel re-reads rows per 64-query tile and
// appends with returning atomics; the wide kernel pays a query-tile staging
// per row chunk).  ANNMI_WIDE_MIN_BATCH overrides it for every type (tuning).
int wide_min_batch(int kind) {
    static const int forced = [] {
        const char* e = getenv("ANNMI_WIDE_MIN_BATCH");
        return e ? atoi(e) : 0;
    }();
    if (forced > 0) return forced;
    switch (kind) {
        case SCAN_F32: return 33;
        case SCAN_FP8: return 4;
        // integer codes: the wide kernel's XCD-aware row chunks stream the
        // codes faster even for one query (10M x 384 scan time a step, one /
        // two / four queries: SQ8 0.733 / 0.764 / 0.850 -> 0.684 / 0.694 /
        // 0.699 ms; SQ4 0.554 / 0.595 / 0.696 -> 0.540 / 0.543 / 0.545 ms);
        // bf16 stays on the streaming kernel below 8 (3% faster there)
        case SCAN_SQ8:
        case SCAN_SQ4: return 1;
        default: return 8;  // bf16
    }
}

hipError_t launch_scan_impl(const ScanArgs& a, hipStream_t st);

hipError_t launch_scan(const ScanArgs& a, hipStream_t st, const char** variant) {
    t_variant = "";
    hipError_t e = launch_scan_impl(a, st);
    if (variant) *variant = t_variant;
    return e;
}

bool wide_cfg_fits(int kind, int kdim, int cfg) {
    const int qs = cfg == kWideCfgQ4R4W8 ? 4 : 8;
    if (cfg <= kWideCfgDefault || cfg >= kWideCfgCount) return false;
    if (cfg == kWideCfgRL) return kind == SCAN_SQ8 && (kdim == 128 || kdim == 256);
    if (cfg == kWideCfg32R1 || cfg == kWideCfg32R2 || cfg == kWideCfg32R4W4 ||
        cfg == kWideCfg32R2W4) {  // 128 queries, whole 64-byte row chunks
        if (kind != SCAN_SQ8 && kind != SCAN_BF16) return false;
        const int qbytes = kind == SCAN_SQ8 ? kdim : 2 * kdim;
        if (qbytes % 64 != 0 || (size_t)128 * (qbytes + 16 + 20) + 8 * 1024 > 160 * 1024) return false;
        // SQ8 on these two takes the lean-epilogue kernel, whose fixed LDS
        // (exact-test slots, a minimum queue) leaves room for fewer dims
        if (kind == SCAN_SQ8 && (cfg == kWideCfg32R4W4 || cfg == kWideCfg32R2W4))
            return i8x32_lds_need(qbytes, 8) <= 160 * 1024;
        return true;
    }
    switch (kind) {
        case SCAN_F32: return wide_fits<PolF32>(qs, kdim);
        case SCAN_BF16: return wide_fits<PolBF16>(qs, kdim);
        case SCAN_FP8: return wide_fits<PolFP8>(qs, kdim);
        case SCAN_SQ8: return wide_fits<PolSQ8>(qs, kdim);
        case SCAN_SQ4: return wide_fits<PolSQ4>(qs, kdim);
    }
    return false;
}

hipError_t launch_scan_impl(const ScanArgs& a, hipStream_t st) {
    if (a.row_end <= a.row_begin || a.B <= 0) return hipSuccess;
    // argmin scans (k-means assignment: few rows, many queries) take the
    // streaming kernel, whose per-lane minima the wide kernel has no
    // registers left for
    // The wide kernel stages a query tile per row chunk: it takes ranges of
    // at least 65536 rows, or smaller ones when the batch is large enough
    // that the streaming kernel's re-reads of every row per 64-query tile
    // cost more (rows x queries >= 2^24).
    const int min_b = a.wide_min_batch > 0 ? a.wide_min_batch : wide_min_batch(a.kind);
    const int64_t nrows = a.row_end - a.row_begin;
    if (a.B >= min_b && (nrows >= 65536 || nrows * a.B >= ((int64_t)1 << 24)) && !a.argmin) {
        switch (a.kind) {
            case SCAN_F32: return launch_wide<PolF32>(a, st);
            case SCAN_BF16: return launch_wide<PolBF16>(a, st);
            case SCAN_FP8: return launch_wide<PolFP8>(a, st);
            case SCAN_SQ8: return launch_wide<PolSQ8>(a, st);
            case SCAN_SQ4: return launch_wide<PolSQ4>(a, st);
        }
        return hipErrorInvalidValue;
    }
    switch (a.kind) {
        case SCAN_F32: return launch_p<PolF32>(a, st);
        case SCAN_BF16: return launch_p<PolBF16>(a, st);
        case SCAN_FP8: return launch_p<PolFP8>(a, st);
        case SCAN_SQ8: return launch_p<PolSQ8>(a, st);
        case SCAN_SQ4: return launch_p<PolSQ4>(a, st);
    }
    return hipErrorInvalidValue;
}

