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

namespace {

constexpr int kThreads = 256;

// ----------------------------------------------------------------- policies
struct PolF32 {
    static constexpr int kDimsPerLoad = 4;     // per lane per 16B load
    static constexpr int kChunk = 16;          // dims per wave-wide load step (4 lane groups)
    static constexpr int kElemBytes = 4;
    typedef v4f Acc;
    __device__ static inline void mma(v4f& acc, const i32x4& a, const i32x4& b) {
        f32x4 af = __builtin_bit_cast(f32x4, a);
        f32x4 bf = __builtin_bit_cast(f32x4, b);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[0], bf[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[1], bf[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[2], bf[2], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[3], bf[3], acc, 0, 0, 0);
    }
};

struct PolBF16 {
    static constexpr int kDimsPerLoad = 8;
    static constexpr int kChunk = 32;
    static constexpr int kElemBytes = 2;
    typedef v4f Acc;
    __device__ static inline void mma(v4f& acc, const i32x4& a, const i32x4& b) {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                      __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
    }
};

struct PolFP8 {
    static constexpr int kDimsPerLoad = 16;
    static constexpr int kChunk = 64;
    static constexpr int kElemBytes = 1;
    typedef v4f Acc;
    __device__ static inline void mma(v4f& acc, const i32x4& a, const i32x4& b) {
        long a0 = ((long)(uint32_t)a[1] << 32) | (uint32_t)a[0];
        long a1 = ((long)(uint32_t)a[3] << 32) | (uint32_t)a[2];
        long b0 = ((long)(uint32_t)b[1] << 32) | (uint32_t)b[0];
        long b1 = ((long)(uint32_t)b[3] << 32) | (uint32_t)b[2];
        acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a1, b1, acc, 0, 0, 0);
    }
};

// int8 MFMA with an int32 accumulator.
struct PolSQ8 {
    static constexpr int kDimsPerLoad = 16;
    static constexpr int kChunk = 64;
    static constexpr int kElemBytes = 1;
    typedef i32x4 Acc;
    __device__ static inline void mma(i32x4& acc, const i32x4& a, const i32x4& b) {
        // codes are uint8 in [0,255]; the MFMA reads signed int8, so flip the
        // top bit: c - 128.  The query side adds 128*sum(w) back.
        i32x4 bs = b ^ (i32x4){(int)0x80808080, (int)0x80808080, (int)0x80808080, (int)0x80808080};
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, bs, acc, 0, 0, 0);
    }
};

// SQ4: 8 packed bytes per lane per chunk of 64 dims.  Unpacked as
// [lo nibbles of bytes 0..7, hi nibbles of bytes 0..7] (values 0..15, exact
// int8); the query operand is stored in the same permuted order.
struct PolSQ4 {
    static constexpr int kDimsPerLoad = 16;
    static constexpr int kChunk = 64;
    static constexpr int kElemBytes = 0;  // packed: 2 dims per byte
    typedef i32x4 Acc;
    __device__ static inline void mma(i32x4& acc, const i32x4& a, const i32x4& b) {
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc, 0, 0, 0);
    }
};


// Row fragment load: 16 bytes of row `row` covering dims [dim0, dim0 + kDimsPerLoad).
template <class P>
__device__ inline i32x4 load_row_frag(const uint8_t* X, int64_t xstride, int64_t row, int dim0) {
    if constexpr (P::kElemBytes == 0) {
        // SQ4: chunk c covers 64 dims = 32 packed bytes; lane group g reads
        // bytes [32c + 8g, +8): dim0 here is 64c + 16g.
        const uint8_t* p = X + row * xstride + (dim0 >> 1);
        uint2 v = *reinterpret_cast<const uint2*>(p);
        // unpack 8 bytes -> 16 int8: lows then highs
        uint32_t lo0 = v.x & 0x0F0F0F0Fu, lo1 = v.y & 0x0F0F0F0Fu;
        uint32_t hi0 = (v.x >> 4) & 0x0F0F0F0Fu, hi1 = (v.y >> 4) & 0x0F0F0F0Fu;
        return (i32x4){(int)lo0, (int)lo1, (int)hi0, (int)hi1};
    } else {
        const uint8_t* p = X + row * xstride + (int64_t)dim0 * P::kElemBytes;
        return *reinterpret_cast<const i32x4*>(p);
    }
}

// Query tile in LDS: 16*QW queries x kdim elements, stored in the scan
// element type, row stride padded by 16 bytes against bank conflicts.
template <class P>
__host__ __device__ inline int qrow_bytes(int kdim) {
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

    for (int64_t t = t_begin + wave; t < t_end; t += 4) {
        const int64_t row = a.row_begin + t * 16 + n;
        const bool row_ok = row < a.row_end;
        const int64_t lrow = row_ok ? row : a.row_end - 1;
        typename P::Acc acc[QW];
#pragma unroll
        for (int w = 0; w < QW; ++w) acc[w] = typename P::Acc{};

        constexpr int U = 8;
        i32x4 buf[U];
        int c = 0;
        for (; c + U <= nchunk; c += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) buf[u] = load_row_frag<P>(X, a.xstride, lrow, (c + u) * P::kChunk + lane_dim);
#pragma unroll
            for (int u = 0; u < U; ++u) {
#pragma unroll
                for (int w = 0; w < QW; ++w) {
                    const int qoff = (16 * w + n) * ldq;
                    int eoff = P::kElemBytes == 0 ? ((c + u) * P::kChunk + lane_dim) : ((c + u) * P::kChunk + lane_dim) * P::kElemBytes;
                    i32x4 qa = *reinterpret_cast<const i32x4*>(smem + qoff + eoff);
                    P::mma(acc[w], qa, buf[u]);
                }
            }
        }
        for (; c < nchunk; ++c) {
            i32x4 b = load_row_frag<P>(X, a.xstride, lrow, c * P::kChunk + lane_dim);
#pragma unroll
            for (int w = 0; w < QW; ++w) {
                const int qoff = (16 * w + n) * ldq;
                int eoff = P::kElemBytes == 0 ? (c * P::kChunk + lane_dim) : (c * P::kChunk + lane_dim) * P::kElemBytes;
                i32x4 qa = *reinterpret_cast<const i32x4*>(smem + qoff + eoff);
                P::mma(acc[w], qa, b);
            }
        }

        // epilogue: distance, filter, append
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
            continue;
        }
#pragma unroll
        for (int w = 0; w < QW; ++w) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float dot = (float)acc[w][i];
                // dist = xn + qn - 2 * qscale * dot  (qscale folds the query/base scaling)
                float dist = xn + qn[w][i] - 2.0f * qs[w][i] * dot;
                bool pass = row_ok && dist <= thr[w][i];
                uint64_t m = __ballot(pass);
                if (m == 0) continue;
                uint32_t gm = (uint32_t)(m >> (16 * g)) & 0xFFFFu;
                int qi = q0 + 16 * w + 4 * g + i;
                uint32_t base = 0;
                if (n == 0 && gm) base = atomicAdd(&a.cand.cnt[qi], (uint32_t)__builtin_popcount(gm));
                base = __shfl(base, lane & ~15);
                if (pass) {
                    uint32_t off = base + __builtin_popcount(gm & ((1u << n) - 1u));
                    if (off < (uint32_t)a.cand.cap) {
                        int64_t slot = (int64_t)qi * a.cand.cap + off;
                        a.cand.key[slot] = dist;
                        a.cand.rid[slot] = a.rid_map ? a.rid_map[row] : (uint32_t)row;
                    }
                }
            }
        }
    }
    if (a.argmin) argmin_flush<QW>(best, a, q0, g, n);
}

template <class P>
const char* pol_name();
template <> const char* pol_name<PolF32>() { return "f32"; }
template <> const char* pol_name<PolBF16>() { return "bf16"; }
template <> const char* pol_name<PolFP8>() { return "fp8"; }
template <> const char* pol_name<PolSQ8>() { return "sq8"; }
template <> const char* pol_name<PolSQ4>() { return "sq4"; }

thread_local const char* t_variant = "";

template <class P, int QW>
hipError_t launch_t(const ScanArgs& a, hipStream_t st) {
    static char name[64] = "";
    if (!name[0]) snprintf(name, sizeof(name), "scan_kernel<%s,QW=%d>", pol_name<P>(), QW);
    t_variant = name;
    const int qtile = 16 * QW;
    const int qtiles = (a.B + qtile - 1) / qtile;
    const int64_t rows = a.row_end - a.row_begin;
    const int64_t tiles = (rows + 15) / 16;
    // aim for ~8 workgroups per CU across all query tiles, >= 4 tiles per wave
    int64_t want = (2048 + qtiles - 1) / qtiles;
    int64_t maxb = (tiles + 15) / 16;
    int64_t gx = want < maxb ? want : maxb;
    if (gx < 1) gx = 1;
    size_t lds = (size_t)qtile * qrow_bytes<P>(a.kdim);
    hipLaunchKernelGGL((scan_kernel<P, QW>), dim3((unsigned)gx, (unsigned)qtiles), dim3(kThreads), lds, st, a);
    return hipGetLastError();
}


// ------------------------------------------------------------------ wide --
// Large batches: query-stationary persistent kernel.  A workgroup (4-8 waves)
// keeps a tile of 16*QS queries in LDS for its whole life and streams a
// contiguous row range; every wave scores 16*QS queries x 16*RS rows per
// step, row fragments loaded from L2/HBM straight into MFMA B registers
// (double-buffered over the K chunks), query fragments from LDS.  Per K
// chunk a wave issues QS*RS MFMA groups for RS global and QS LDS fragment
// loads, which keeps LDS traffic at ~half its bandwidth at full MFMA rate.
//
// Workgroup -> work mapping is XCD-aware: dispatch places workgroup b on XCD
// b % 8, so the QTILES query tiles of one row range are given consecutive
// slots b/8 on the same XCD.  They stream the same rows at the same time and
// share that XCD's L2, so HBM sees each row once per launch.
template <class P, int QS, int RS, int kWideWaves>
__global__ __launch_bounds__(64 * kWideWaves) void scan_wide_kernel(ScanArgs a, int qtiles, int rpx, int64_t range_len) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int n = lane & 15;
    const int xcd = blockIdx.x & 7;
    const int slot = blockIdx.x >> 3;
    if (slot >= qtiles * rpx) return;
    const int qtile = slot % qtiles;
    const int rsplit = xcd * rpx + slot / qtiles;
    const int64_t r_begin = a.row_begin + (int64_t)rsplit * range_len;
    if (r_begin >= a.row_end) return;
    const int64_t r_end = r_begin + range_len < a.row_end ? r_begin + range_len : a.row_end;

    constexpr int QT = 16 * QS;
    const int q0 = qtile * QT;
    const int qbytes = P::kElemBytes == 0 ? a.kdim : a.kdim * P::kElemBytes;
    const int ldq = qrow_bytes<P>(a.kdim);
    float* thr_s = reinterpret_cast<float*>(smem + QT * ldq);
    float* qn_s = thr_s + QT;
    float* qsc_s = qn_s + QT;
    for (int idx = threadIdx.x; idx < QT * (qbytes / 16); idx += 64 * kWideWaves) {
        int r = idx / (qbytes / 16);
        int c = idx - r * (qbytes / 16);
        i32x4 v = (i32x4){0, 0, 0, 0};
        if (q0 + r < a.B) v = reinterpret_cast<const i32x4*>(static_cast<const uint8_t*>(a.Q) + (int64_t)(q0 + r) * qbytes)[c];
        *reinterpret_cast<i32x4*>(smem + r * ldq + c * 16) = v;
    }
    for (int m = threadIdx.x; m < QT; m += 64 * kWideWaves) {
        int qi = q0 + m;
        bool ok = qi < a.B;
        thr_s[m] = ok ? a.thr[qi] : -__builtin_inff();
        qn_s[m] = ok ? a.qn[qi] : 0.f;
        qsc_s[m] = (ok && a.qscale) ? a.qscale[qi] : 1.f;
    }
    __syncthreads();

    const uint8_t* X = static_cast<const uint8_t*>(a.X);
    const int nchunk = a.kdim / P::kChunk;
    const int lane_dim = P::kDimsPerLoad * g;
    constexpr int TR = 16 * RS;  // rows per wave step
    const int64_t ntiles = (r_end - r_begin + TR - 1) / TR;
    unsigned long long best[QS][4];
#pragma unroll
    for (int q = 0; q < QS; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) best[q][i] = ~0ull;

    for (int64_t t = wave; t < ntiles; t += kWideWaves) {
        const int64_t r0 = r_begin + t * TR;
        int64_t lrow[RS];
        bool rok[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            int64_t row = r0 + 16 * r + n;
            rok[r] = row < r_end;
            lrow[r] = rok[r] ? row : r_end - 1;
        }
        typename P::Acc acc[QS][RS];
#pragma unroll
        for (int q = 0; q < QS; ++q)
#pragma unroll
            for (int r = 0; r < RS; ++r) acc[q][r] = typename P::Acc{};
        i32x4 bc[RS], bn[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) bc[r] = load_row_frag<P>(X, a.xstride, lrow[r], lane_dim);
        for (int c = 0; c < nchunk; ++c) {
            if (c + 1 < nchunk) {
#pragma unroll
                for (int r = 0; r < RS; ++r) bn[r] = load_row_frag<P>(X, a.xstride, lrow[r], (c + 1) * P::kChunk + lane_dim);
            }
            const int eoff = P::kElemBytes == 0 ? (c * P::kChunk + lane_dim) : (c * P::kChunk + lane_dim) * P::kElemBytes;
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                i32x4 qa = *reinterpret_cast<const i32x4*>(smem + (16 * q + n) * ldq + eoff);
#pragma unroll
                for (int r = 0; r < RS; ++r) P::mma(acc[q][r], qa, bc[r]);
            }
#pragma unroll
            for (int r = 0; r < RS; ++r) bc[r] = bn[r];
        }
        // epilogue: cheap test first, append only when something passes
        float xn[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) xn[r] = a.xn[lrow[r]];
        if (a.argmin) {
#pragma unroll
            for (int q = 0; q < QS; ++q)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int m = 16 * q + 4 * g + i;
                    const float qv = qn_s[m], sc = 2.0f * qsc_s[m];
#pragma unroll
                    for (int r = 0; r < RS; ++r) {
                        float dist = xn[r] + qv - sc * (float)acc[q][r][i];
                        unsigned long long kv = ((unsigned long long)f2key(dist) << 32) | (uint32_t)(r0 + 16 * r + n);
                        if (rok[r] && kv < best[q][i]) best[q][i] = kv;
                    }
                }
            continue;
        }
        bool any = false;
        v4f dist4[QS][RS];
#pragma unroll
        for (int q = 0; q < QS; ++q)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * q + 4 * g + i;
                const float th = thr_s[m], qv = qn_s[m], sc = 2.0f * qsc_s[m];
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    float dist = xn[r] + qv - sc * (float)acc[q][r][i];
                    dist4[q][r][i] = dist;
                    any |= rok[r] && dist <= th;
                }
            }
        if (__ballot(any) == 0) continue;
#pragma unroll
        for (int q = 0; q < QS; ++q)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * q + 4 * g + i;
                const float th = thr_s[m];
                const int qi = q0 + m;
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    const float dist = dist4[q][r][i];
                    const bool pass = rok[r] && dist <= th;
                    uint64_t mb = __ballot(pass);
                    if (mb == 0) continue;
                    uint32_t gm = (uint32_t)(mb >> (16 * g)) & 0xFFFFu;
                    uint32_t base = 0;
                    if (n == 0 && gm) base = atomicAdd(&a.cand.cnt[qi], (uint32_t)__builtin_popcount(gm));
                    base = __shfl(base, lane & ~15);
                    if (pass) {
                        uint32_t off = base + __builtin_popcount(gm & ((1u << n) - 1u));
                        if (off < (uint32_t)a.cand.cap) {
                            int64_t sl = (int64_t)qi * a.cand.cap + off;
                            const uint32_t row = (uint32_t)(r0 + 16 * r + n);
                            a.cand.key[sl] = dist;
                            a.cand.rid[sl] = a.rid_map ? a.rid_map[row] : row;
                        }
                    }
                }
            }
    }
    if (a.argmin) argmin_flush<QS>(best, a, q0, g, n);
}

template <class P, int QS, int RS, int NW>
hipError_t launch_wide_t(const ScanArgs& a, hipStream_t st) {
    static char name[64] = "";
    if (!name[0]) snprintf(name, sizeof(name), "scan_wide_kernel<%s,QS=%d,RS=%d,NW=%d>", pol_name<P>(), QS, RS, NW);
    t_variant = name;
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 256;
    }
    constexpr int QT = 16 * QS;
    const size_t lds = (size_t)QT * qrow_bytes<P>(a.kdim) + 3 * QT * sizeof(float);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&scan_wide_kernel<P, QS, RS, NW>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    const int qtiles = (a.B + QT - 1) / QT;
    int S = ncu / 8;                       // workgroup slots per XCD (one workgroup per CU)
    int rpx = S / qtiles;                  // row ranges per XCD
    if (rpx < 1) rpx = 1;
    const int slots = qtiles * rpx;
    const int64_t rows = a.row_end - a.row_begin;
    constexpr int TR = 16 * RS;
    int64_t range_len = (rows + 8 * rpx - 1) / (8 * rpx);
    range_len = (range_len + TR - 1) / TR * TR;
    hipLaunchKernelGGL((scan_wide_kernel<P, QS, RS, NW>), dim3((unsigned)(8 * slots)), dim3(64 * NW), lds, st, a,
                       qtiles, rpx, range_len);
    return hipGetLastError();
}

// Query tile sizes (16*QS queries) per element type and dim, LDS <= 160 KB.
template <class P>
hipError_t launch_wide(const ScanArgs& a, hipStream_t st) {
    const size_t row = (size_t)qrow_bytes<P>(a.kdim) + 12;
    // 128-query tiles need ~300 registers per lane: one wave per SIMD
    if (128 * row <= 160 * 1024) return launch_wide_t<P, 8, 4, 4>(a, st);
    if (64 * row <= 160 * 1024) return launch_wide_t<P, 4, 4, 8>(a, st);
    if (32 * row <= 160 * 1024) return launch_wide_t<P, 2, 4, 8>(a, st);
    return hipErrorInvalidValue;
}

// Largest query tile (16, 32 or 64 queries) that the batch can use and
// whose LDS staging stays within 64 KB.
template <class P>
hipError_t launch_p(const ScanArgs& a, hipStream_t st) {
    const size_t row = (size_t)qrow_bytes<P>(a.kdim);
    if (a.B > 32 && 64 * row <= 65536) return launch_t<P, 4>(a, st);
    if (a.B > 16 && 32 * row <= 65536) return launch_t<P, 2>(a, st);
    if (16 * row <= 65536) return launch_t<P, 1>(a, st);
    return hipErrorInvalidValue;
}

}  // namespace

int g_wide_min_batch = 128;

hipError_t launch_scan_impl(const ScanArgs& a, hipStream_t st);

hipError_t launch_scan(const ScanArgs& a, hipStream_t st, const char** variant) {
    t_variant = "";
    hipError_t e = launch_scan_impl(a, st);
    if (variant) *variant = t_variant;
    return e;
}

hipError_t launch_scan_impl(const ScanArgs& a, hipStream_t st) {
    if (a.row_end <= a.row_begin || a.B <= 0) return hipSuccess;
    if (a.B >= g_wide_min_batch && (a.row_end - a.row_begin) >= 65536) {
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

