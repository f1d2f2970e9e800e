// Base-row encoders (fit time) and per-batch query preparation.
//
// Every scan element type stores rows zero-padded to `kdim` (a multiple of
// the MFMA K chunk), so padded dims contribute exactly 0 to every dot and
// norm.  Each encoder also writes the per-row norm term the scan adds:
//   F32  : ||x||^2                      (fp32 accumulate)
//   BF16 : ||bf16(x)||^2
//   FP8  : ||e4m3(alpha x)||^2 / alpha^2 (alpha a power of two)
//   SQ8  : sum_d (code_d * scale_d)^2    (dequantised row, offset folded in the query)
//   SQ4  : same with 4-bit codes
//
// SQ8 code rule: competitors/ivf_lsh/src/lsh_index_optimized.cpp:206-213
// (variant 0, std::round = half away from zero) or gpu_filter.cu:119-126
// (variant 1, lrintf = half to even); clamp to [0, 255].
// SQ4 code rule and nibble layout: submission/src/hnsw.cpp:170-182.
#include "common.h"

namespace {

__device__ inline float warp16_sum(float s) {
    s += __shfl_xor(s, 8, 16);
    s += __shfl_xor(s, 4, 16);
    s += __shfl_xor(s, 2, 16);
    s += __shfl_xor(s, 1, 16);
    return s;
}

// 16 lanes per row, 16 rows per 256-thread block.
__global__ __launch_bounds__(256) void enc_f32(const float* src, int64_t sstride, int64_t n, int d, int kdim,
                                               float* dst, float* xn, uint32_t* maxkey) {
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (row >= n) return;
    const float* p = src + row * sstride;
    float s = 0.f;
    for (int j = l; j < kdim; j += 16) {
        float v = j < d ? p[j] : 0.f;
        if (dst != src) dst[row * kdim + j] = v;
        s = fmaf(v, v, s);
    }
    s = warp16_sum(s);
    if (l == 0) {
        xn[row] = s;
        if (maxkey) atomicMax(maxkey, f2key(sqrtf(s)));
    }
}

// Destination of element j of row `row` (element size es bytes): row-major,
// or tiled (cb > 0: chunk bytes of the tiled layout).
__device__ inline int64_t dst_elem(int64_t row, int j, int kdim, int es, int cb) {
    return cb > 0 ? tiled_off(row, (int64_t)j * es, (int64_t)kdim * es, cb) / es : row * kdim + j;
}

__global__ __launch_bounds__(256) void enc_bf16(const float* src, int64_t sstride, int64_t n, int d, int kdim,
                                                __bf16* dst, float* xn, int cb) {
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (row >= n) return;
    const float* p = src + row * sstride;
    float s = 0.f;
    for (int j = l; j < kdim; j += 16) {
        __bf16 b = (__bf16)(j < d ? p[j] : 0.f);
        dst[dst_elem(row, j, kdim, 2, cb)] = b;
        float v = (float)b;
        s = fmaf(v, v, s);
    }
    s = warp16_sum(s);
    if (l == 0) xn[row] = s;
}

// OCP e4m3 via the hardware converter (round to nearest even, saturating),
// so the encoding is the one the fp8 MFMA reads.
__device__ inline uint8_t to_e4m3(float v) {
    int packed = __builtin_amdgcn_cvt_pk_fp8_f32(v, 0.0f, 0, false);
    return (uint8_t)(packed & 0xFF);
}
__device__ inline float from_e4m3(uint8_t b) {
    return __builtin_amdgcn_cvt_f32_fp8((int)b, 0);
}

__global__ __launch_bounds__(256) void enc_fp8(const float* src, int64_t sstride, int64_t n, int d, int kdim,
                                               float alpha, uint8_t* dst, float* xn, int cb) {
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (row >= n) return;
    const float* p = src + row * sstride;
    float s = 0.f;
    for (int j = l; j < kdim; j += 16) {
        uint8_t b = to_e4m3(j < d ? p[j] * alpha : 0.f);
        dst[dst_elem(row, j, kdim, 1, cb)] = b;
        float v = from_e4m3(b);
        s = fmaf(v, v, s);
    }
    s = warp16_sum(s);
    if (l == 0) xn[row] = s / (alpha * alpha);
}

__device__ inline int sq_round(float t, int variant) {
    return variant == 0 ? (int)roundf(t) : (int)rintf(t);
}

__global__ __launch_bounds__(256) void enc_sq8(const float* src, int64_t sstride, int64_t n, int d, int kdim,
                                               const float* scale, const float* offset, int variant, uint8_t* dst,
                                               float* xn, const uint32_t* gather, int cb) {
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (row >= n) return;
    const float* p = src + (gather ? (int64_t)gather[row] : row) * sstride;
    float s = 0.f;
    for (int j = l; j < kdim; j += 16) {
        uint8_t c = 0;
        if (j < d) {
            int qv = sq_round((p[j] - offset[j]) / scale[j], variant);
            c = (uint8_t)(qv < 0 ? 0 : (qv > 255 ? 255 : qv));
            float v = fmaf((float)c, scale[j], offset[j]);  // reconstructed value
            s = fmaf(v, v, s);
        }
        // the tiled scan copy stores c ^ 0x80, i.e. c - 128 as the int8 the
        // MFMA reads (the query side adds 128 * sum(w) back); row-major
        // copies (IVF lists) keep the plain codes
        dst[dst_elem(row, j, kdim, 1, cb)] = cb ? (uint8_t)(c ^ 0x80u) : c;
    }
    s = warp16_sum(s);
    if (l == 0) xn[row] = s;
}

// SQ4: one thread per output byte (two dims).  kdim is a multiple of 64;
// dims past d are code 0 with scale 0.
__global__ __launch_bounds__(256) void enc_sq4(const float* src, int64_t sstride, int64_t n, int d, int kdim,
                                               const float* scale, const float* offset, uint8_t* dst, float* xn,
                                               int cb) {
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (row >= n) return;
    const float* p = src + row * sstride;
    const int stride = kdim / 2;
    float s = 0.f;
    for (int b = l; b < stride; b += 16) {
        int blk = b / 16, j = b % 16;
        int dlo = 32 * blk + j, dhi = 32 * blk + 16 + j;
        int nlo = 0, nhi = 0;
        if (dlo < d) {
            int v = (int)roundf((p[dlo] - offset[dlo]) / scale[dlo]);
            nlo = v < 0 ? 0 : (v > 15 ? 15 : v);
            float f = fmaf((float)nlo, scale[dlo], offset[dlo]);
            s = fmaf(f, f, s);
        }
        if (dhi < d) {
            int v = (int)roundf((p[dhi] - offset[dhi]) / scale[dhi]);
            nhi = v < 0 ? 0 : (v > 15 ? 15 : v);
            float f = fmaf((float)nhi, scale[dhi], offset[dhi]);
            s = fmaf(f, f, s);
        }
        dst[cb > 0 ? tiled_off(row, b, stride, cb) : row * stride + b] = (uint8_t)(nlo | (nhi << 4));
    }
    s = warp16_sum(s);
    if (l == 0) xn[row] = s;
}

// Per-dim min/max over rows (orderable keys, atomics), for SQ training.
__global__ __launch_bounds__(256) void dim_minmax(const float* src, int64_t sstride, int64_t n, int d,
                                                  uint32_t* mn, uint32_t* mx) {
    const int j = blockIdx.y * 256 + threadIdx.x;
    if (j >= d) return;
    const int64_t rows_per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per;
    const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
    float lo = __builtin_inff(), hi = -__builtin_inff();
    for (int64_t r = r0; r < r1; ++r) {
        float v = src[r * sstride + j];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    if (r0 < r1) {
        atomicMin(&mn[j], f2key(lo));
        atomicMax(&mx[j], f2key(hi));
    }
}

__global__ void fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_f32(float* p, float v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ queries
// One 256-thread block per query.  Writes the scan operand (row `q` of
// Qscan, kdim elements of the scan type), qn (norm term), qscale and the
// exact-mode margin.
struct QPrepArgs {
    int kind;
    const float* Q;       // [B x d] fp32
    int B, d, kdim;
    void* Qscan;          // [B x kdim] scan element type
    float* qn;
    float* qscale;
    float* margin;        // [B] rigorous |approx - exact| bound (F32 kind), else 0
    float maxnorm;        // max ||x|| over base rows (F32 margin)
    float alpha_x;        // FP8 base scale
    const float* scale;   // SQ8/SQ4 per-dim scale (kdim entries for SQ4 padding)
    const float* offset;
    // per-query step state reset here when set (saves three fill launches):
    // candidate count 0, flags 0, filter threshold +inf
    uint32_t* cnt;
    uint32_t* flags;
    float* thr;
};

__device__ float block_reduce_sum(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return t;
}
__device__ float block_reduce_max(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    __syncthreads();
    return t;
}

// SQ4 storage order: within chunk c (64 dims), lane group g, slot s (0..15)
// holds dim 64c + 32(g/2) + 8(g%2) + (s < 8 ? s : 8 + s).
__device__ inline int sq4_dim_of_slot(int p) {
    int c = p / 64, r = p % 64, g = r / 16, s = r % 16;
    return 64 * c + 32 * (g / 2) + 8 * (g % 2) + (s < 8 ? s : 8 + s);
}

__global__ __launch_bounds__(256) void prep_queries(QPrepArgs a) {
    __shared__ float red[4];
    const int q = blockIdx.x;
    const float* x = a.Q + (int64_t)q * a.d;
    if (threadIdx.x == 0 && a.cnt) {
        a.cnt[q] = 0u;
        a.flags[q] = 0u;
        a.thr[q] = __builtin_inff();
    }
    if (a.kind == SCAN_F32) {
        float* o = static_cast<float*>(a.Qscan) + (int64_t)q * a.kdim;
        float s = 0.f;
        for (int j = threadIdx.x; j < a.kdim; j += 256) {
            float v = j < a.d ? x[j] : 0.f;
            o[j] = v;
            s = fmaf(v, v, s);
        }
        s = block_reduce_sum(s, red);
        if (threadIdx.x == 0) {
            a.qn[q] = s;
            a.qscale[q] = 1.0f;
            // |fl(xn + qn - 2 x.q) - ||x-q||^2| <= gamma * (||x|| + ||q||)^2 with
            // gamma = (d + 4) u for u = 2^-24 (fp32 dot/norm accumulation, two
            // final roundings); doubled for safety.
            float r = a.maxnorm + sqrtf(s);
            a.margin[q] = 2.0f * (float)(a.kdim + 4) * 5.9604645e-8f * r * r * 1.0001f;
        }
    } else if (a.kind == SCAN_BF16) {
        __bf16* o = static_cast<__bf16*>(a.Qscan) + (int64_t)q * a.kdim;
        float s = 0.f;
        for (int j = threadIdx.x; j < a.kdim; j += 256) {
            __bf16 b = (__bf16)(j < a.d ? x[j] : 0.f);
            o[j] = b;
            float v = (float)b;
            s = fmaf(v, v, s);
        }
        s = block_reduce_sum(s, red);
        if (threadIdx.x == 0) {
            a.qn[q] = s;
            a.qscale[q] = 1.0f;
            a.margin[q] = 0.f;
        }
    } else if (a.kind == SCAN_FP8) {
        float m = 0.f;
        for (int j = threadIdx.x; j < a.d; j += 256) m = fmaxf(m, fabsf(x[j]));
        m = block_reduce_max(m, red);
        // power-of-two scale putting max |q| in [128, 256)
        float aq = m > 0.f ? exp2f(floorf(log2f(255.0f / m))) : 1.0f;
        while (m * aq >= 256.0f) aq *= 0.5f;
        uint8_t* o = static_cast<uint8_t*>(a.Qscan) + (int64_t)q * a.kdim;
        float s = 0.f;
        for (int j = threadIdx.x; j < a.kdim; j += 256) {
            float v = j < a.d ? x[j] : 0.f;
            o[j] = to_e4m3(v * aq);
            s = fmaf(v, v, s);
        }
        s = block_reduce_sum(s, red);
        if (threadIdx.x == 0) {
            a.qn[q] = s;
            a.qscale[q] = 1.0f / (a.alpha_x * aq);
            a.margin[q] = 0.f;
        }
    } else {
        // SQ8 / SQ4 against the reconstructed rows x^ = offset + scale * c:
        // ||x^ - q||^2 = ||x^||^2 + (||q||^2 - 2 q.offset) - 2 sum_d (scale_d q_d) c_d,
        // the last sum as int8 weights w = scale * q with a per-query scale
        // beta (row norm ||x^||^2 from the encoder).
        float m = 0.f, qq = 0.f, qo = 0.f;
        for (int j = threadIdx.x; j < a.d; j += 256) {
            float w = a.scale[j] * x[j];
            m = fmaxf(m, fabsf(w));
            qq = fmaf(x[j], x[j], qq);
            qo = fmaf(x[j], a.offset[j], qo);
        }
        m = block_reduce_max(m, red);
        qq = block_reduce_sum(qq, red);
        qo = block_reduce_sum(qo, red);
        float beta = m > 0.f ? m / 127.0f : 1.0f;
        int8_t* o = static_cast<int8_t*>(a.Qscan) + (int64_t)q * a.kdim;
        float wsum = 0.f;
        for (int p = threadIdx.x; p < a.kdim; p += 256) {
            int j = a.kind == SCAN_SQ4 ? sq4_dim_of_slot(p) : p;
            int wi = 0;
            if (j < a.d) {
                float w = a.scale[j] * x[j];
                wi = (int)rintf(w / beta);
                wi = wi < -127 ? -127 : (wi > 127 ? 127 : wi);
            }
            o[p] = (int8_t)wi;
            wsum += (float)wi;
        }
        wsum = block_reduce_sum(wsum, red);
        if (threadIdx.x == 0) {
            // SQ8 codes enter the MFMA as c - 128: add 128 * sum(w) back
            // through the norm term: -2 beta (acc + 128 sum w).
            const float base = qq - 2.0f * qo;
            a.qn[q] = a.kind == SCAN_SQ8 ? base - 256.0f * beta * wsum : base;
            a.qscale[q] = beta;
            a.margin[q] = 0.f;
        }
    }
}

}  // namespace

hipError_t launch_encode(int kind, const float* src, int64_t sstride, int64_t n, int d, int kdim, void* dst, float* xn,
                         uint32_t* maxkey, float alpha, const float* scale, const float* offset, int variant,
                         hipStream_t st, const uint32_t* gather, int tile_cb) {
    if (gather && kind != SCAN_SQ8) return hipErrorInvalidValue;
    if (tile_cb < 0 || (tile_cb > 0 && (kind == SCAN_F32 || gather))) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    dim3 grid((unsigned)((n + 15) / 16));
    switch (kind) {
        case SCAN_F32:
            hipLaunchKernelGGL(enc_f32, grid, dim3(256), 0, st, src, sstride, n, d, kdim, static_cast<float*>(dst), xn,
                               maxkey);
            break;
        case SCAN_BF16:
            hipLaunchKernelGGL(enc_bf16, grid, dim3(256), 0, st, src, sstride, n, d, kdim, static_cast<__bf16*>(dst),
                               xn, tile_cb);
            break;
        case SCAN_FP8:
            hipLaunchKernelGGL(enc_fp8, grid, dim3(256), 0, st, src, sstride, n, d, kdim, alpha,
                               static_cast<uint8_t*>(dst), xn, tile_cb);
            break;
        case SCAN_SQ8:
            hipLaunchKernelGGL(enc_sq8, grid, dim3(256), 0, st, src, sstride, n, d, kdim, scale, offset, variant,
                               static_cast<uint8_t*>(dst), xn, gather, tile_cb);
            break;
        case SCAN_SQ4:
            hipLaunchKernelGGL(enc_sq4, grid, dim3(256), 0, st, src, sstride, n, d, kdim, scale, offset,
                               static_cast<uint8_t*>(dst), xn, tile_cb);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_dim_minmax(const float* src, int64_t sstride, int64_t n, int d, uint32_t* mn, uint32_t* mx,
                             hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int64_t gx = (n + 4095) / 4096;
    if (gx > 512) gx = 512;
    hipLaunchKernelGGL(dim_minmax, dim3((unsigned)gx, (unsigned)((d + 255) / 256)), dim3(256), 0, st, src, sstride, n,
                       d, mn, mx);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_prep_queries(int kind, const float* Q, int B, int d, int kdim, void* Qscan, float* qn, float* qscale,
                               float* margin, float maxnorm, float alpha_x, const float* scale, const float* offset,
                               hipStream_t st, uint32_t* cnt, uint32_t* flags, float* thr) {
    if (B <= 0) return hipSuccess;
    if ((cnt != nullptr) != (flags != nullptr) || (cnt != nullptr) != (thr != nullptr)) return hipErrorInvalidValue;
    QPrepArgs a{kind, Q, B, d, kdim, Qscan, qn, qscale, margin, maxnorm, alpha_x, scale, offset, cnt, flags, thr};
    ANNMI_LAUNCH(prep_queries, dim3((unsigned)B), dim3(256), 0, st, a);
    return hipGetLastError();
}
