// Shared device/host definitions for the MI355X (gfx950) k-NN search library.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

#define ANNMI_WAVE 64

// Per-launch timing (the library's profiling): a caller arms a start/stop
// event pair for the next ANNMI_LAUNCH on this thread, which then has the
// dispatch itself record them (hipExtLaunchKernelGGL) instead of separate
// event-record packets around the launch, which stalled small steps (one
// query over 1M x 384: 0.50-0.57 ms a step timed that way, 0.44 untimed).
// The launch consumes (clears) the pair.
extern thread_local hipEvent_t t_launch_ev[2];
#define ANNMI_LAUNCH(K, grid, block, lds, st, ...)                                                          \
    do {                                                                                                    \
        if (t_launch_ev[0]) {                                                                               \
            hipExtLaunchKernelGGL(K, grid, block, lds, st, t_launch_ev[0], t_launch_ev[1], 0, __VA_ARGS__); \
            t_launch_ev[0] = t_launch_ev[1] = nullptr;                                                      \
        } else {                                                                                            \
            hipLaunchKernelGGL(K, grid, block, lds, st, __VA_ARGS__);                                       \
        }                                                                                                   \
    } while (0)

// Compute units of the current device, read once per process (launchers on
// several host threads share it; a function-local static is initialised
// exactly once).  256 on MI355X; 256 if the query fails.
inline int device_cu_count() {
    static const int n = [] {
        int dev = 0, c = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
            c = 256;
        return c;
    }();
    return n;
}

// Orderable unsigned image of a float: a < b (as floats, no NaN) iff
// key(a) < key(b) (as unsigned).  Negative approximate distances (possible in
// the ||x||^2 - 2x.q + ||q||^2 form) sort below zero.
__host__ __device__ inline uint32_t f2key(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float key2f(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __builtin_bit_cast(float, u);
}

// Smallest float above f (f finite); +inf stays +inf.
__host__ __device__ inline float next_up(float f) {
    if (!(f < __builtin_inff())) return f;
    return key2f(f2key(f) + 1u);
}

// Tiled scan-row layout (the quantized scan copies): rows are grouped in
// tiles of 16, and a tile is stored chunk-major: the cb bytes of chunk c of
// its 16 rows are contiguous (16 * cb bytes), then chunk c + 1.  One 16-byte
// load per lane of a wave then reads one contiguous 1 KB block (16 rows x
// 64 bytes) instead of 16 half cache lines 1 row apart.  A partial last tile
// occupies a whole tile of memory.  Byte offset of byte b of row `row`:
__host__ __device__ inline int64_t tiled_off(int64_t row, int64_t b, int64_t row_bytes, int cb) {
    return (row >> 4) * 16 * row_bytes + (b / cb) * 16 * cb + (row & 15) * cb + (b % cb);
}

// Scan modes (element type of the candidate-generation pass).
enum annmi_scan_kind : int {
    SCAN_F32 = 0,
    SCAN_BF16 = 1,
    SCAN_FP8 = 2,
    SCAN_SQ8 = 3,
    SCAN_SQ4 = 4,
};

// Per-batch candidate workspace, struct-of-arrays, per query q:
//   cnt[q]                 number of candidates appended (may exceed cap)
//   key[q*cap + j]         approximate distance (float bits) of candidate j
//   rid[q*cap + j]         row id of candidate j
struct CandBuf {
    uint32_t* cnt;
    float* key;
    uint32_t* rid;
    int32_t cap;
};

// Per-query filter: a candidate passes when its approximate distance is
// <= thr[q]; thr = tau + slack, computed by the select kernel.
struct QueryPrep {
    const float* qn;      // [B] squared norm of the query in the scan's element type
    const float* thr;     // [B] current filter threshold
};

// ---------------------------------------------------------------- launches
// All launch helpers return hipSuccess or the launch error; all run on `st`.
struct ScanArgs {
    int kind;                 // annmi_scan_kind
    const void* X;            // base rows in the scan element type (row-major, row stride = xstride bytes)
    int64_t xstride;          // bytes per row
    const float* xn;          // [N] per-row squared norm term (scan element type)
    int64_t row_begin, row_end;
    int d;                    // logical dims
    int kdim;                 // padded contraction length in elements (multiple of the MFMA K chunk)
    const void* Q;            // [B x kdim] queries in the scan element type (row-major)
    const float* qn;          // [B]
    const float* qscale;      // [B] multiplier applied to the raw dot (fp8/sq8/sq4), or nullptr
    const float* thr;         // [B]
    int B;
    CandBuf cand;
    // argmin mode (k-means assignment): when set, no candidates are appended;
    // argmin[q] receives min over rows of (f2key(dist) << 32 | row) via a
    // 64-bit atomicMin (smallest distance, then lowest row).
    unsigned long long* argmin = nullptr;
    // candidate ids are rid_map[row] when set (list-ordered rows -> original ids)
    const uint32_t* rid_map = nullptr;
    // bounds of xn over all rows: min, max of |xn| (the wide scan's
    // prefilter; the defaults make every pair take the exact test), and the
    // prefilter mode: 1 per-row bound, 2 one bound per query from xn_lo
    // (chosen when the row norms are nearly equal), 3 (integer dots) one
    // bound per query on dot minus each row's scaled norm excess
    float xn_lo = -__builtin_inff();
    float xn_amax = __builtin_inff();
    int prefilter = 1;
    // kernel shape (the autotuner's choice; 0 = the built-in rule):
    // wide_min_batch  batches from this size up take the wide kernel
    // wide_cfg        wide-kernel tile configuration (kWideCfg*)
    // chunk_max       most rows per wide-kernel row chunk
    int wide_min_batch = 0;
    int wide_cfg = 0;
    int64_t chunk_max = 0;
    // X in the tiled layout (tiled_off, chunk bytes of the element type)
    // instead of row-major
    int tiled = 0;
};

// Wide-kernel tile configurations: queries per workgroup tile, rows per wave
// step, waves per workgroup, query fragments prefetched, MFMA tile shape.
enum annmi_wide_cfg : int {
    kWideCfgDefault = 0,   // the built-in choice per element type and dim
    kWideCfgQ8R4W4 = 1,    // 128 queries x 64 rows a wave, 4 waves, prefetched query fragments
    kWideCfgQ8R2W8 = 2,    // 128 x 32, 8 waves (two a SIMD)
    kWideCfgQ4R4W8 = 3,    // 64 x 64, 8 waves
    kWideCfg32R1 = 4,      // 32 x 32 MFMA tiles (SQ8, BF16): 128 x 32, 8 waves
    kWideCfg32R2 = 5,      // 32 x 32 MFMA tiles (SQ8, BF16): 128 x 64, 8 waves
    kWideCfg32R4W4 = 6,    // 32 x 32 MFMA tiles: SQ8 128 x 32 on 8 waves, lean epilogue; BF16 128 x 128 on 4 waves
    kWideCfg32R2W4 = 7,    // 32 x 32 MFMA tiles: SQ8 128 x 64 on 8 waves, lean epilogue; BF16 128 x 64 on 4 waves
    kWideCfgRL = 8,        // SQ8, d padded to 128 or 256: rows staged in LDS by direct loads, 512 queries in registers
    kWideCfgCount = 9,
};
// Whether a wide configuration fits the LDS for this element type and dim.
bool wide_cfg_fits(int kind, int kdim, int cfg);
// Launches the scan; *variant (if not null) receives a static name of the
// kernel instantiation launched (for per-kernel timing).
hipError_t launch_scan(const ScanArgs& a, hipStream_t st, const char** variant = nullptr);


struct SelectArgs {
    CandBuf cand;
    int B;
    int rsel;                 // rank to select (k or rerank R)
    const float* slack;       // [B] added to the selected key to form the threshold
    float* thr;               // [B] out: new threshold (unchanged when cnt < rsel)
    uint32_t* overflow;       // [B] out: set to 1 when cnt > cap
};
hipError_t launch_select_threshold(const SelectArgs& a, hipStream_t st);

struct FinalArgs {
    CandBuf cand;
    int B;
    int k;                    // output columns (rows past the candidate count get id -1)
    int rsel;                 // k (exact) or R (over-fetch)
    const float* slack;       // [B] collect slack (2*margin for exact mode, 0 otherwise)
    const float* X32;         // fp32 base rows for the rerank
    int64_t x32_stride;       // floats per row of X32
    int d;
    const float* Q32;         // [B x d] fp32 queries
    int64_t* out_ids;         // [B x k]
    float* out_d;             // [B x k]
    uint32_t* overflow;       // [B] out flags: bit0 cand overflow, bit1 rerank truncated
    uint32_t* n_reranked;     // [B] out: number of rows re-scored in fp32
    double eps_rel;           // FP32 near-tie window (relative) for the FP64 boundary step
};
hipError_t launch_finalize(const FinalArgs& a, hipStream_t st);
// The finalize kernel's function (to find its node in a captured step).
const void* finalize_kernel_fn();
