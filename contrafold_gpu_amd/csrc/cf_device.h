// cf_device.h -- device-side building blocks of the gfx950 fold engine.
//
// Shared by every kernel: the FP32 log-space operators (CONTRAfold's
// Fast_LogPlusEquals / Fast_Exp, restated from gpu_contrafold/cpu.py:527-580),
// the triangular diagonal-major table layout, and the parameter block.
//
// All HIP sources are compiled with -ffp-contract=off: every a*b+c below is a
// rounded multiply followed by a rounded add, exactly like the reference's
// float32 path, which is what makes the FP32 tables bit-identical to it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CF_NEG (-1e30f)
#define CF_HALF (-5e29f)
#define CF_NEG64 (-1e30)
#define CF_HALF64 (-5e29)
#define CF_CMAX 30
#define CF_RING 34  // FCo diagonals kept in LDS: two diagonals written, d+2 .. d+32 read

// Flat parameter-table layout shared with the host (2577 values; the order
// of the dict returned by cpu.load, cpu.py:132-135).
#define CF_O_BP 0
#define CF_O_STACK 25
#define CF_O_TM 650
#define CF_O_HC 1275
#define CF_O_DL 1300
#define CF_O_DR 1425
#define CF_O_HP 1550
#define CF_O_CS 1581
#define CF_O_I11 2542
#define CF_O_B01 2567
#define CF_O_MB 2572
#define CF_O_MU 2573
#define CF_O_MP 2574
#define CF_O_EU 2575
#define CF_O_EP 2576
#define CF_NTAB 2577

// Device parameter block.  `t` holds the 2577 tables in the flat order; JB is
// the derived junction table hc[a][b] + tm[a][b][c][d] (one rounded add, the
// same add the reference performs for every ScoreJunctionB).
template <typename T>
struct CfParams {
    T t[CF_NTAB];
    T JB[625];
};

__host__ __device__ constexpr int cf_i2(int a, int b) { return a * 5 + b; }
__host__ __device__ constexpr int cf_i3(int a, int b, int c) { return (a * 5 + b) * 5 + c; }
__host__ __device__ constexpr int cf_i4(int a, int b, int c, int d) { return ((a * 5 + b) * 5 + c) * 5 + d; }

// canonical pairs AU UA CG GC GU UG: bit (5*a+b) of a 25-bit mask
#define CF_CANON_MASK ((1u << (0 * 5 + 3)) | (1u << (3 * 5 + 0)) | (1u << (1 * 5 + 2)) | (1u << (2 * 5 + 1)) | \
                       (1u << (2 * 5 + 3)) | (1u << (3 * 5 + 2)))
__host__ __device__ inline bool cf_canon(int a, int b) { return (CF_CANON_MASK >> (a * 5 + b)) & 1u; }

// ---------------------------------------------------------------- layout
// Triangular, diagonal-major: cell (i, j), 0 <= i <= j <= L, lives at
// dbase(j - i) + i with dbase(d) = d*(L+1) - d*(d-1)/2.  Lanes that walk a
// diagonal touch consecutive addresses, and so do the FM1-row / FM-column
// reads of the multiloop split at a fixed split offset.
__host__ __device__ inline int64_t cf_dbase(int L, int d) {
    return (int64_t)d * (L + 1) - (int64_t)d * (d - 1) / 2;
}
__host__ __device__ inline int64_t cf_tri(int L, int i, int j) { return cf_dbase(L, j - i) + i; }
// 32-bit form for per-sequence offsets (exact for L <= 32000: at most ~5.1e8)
__host__ __device__ inline int cf_tri32(int L, int i, int j) {
    const int d = j - i;
    return d * (L + 1) - ((d * (d - 1)) >> 1) + i;
}
__host__ __device__ inline int64_t cf_tri_size(int L) { return (int64_t)(L + 1) * (L + 2) / 2; }
// Load base[idx] through a 32-bit byte offset: with base in scalar registers
// the address is one shift and a scalar-base load, no 64-bit lane arithmetic.
// Valid while idx * sizeof(T) < 2^32 (every per-sequence table offset: a
// triangle of L <= 32000 has < 5.2e8 cells).
template <typename T, typename I>
__device__ __forceinline__ T cf_ld(const T* __restrict__ base, I idx) {
    return *(const T*)((const char*)base + (size_t)((unsigned)idx * (unsigned)sizeof(T)));
}

// Exterior (column-major) copy of FC used by the F5 / F5o chains: entry
// (k, j), 0 <= k < j <= L, holds FC[k+1][j-1].
__host__ __device__ inline int64_t cf_xbase(int j) { return (int64_t)j * (j - 1) / 2; }
__host__ __device__ inline int64_t cf_x_size(int L) { return (int64_t)L * (L + 1) / 2; }

// Per-sequence descriptor (host builds it; one per sequence of a batch).
struct CfSeq {
    int32_t L;
    int32_t coff;   // offset of s[0] in codes / masks (L+2 entries), also F5/F5o/up arrays
    int64_t toff;   // offset of cell (0,0) in every triangular table
    int64_t xoff;   // offset of the exterior FC copy
    int64_t moff;   // offset of cell (0,0)'s tile in one gamma's tiled MEA score table
};

// ---------------------------------------------------------------- LSE
// CONTRAfold Fast_LogExpPlusOne segments for RealT=float (cpu.py:538-556):
// r(d) = ((c0*d + c1)*d + c2)*d + c3 on 8 consecutive d-intervals.
struct CfLseTab {
    float4 c[8];
};
__constant__ const float cf_lse_coef[8][4] = {
    {-0.0065591595f, 0.1276442762f, 0.4996554598f, 0.6931542306f},
    {-0.0155157557f, 0.1446775699f, 0.4882939746f, 0.6958092989f},
    {-0.0128909247f, 0.1301028251f, 0.5150398748f, 0.6795585882f},
    {-0.0072142647f, 0.0877540853f, 0.6208708362f, 0.5909675829f},
    {-0.0031455354f, 0.0467229449f, 0.7592532310f, 0.4348794399f},
    {-0.0010110698f, 0.0185943421f, 0.8831730747f, 0.2523695427f},
    {-0.0001962780f, 0.0046084408f, 0.9634431978f, 0.0983148903f},
    {-0.0000113994f, 0.0003734731f, 0.9959107193f, 0.0149855051f},
};

// The 7 interior breakpoints of those segments: segment k holds
// bp[k-1] <= d < bp[k] (the reference's `if (x < bp)` tests).
__constant__ const float cf_lse_bp[7] = {0.6615367791f, 1.6320158198f, 2.4912588184f, 3.3792499610f,
                                         4.4261691294f, 5.7890710412f, 7.8162726752f};
// Segment lookup by quarter-unit bins of d: bin b covers [b/4, (b+1)/4) and
// holds at most one breakpoint (they are more than 1/4 apart).  Entry b =
// {breakpoint in the bin as float bits (or +inf), byte offset of the segment
// below it}; the segment from the breakpoint on is the next one.  Bin 32
// stands for every d >= 8 (the last segment).  Staged in LDS right after the
// 8 coefficient rows (cf_lse_stage); 8-byte entries keep the sweeps' LDS plan
// within six workgroups per CU at 200 nt.
#define CF_LSE_BINS 33
__device__ inline void cf_lse_stage(float4* C, int t) {
    if (t < 8) C[t] = make_float4(cf_lse_coef[t][0], cf_lse_coef[t][1], cf_lse_coef[t][2], cf_lse_coef[t][3]);
    if (t < CF_LSE_BINS) {
        const float lo = 0.25f * (float)t, hi = 0.25f * (float)(t + 1);
        int below = 0;
        float in_bin = __int_as_float(0x7f800000);  // +inf: no breakpoint in this bin
        for (int k = 0; k < 7; ++k) {
            if (cf_lse_bp[k] < lo) ++below;
            else if (cf_lse_bp[k] < hi) in_bin = cf_lse_bp[k];
        }
        if (t == CF_LSE_BINS - 1) in_bin = __int_as_float(0x7f800000);
        ((int2*)(C + 8))[t] = make_int2(__float_as_int(in_bin), 16 * below);
    }
}
// LDS bytes of the staged table (coefficients + bins)
#define CF_LSE_TABLE_BYTES (8 * 16 + ((CF_LSE_BINS * 8 + 15) / 16) * 16)

// log(exp(a)+exp(b)), float32, branch-free.  Same result bits as the
// reference's branchy lse32 (cpu.py:527-557): the interval of d is found from
// its quarter-unit bin (d * 4 is exact) and one compare with the bin's
// breakpoint, the cubic's coefficients come from a table in LDS, and the two
// early exits become one select.
__device__ __forceinline__ float cf_lse(float a, float b, const float4* __restrict__ C) {
    const float hi = fmaxf(a, b);
    const float lo = fminf(a, b);
    const float d = hi - lo;
    const int bin = (int)fminf(d * 4.0f, (float)(CF_LSE_BINS - 1));
    const int2 e = ((const int2*)(C + 8))[bin];
    const int off = e.y + ((d >= __int_as_float(e.x)) ? 16 : 0);
    const float4 c = *(const float4*)((const char*)C + off);
    float r = c.x * d;
    r = r + c.y;
    r = r * d;
    r = r + c.z;
    r = r * d;
    r = r + c.w;
    const float res = lo + r;
    // The reference also returns hi when lo < -5e29.  Values below that are
    // exactly NEG = -1e30 (scores never come near): then either hi is finite
    // and d >= 11.86 anyway, or hi = lo = NEG, d = 0 and lo + r rounds back to
    // NEG (|r| < 1 is far below half an ulp of 1e30), so one test suffices.
    return (d >= 11.8624794162f) ? hi : res;
}
// Skipped terms.  The reference skips a term when one of its table parts is
// not finite (below -5e29).  Every such part is exactly NEG, and a term is
// that part plus finite scores or plus another NEG, so the term itself is
// -1e30 or -2e30 (the finite addends vanish in rounding).  cf_lse(acc, t)
// with t <= -1e30 returns acc bit for bit: for finite acc, d >= 11.86 and hi
// = acc is returned; for acc = NEG either d >= 11.86 again or d = 0 and NEG +
// r rounds back to NEG.  So the sweeps fold such terms unconditionally, with
// no compare or select per term.

// CONTRAfold Fast_Exp for RealT=float (cpu.py:560-580), branch-free.  For
// x >= 0 the reference calls expf (>= 1); the only consumer here is the
// posterior, a sum of non-negative terms clamped to 1, which any term >= 1
// sends to exactly 1, so 1 stands in for that tail and the result bits are the
// same.
__constant__ const float cf_exp_coef[6][4] = {
    {0.0000803850f, 0.0021627428f, 0.0194708555f, 0.0588080014f},
    {0.0013889414f, 0.0244676474f, 0.1471290604f, 0.3042757740f},
    {0.0072335607f, 0.0906002677f, 0.3983111356f, 0.6245959221f},
    {0.0232410351f, 0.2085645908f, 0.6906367911f, 0.8682322329f},
    {0.0573782771f, 0.3580258429f, 0.9121133217f, 0.9793091728f},
    {0.1199175927f, 0.4815668234f, 0.9975991939f, 0.9999505077f},
};
__device__ __forceinline__ float cf_fexp(float x, const float4* __restrict__ E) {
    // seg = number of breakpoints <= x, by a two-level tree on the middle one
    const bool c3 = x >= -2.4915033807f;
    const bool c2 = x >= (c3 ? -1.4805375919f : -5.8622823336f);
    const bool c1 = x >= (c3 ? -0.6725053211f : -3.8396630909f);
    const int off = (c3 ? 48 : 0) + (c2 ? 16 : 0) + (c1 ? 16 : 0);  // byte offset of E[seg]
    const float4 c = *(const float4*)((const char*)E + off);
    float r = c.x * x;
    r = r + c.y;
    r = r * x;
    r = r + c.z;
    r = r * x;
    r = r + c.w;
    if (x < -9.91152f) r = 0.0f;
    if (x >= 0.0f) r = 1.0f;
    return r;
}

// ---------------------------------------------------------------- bitmasks
// Partner masks: bit q of row x (x = base code of the left position, or 4 if
// it is N/forced) says whether position q may pair with base x.  A window of
// up to 31 consecutive q is cut out with one 64-bit shift.
__device__ __forceinline__ uint32_t cf_window(const uint32_t* __restrict__ row, int qlo, int qhi) {
    if (qhi < qlo) return 0u;
    const int w = qlo >> 5, sh = qlo & 31;
    const uint64_t v = ((uint64_t)row[w + 1] << 32) | (uint64_t)row[w];
    const uint32_t bits = (uint32_t)(v >> sh);
    const int n = qhi - qlo;  // <= 30
    return bits & ((2u << n) - 1u);
}
