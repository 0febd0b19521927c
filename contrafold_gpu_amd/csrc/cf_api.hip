// cf_api.hip -- host side of the fold engine: context, device memory, batch
// packing, kernel launches, result readback, and the extern "C" boundary
// declared in include/contrafold_amd.h.
#include "cf_kernels.hip"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "contrafold_amd.h"

#define CF_VERSION 1

#include "cf_internal.h"

namespace cfi {

int fail(cf_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
int hipfail(cf_ctx* c, hipError_t e, const char* what) {
    return fail(c, CF_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define CF_HIP(ctx, call, what)                                \
    do {                                                       \
        hipError_t _e = (call);                                \
        if (_e != hipSuccess) return hipfail((ctx), _e, what); \
    } while (0)

inline uint8_t encode_base(char ch) {
    switch (ch) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'U': case 'u': case 'T': case 't': return 3;
        default: return 4;
    }
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Launch order: longest first, and among equal lengths the folds with the
// most possible base pairs first (the A-U, C-G and G-U products of the free
// bases' counts, a work estimate), so the fold queues hand out the lightest
// folds last and the sweeps drain sooner (10,000 x 200 nt, 10% masked:
// inside 52.2 -> 49.3 ms, outside 71.0 -> 68.3 ms).
// The counts are four 16-bit fields of one 64-bit sum (A, C, G, U/T; a
// sequence has at most 32,000 bases), added from a per-character table with
// the mask applied as a bit mask: no branch and no store per base (a counter
// array indexed by base, or a branch on the mask, is several times slower).
struct BaseIncrements {
    uint64_t v[256];
    BaseIncrements() {
        for (int ch = 0; ch < 256; ++ch) {
            const uint8_t b = encode_base((char)ch);
            v[ch] = b < 4 ? 1ull << (16 * b) : 0ull;
        }
    }
};
const BaseIncrements kBaseInc;

void launch_order(int n, const std::vector<int>& L, const char* seqs, const int64_t* offsets,
                  const uint8_t* masks, std::vector<int>& order) {
    std::vector<int64_t> pairs((size_t)n);
    for (int i = 0; i < n; ++i) {
        const uint8_t* s = (const uint8_t*)seqs + offsets[i];
        const int64_t l = offsets[i + 1] - offsets[i];
        uint64_t acc = 0;
        if (masks) {
            const uint8_t* m = masks + offsets[i];
            for (int64_t q = 0; q < l; ++q) acc += kBaseInc.v[s[q]] & (0ull - (uint64_t)(m[q] == 0));
        } else {
            for (int64_t q = 0; q < l; ++q) acc += kBaseInc.v[s[q]];
        }
        const int64_t a = (int64_t)(acc & 0xffff), c = (int64_t)((acc >> 16) & 0xffff);
        const int64_t g = (int64_t)((acc >> 32) & 0xffff), u = (int64_t)(acc >> 48);
        pairs[i] = a * u + c * g + g * u;
    }
    order.resize(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return L[x] != L[y] ? L[x] > L[y] : pairs[x] > pairs[y]; });
}

// carve consecutive sub-buffers out of one region
struct Carver {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(int64_t count) {
        off = align_up(off, 256);
        T* p = (T*)(base + off);
        off += (size_t)std::max<int64_t>(count, 1) * sizeof(T);
        return p;
    }
};
template <typename F>
size_t measure(F f) {
    Carver c{nullptr};
    f(c);
    return align_up(c.off, 256);
}

// `gen` counts pointer changes of the regions of one batch: a captured graph
// holds the addresses it was recorded with and is replayed only while the
// count is unchanged.
int region_get(cf_ctx* c, Region& r, size_t n, uint64_t* gen = nullptr) {
    if (r.p && r.n >= n) return CF_OK;
    if (r.p) c->pool.put(r.p, r.n);
    r = Region();
    if (gen) ++*gen;
    void* p = nullptr;
    size_t got = 0;
    hipError_t e = c->pool.get(&p, n, &got);
    if (e != hipSuccess) return fail(c, CF_ENOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
    r.p = p;
    r.n = got;
    return CF_OK;
}
void region_put(cf_ctx* c, Region& r) {
    if (r.p) c->pool.put(r.p, r.n);
    r = Region();
}

// Calibration copy for hardware traffic counters: each thread moves 16 bytes
// per step, read once and written once.
__global__ void cf_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void cf_fill_u32(unsigned* p, int64_t n, unsigned v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Zero the barrier counters of a split launch (a fill kernel, not a memset,
// so the run stays capturable as a hipGraph).
void clear_xsync(cf_batch* b, const cf_batch::Bucket& bk, hipStream_t st) {
    const int64_t n = (int64_t)(bk.end - bk.begin) * 16;
    hipLaunchKernelGGL(cf_fill_u32, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0, st,
                       b->d_xsync + (size_t)bk.begin * 16, n, 0u);
}

// Persistent packed launch: workgroups the device holds at once (at most
// one per G folds), and the fold queue's counter cleared in stream order.
// `spare` workgroups per CU are left to a kernel running beside it, when
// the launch keeps at least twice as many (a sweep with few resident
// workgroups per CU, long folds, keeps them all).
template <typename Kern>
int persistent_grid(cf_ctx* c, Kern kern, int G, size_t lds, int nfold, int spare = 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * G, lds) != hipSuccess || per_cu <= 0) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    if (spare > 0 && per_cu >= 3 * spare) per_cu -= spare;
    return (int)std::max<int64_t>(1, std::min<int64_t>((nfold + G - 1) / G, (int64_t)per_cu * c->n_cu));
}
int* clear_queue(cf_batch* b, const cf_batch::Bucket& bk, int stage, hipStream_t st) {
    int* q = b->d_queue + (size_t)bk.begin * 4 + stage;
    hipLaunchKernelGGL(cf_fill_u32, dim3(1), dim3(64), 0, st, (unsigned*)q, (int64_t)1, 0u);
    return q;
}

int vit_lds(int Lmax) { return cf_vit_plan(Lmax).total; }

// ---------------------------------------------------------------- stages
// A bucket whose tuned width is one wave (64 threads) runs packed: several
// sequences per workgroup, one wave each, sharing the staged parameter
// tables, and the workgroups take folds from a queue until it is empty.
// Wider buckets give the whole workgroup to one sequence.  Sequences per
// workgroup of the packed inside / outside sweeps (fewer when LDS is short),
// alone, at 10,000 x 200 nt (inside / outside ms): 1 per workgroup
// 90.7 / 120.3, 2: 65.4 / 87.1, 4: 58.3 / 78.8, 8: 57.0 / 77.0 --
// default packing (cf_ctx::pack_in / pack_out): eight one-wave teams per
// workgroup keep 32 teams on a CU (LDS-limited) against 24, and more resident
// folds hide more of the split reads' latency.  But the outside sweep shares
// the GPU with the sampler in a logZ + bpp + samples run, and there its
// 512-thread, 40 KB workgroups found CU room less often: 80.7 -> 104.7 ms and
// 58.2k -> 53.3k folds/s.  So the inside packs eight and the outside four.
constexpr int kLanes = 4;  // streams a multi-bucket batch forks onto

// Workgroup width of the kernels that have no packed layout (MEA, Viterbi,
// gradient): a packed bucket (width 64) gives them a whole workgroup per
// sequence, sized to the bucket's longest sequence.
int solo_threads(const cf_batch::Bucket& bk) {
    if (bk.threads > 64) return std::min(bk.threads, 256);
    return bk.Lmax <= 96 ? 128 : 256;
}

using Bucket = cf_batch::Bucket;

// ---- per-bucket launches: every kernel of one length bucket goes to stream `st`

int launch_inside(cf_batch* b, const Bucket& bk, hipStream_t st) {
    cf_ctx* c = b->ctx;
    const int nl = bk.end - bk.begin;
    if (nl <= 0) return CF_OK;
    // The inside sweep reads its single-branch window from the global FC
    // table: without an LDS ring of FC diagonals a packed workgroup needs
    // ~26 KB of LDS at 200 nt, and more resident waves per CU beat the ring
    // on throughput (54 vs 87 ms for 3000 x 200 nt).
    const bool pack = bk.threads <= 64;
    // packed: as many sequences per workgroup (up to pack_in) as LDS allows
    int G = pack ? c->pack_in : 1;
    CfLds plan = cf_lds_plan(bk.Lmax, false, 0, G);
    while (plan.total > c->lds_max && G > 1) plan = cf_lds_plan(bk.Lmax, false, 0, G >>= 1);
    if (plan.total > c->lds_max) return fail(c, CF_EINVAL, "sequence too long for the inside kernel's LDS plan");
    const int K = pack ? 1 : bk.split;
    if (K > 1) clear_xsync(b, bk, st);
    if (pack) {
        int* q = clear_queue(b, bk, 0, st);
        const int grid = persistent_grid(c, cf_inside_kernel<true, false>, G, plan.total, nl);
        hipLaunchKernelGGL((cf_inside_kernel<true, false>), dim3(grid), dim3(64 * G), plan.total, st, b->d_seqs,
                           b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32, b->FC, b->FM, b->FM1, b->FM2,
                           b->FCX, bk.Lmax, 1, (unsigned*)nullptr, b->d_status, q);
    }
    else if (K > 1)
        hipLaunchKernelGGL((cf_inside_kernel<false, true>), dim3(cf_split_grid(nl, K)), dim3(bk.threads), plan.total,
                           st, b->d_seqs, b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32, b->FC, b->FM,
                           b->FM1, b->FM2, b->FCX, bk.Lmax, K, b->d_xsync + (size_t)bk.begin * 16, b->d_status,
                           (int*)nullptr);
    else
        hipLaunchKernelGGL((cf_inside_kernel<false, false>), dim3(nl), dim3(bk.threads), plan.total, st, b->d_seqs,
                           b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32, b->FC, b->FM, b->FM1, b->FM2,
                           b->FCX, bk.Lmax, 1, (unsigned*)nullptr, b->d_status, (int*)nullptr);
    CF_HIP(c, hipGetLastError(), "inside launch");
    return CF_OK;
}

int launch_exterior(cf_batch* b, const Bucket& bk, int mode, hipStream_t st) {
    cf_ctx* c = b->ctx;
    const int nl = bk.end - bk.begin;
    if (nl <= 0) return CF_OK;
    // one wave per (sequence, chain); up to four waves per workgroup as LDS allows
    const int per = cf_ext_wave_bytes(bk.Lmax);
    const int avail = c->lds_max - 4096;  // static tables (LSE coefficients and bins, junction terms: 2.1 KB)
    if (per > avail) return fail(c, CF_EINVAL, "sequence too long for the exterior kernel");
    const int nw = std::max(1, std::min(4, avail / per));
    const int tasks = nl * (mode == 3 ? 2 : 1);
    hipLaunchKernelGGL(cf_exterior_kernel, dim3((tasks + nw - 1) / nw), dim3(64 * nw), nw * per, st, b->d_seqs,
                       b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32, b->FCX, b->F5, b->F5o, b->logz,
                       mode, bk.Lmax);
    CF_HIP(c, hipGetLastError(), "exterior launch");
    return CF_OK;
}

// beside_sampler: the sampler runs on a lane stream beside this launch.  A
// packed sweep then leaves about two waves per SIMD to it (eight per CU):
// the sampler's waves use the sweep's idle issue slots better than the
// sweep's own last two waves do (10,000 x 200 nt, bench step with the
// sampler: outside 70.1 -> 76.2 ms, sampler 92.7 -> 81.7 ms, step 152.1 ->
// 147.9 ms; leaving one wave per SIMD: 151.5 ms; three: 157.9 ms against
// 146.6 for two on another GPU).  At 100, 300 and 380 nt (30,000, 4,000 and
// 2,000 folds) the step time is within 1% either way.
int launch_outside(cf_batch* b, const Bucket& bk, hipStream_t st, bool beside_sampler = false) {
    cf_ctx* c = b->ctx;
    const int nl = bk.end - bk.begin;
    if (nl <= 0) return CF_OK;
    if (bk.threads <= 64) {
        // packed: no room for a per-sequence FCo ring, the window is read from L2
        int G = c->pack_out;
        CfLds plan = cf_lds_plan(bk.Lmax, false, 0, G);
        while (plan.total > c->lds_max && G > 1) plan = cf_lds_plan(bk.Lmax, false, 0, G >>= 1);
        if (plan.total > c->lds_max) return fail(c, CF_EINVAL, "sequence too long for the outside kernel's LDS plan");
        int* q = clear_queue(b, bk, 1, st);
        const int grid = persistent_grid(c, cf_outside_kernel<false, true, false>, G, plan.total, nl,
                                         beside_sampler ? (8 + G - 1) / G : 0);
        hipLaunchKernelGGL((cf_outside_kernel<false, true, false>), dim3(grid), dim3(64 * G),
                           plan.total, st, b->d_seqs, b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32,
                           b->FC, b->FM, b->FM1, b->F5, b->F5o, b->FCo, b->FMo, b->FM1o, b->FM2o, bk.Lmax, 1,
                           (unsigned*)nullptr, b->d_status, q);
        CF_HIP(c, hipGetLastError(), "outside launch");
        return CF_OK;
    }
    const int K = bk.split;
    if (K > 1) {
        // a split team reads every FCo diagonal from the global table (no ring)
        CfLds plan = cf_lds_plan(bk.Lmax, false, 0);
        if (plan.total > c->lds_max) return fail(c, CF_EINVAL, "sequence too long for the outside kernel's LDS plan");
        clear_xsync(b, bk, st);
        hipLaunchKernelGGL((cf_outside_kernel<false, false, true>), dim3(cf_split_grid(nl, K)), dim3(bk.threads),
                           plan.total, st, b->d_seqs, b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32,
                           b->FC, b->FM, b->FM1, b->F5, b->F5o, b->FCo, b->FMo, b->FM1o, b->FM2o, bk.Lmax, K,
                           b->d_xsync + (size_t)bk.begin * 16, b->d_status, (int*)nullptr);
        CF_HIP(c, hipGetLastError(), "outside launch");
        return CF_OK;
    }
    CfLds plan = cf_lds_plan(bk.Lmax, true, 0);
    const bool ring = plan.total <= c->lds_max;
    if (!ring) plan = cf_lds_plan(bk.Lmax, false, 0);
    if (plan.total > c->lds_max) return fail(c, CF_EINVAL, "sequence too long for the outside kernel's LDS plan");
    if (ring)
        hipLaunchKernelGGL((cf_outside_kernel<true, false, false>), dim3(nl), dim3(bk.threads), plan.total, st, b->d_seqs,
                           b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32, b->FC, b->FM, b->FM1, b->F5,
                           b->F5o, b->FCo, b->FMo, b->FM1o, b->FM2o, bk.Lmax, 1, (unsigned*)nullptr, b->d_status,
                           (int*)nullptr);
    else
        hipLaunchKernelGGL((cf_outside_kernel<false, false, false>), dim3(nl), dim3(bk.threads), plan.total, st, b->d_seqs,
                           b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32, b->FC, b->FM, b->FM1, b->F5,
                           b->F5o, b->FCo, b->FMo, b->FM1o, b->FM2o, bk.Lmax, 1, (unsigned*)nullptr, b->d_status,
                           (int*)nullptr);
    CF_HIP(c, hipGetLastError(), "outside launch");
    return CF_OK;
}

int launch_post(cf_batch* b, const Bucket& bk, hipStream_t st) {
    cf_ctx* c = b->ctx;
    const int nl = bk.end - bk.begin;
    if (nl <= 0) return CF_OK;
    // The waves of a workgroup split one sequence's diagonals: the kernel
    // waits on global FCo reads, which more waves per sequence hide better.
    // Eight per sequence up to 256 nt (10,000 x 200 nt: 26.0 ms, against 28.1
    // for four and 41 for two); four above, where a workgroup's per-wave pools
    // and the longer sequence staging would hold more LDS than they hide
    // latency (2,000 x 500 nt: 55.8 against 59.4 ms; 20,000 x U[30, 500] nt:
    // 135 against 155 ms).  A wide bucket's workgroup width (up to 16 waves)
    // splits long folds further, as far as the per-wave LDS areas fit.
    if (bk.Lmax >= 8192) return fail(c, CF_EINVAL, "sequence too long for the posterior kernel (8191 nt at most)");
    int split = bk.threads > 64 ? std::max(8, bk.threads / 64) : (bk.Lmax > 256 ? 4 : 8);
    const size_t seq_lds = cf_lds_plan(bk.Lmax, false, 0, 1).total;
    auto lds = [&](int w) { return seq_lds + (size_t)w * CF_POST_WORK; };
    while (lds(split) > (size_t)c->lds_max && split > 1) --split;
    if (lds(split) > (size_t)c->lds_max)
        return fail(c, CF_EINVAL, "sequence too long for the posterior kernel's LDS plan");
    // a split team's folds take K workgroups here too (no waits: cells are independent)
    const int K = bk.threads > 64 ? bk.split : 1;
    hipLaunchKernelGGL(cf_post_kernel, dim3(K > 1 ? cf_split_grid(nl, K) : nl), dim3(64 * split), lds(split), st,
                       b->d_seqs, b->d_order + bk.begin, nl, b->d_codes, b->d_masks, c->dP32, b->FC, b->F5, b->F5o,
                       b->FCo, b->FM1o, b->POST, bk.Lmax, K);
    CF_HIP(c, hipGetLastError(), "posterior launch");
    return CF_OK;
}

int launch_mea(cf_batch* b, const Bucket& bk, hipStream_t st) {
    cf_ctx* c = b->ctx;
    const int nl = bk.end - bk.begin;
    if (nl <= 0) return CF_OK;
    if (bk.threads <= 64) {
        // packed: one wave per (fold, gamma), eight per workgroup (fewer when
        // LDS is short), taking items from the bucket's MEA queue
        const int team = cf_mea_plan(bk.Lmax, 1).team;
        int G = 8;
        while (G > 1 && G * team > c->lds_max) G >>= 1;
        if (team > c->lds_max) return fail(c, CF_EINVAL, "sequence too long for the MEA kernel");
        int* q = clear_queue(b, bk, 2, st);
        const int grid = persistent_grid(c, cf_mea_kernel<true>, G, (size_t)G * team, nl * b->n_gamma);
        hipLaunchKernelGGL(cf_mea_kernel<true>, dim3(grid), dim3(64 * G), (size_t)G * team, st, b->d_seqs,
                           b->d_order + bk.begin, nl, b->n_gamma, b->d_codes, b->d_masks, b->POST, b->d_gin, b->SC,
                           b->total_M, b->mea_out, b->d_charoff, b->total_chars, bk.Lmax, q);
    } else {
        // wide: the workgroup (up to 8 waves) decodes one (fold, gamma)
        const int nw = std::min(8, bk.threads / 64);
        const int lds = cf_mea_plan(bk.Lmax, nw).team;
        if (lds > c->lds_max) return fail(c, CF_EINVAL, "sequence too long for the MEA kernel");
        hipLaunchKernelGGL(cf_mea_kernel<false>, dim3(nl * b->n_gamma), dim3(64 * nw), lds, st, b->d_seqs,
                           b->d_order + bk.begin, nl, b->n_gamma, b->d_codes, b->d_masks, b->POST, b->d_gin, b->SC,
                           b->total_M, b->mea_out, b->d_charoff, b->total_chars, bk.Lmax, (int*)nullptr);
    }
    CF_HIP(c, hipGetLastError(), "mea launch");
    return CF_OK;
}

int launch_viterbi(cf_batch* b, const Bucket& bk, bool lse, hipStream_t st) {
    cf_ctx* c = b->ctx;
    const int nl = bk.end - bk.begin;
    if (nl <= 0) return CF_OK;
    const int lds = vit_lds(bk.Lmax);
    if (lds > c->lds_max) return fail(c, CF_EINVAL, "sequence too long for the Viterbi kernel");
    if (lse)
        hipLaunchKernelGGL(cf_viterbi_kernel<true>, dim3(nl), dim3(solo_threads(bk)), lds, st, b->d_seqs,
                           b->d_order + bk.begin, b->d_codes, b->d_masks, c->dP64, b->VFC, b->VFM, b->VFM1, b->logz64,
                           b->mfe_out, b->d_charoff, bk.Lmax);
    else
        hipLaunchKernelGGL(cf_viterbi_kernel<false>, dim3(nl), dim3(solo_threads(bk)), lds, st, b->d_seqs,
                           b->d_order + bk.begin, b->d_codes, b->d_masks, c->dP64, b->VFC, b->VFM, b->VFM1, b->vscore,
                           b->mfe_out, b->d_charoff, bk.Lmax);
    CF_HIP(c, hipGetLastError(), "viterbi launch");
    return CF_OK;
}

// ---- whole-batch launches

// The sampler's per-wave LDS slice holds the F5 row, codes and mask as well
// as the stack top when four 512-thread workgroups still fit a CU with it.
bool samp_local(int Lmax, bool masks) {
    const size_t fixed = sizeof(CfParams<float>) + 2 * 496;
    return 4 * (fixed + (size_t)CF_SAMP_WAVES * cf_samp_pitch(Lmax, true, masks)) <= 160 * 1024;
}

int launch_sample(cf_batch* b, uint64_t seed, hipStream_t st) {
    cf_ctx* c = b->ctx;
    const bool masks = b->d_masks != nullptr;
    const bool local = samp_local(b->Lmax, masks);
    const size_t lds = (size_t)CF_SAMP_WAVES * cf_samp_pitch(b->Lmax, local, masks);
    // the task counter: the first bucket's fourth queue slot (the sweeps use
    // slots 0-2), cleared in stream order.  Waves take four tasks at a time
    // when there are many per wave, one when there are few (100 x 200 nt x
    // 100 samples: chunks of four left most waves idle and the rest four
    // samples deep, 0.3 -> 0.7 ms).
    int* q = b->d_queue + 3;
    hipLaunchKernelGGL(cf_fill_u32, dim3(1), dim3(64), 0, st, (unsigned*)q, (int64_t)1, 0u);
    const int64_t tasks = (int64_t)b->n * b->n_samples;
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(4, tasks / ((int64_t)b->sgrid * CF_SAMP_WAVES * 8)));
    if (local)
        hipLaunchKernelGGL(cf_sample_kernel<true>, dim3(b->sgrid), dim3(CF_SAMP_WAVES * WAVE), lds, st, b->d_seqs,
                           b->d_order, b->n, b->n_samples, seed, b->d_keys, b->d_codes, b->d_masks, c->dP32, b->FC,
                           b->FM, b->FM1, b->FM2, b->F5, b->stack, b->stack_cap, b->samp_out, b->d_soff, b->Lmax,
                           b->d_status, q, chunk);
    else
        hipLaunchKernelGGL(cf_sample_kernel<false>, dim3(b->sgrid), dim3(CF_SAMP_WAVES * WAVE), lds, st, b->d_seqs,
                           b->d_order, b->n, b->n_samples, seed, b->d_keys, b->d_codes, b->d_masks, c->dP32, b->FC,
                           b->FM, b->FM1, b->FM2, b->F5, b->stack, b->stack_cap, b->samp_out, b->d_soff, b->Lmax,
                           b->d_status, q, chunk);
    CF_HIP(c, hipGetLastError(), "sample launch");
    return CF_OK;
}

// Per-sequence expected counts (one row each), then their weighted sum in
// sequence order: the gradient is the same bit for bit whatever the buckets,
// team widths or scheduling.
int launch_grad(cf_batch* b, hipStream_t st) {
    cf_ctx* c = b->ctx;
    for (const auto& bk : b->buckets) {
        const int nl = bk.end - bk.begin;
        if (nl <= 0) continue;
        hipLaunchKernelGGL(cf_grad_kernel, dim3(nl), dim3(solo_threads(bk)), 0, st, b->d_seqs, b->d_order + bk.begin,
                           b->d_codes, b->d_masks, c->dP32, b->FC, b->FM, b->FM1, b->FM2, b->F5, b->FCo, b->FMo,
                           b->FM1o, b->F5o, b->seqcnt);
        CF_HIP(c, hipGetLastError(), "grad launch");
    }
    hipLaunchKernelGGL(cf_grad_reduce_kernel, dim3((CF_NTAB + 255) / 256), dim3(256), 0, st, b->seqcnt, b->n,
                       (const float*)(b->d_gin + kMaxGamma), b->grad);
    CF_HIP(c, hipGetLastError(), "grad reduce launch");
    return CF_OK;
}

// Allocate whatever the requested stages need.
int ensure_buffers(cf_batch* b, unsigned flags, int n_gamma, int n_samples) {
    cf_ctx* c = b->ctx;
    int rc;
    if (flags & (CF_BPP | CF_MEA | CF_GRAD)) {
        if (!b->r_out.p) {
            size_t need = measure([&](Carver& k) {
                for (int t = 0; t < 5; ++t) k.take<float>(b->total_T);
            });
            if ((rc = region_get(c, b->r_out, need, &b->buf_gen))) return rc;
            Carver k{(char*)b->r_out.p};
            b->FCo = k.take<float>(b->total_T);
            b->FMo = k.take<float>(b->total_T);
            b->FM1o = k.take<float>(b->total_T);
            b->FM2o = k.take<float>(b->total_T);
            b->POST = k.take<float>(b->total_T);
        }
    }
    if (flags & CF_MEA) {
        if (n_gamma > b->mea_cap) {
            region_put(c, b->r_mea);
            auto lay = [&](Carver& k) {
                k.take<float>((int64_t)n_gamma * b->total_M);
                k.take<uint8_t>((int64_t)n_gamma * b->total_chars);
            };
            if ((rc = region_get(c, b->r_mea, measure(lay), &b->buf_gen))) return rc;
            Carver k{(char*)b->r_mea.p};
            b->SC = k.take<float>((int64_t)n_gamma * b->total_M);
            b->mea_out = k.take<uint8_t>((int64_t)n_gamma * b->total_chars);
            b->mea_cap = n_gamma;
        }
    }
    if ((flags & (CF_MFE | CF_LOGZ64)) && !b->r_vit.p) {
        auto lay = [&](Carver& k) {
            for (int t = 0; t < 3; ++t) k.take<double>(b->total_T);
            k.take<double>(b->n);
            k.take<double>(b->n);
            k.take<uint8_t>(b->total_chars);
        };
        if ((rc = region_get(c, b->r_vit, measure(lay), &b->buf_gen))) return rc;
        Carver k{(char*)b->r_vit.p};
        b->VFC = k.take<double>(b->total_T);
        b->VFM = k.take<double>(b->total_T);
        b->VFM1 = k.take<double>(b->total_T);
        b->vscore = k.take<double>(b->n);
        b->logz64 = k.take<double>(b->n);
        b->mfe_out = k.take<uint8_t>(b->total_chars);
    }
    if (flags & CF_SAMPLE) {
        const int64_t tasks = (int64_t)b->n * n_samples;
        // one wave per (sequence, sample), eight waves per workgroup, a persistent
        // grid of up to 4 workgroups per CU; each wave owns a traceback stack
        const int sgrid = (int)std::max<int64_t>(
            1, std::min<int64_t>((tasks + CF_SAMP_WAVES - 1) / CF_SAMP_WAVES, (int64_t)c->n_cu * 4));
        const int cap = 2 * b->Lmax + 8;
        if (n_samples > b->smp_cap || sgrid > b->sgrid || cap > b->stack_cap) {
            region_put(c, b->r_smp);
            const int ns_cap = std::max(n_samples, b->smp_cap);
            const int g_cap = std::max(sgrid, b->sgrid);
            auto lay = [&](Carver& k) {
                k.take<uint8_t>((int64_t)ns_cap * b->total_chars);
                k.take<uint32_t>((int64_t)g_cap * CF_SAMP_WAVES * cap);
            };
            if ((rc = region_get(c, b->r_smp, measure(lay), &b->buf_gen))) return rc;
            Carver k{(char*)b->r_smp.p};
            b->samp_out = k.take<uint8_t>((int64_t)ns_cap * b->total_chars);
            b->stack = k.take<uint32_t>((int64_t)g_cap * CF_SAMP_WAVES * cap);
            b->smp_cap = ns_cap;
            b->stack_cap = cap;
        }
        b->sgrid = sgrid;
    }
    if ((flags & CF_GRAD) && !b->r_grad.p) {
        auto lay = [&](Carver& k) {
            k.take<double>(CF_NTAB);
            k.take<double>((int64_t)b->n * CF_NTAB);
        };
        if ((rc = region_get(c, b->r_grad, measure(lay), &b->buf_gen))) return rc;
        Carver k{(char*)b->r_grad.p};
        b->grad = k.take<double>(CF_NTAB);
        b->seqcnt = k.take<double>((int64_t)b->n * CF_NTAB);
    }
    return CF_OK;
}

// Enqueue the stages of one run on the context stream (no host sync, no
// memset/memcpy: the sequence is capturable as a hipGraph).
// The per-bucket stages (inside, exterior, outside, posterior, MEA, Viterbi)
// of one length bucket, in order, on stream st.
int enqueue_bucket(cf_batch* b, const Bucket& bk, unsigned need, hipStream_t st) {
    int rc;
    if (need & CF_LOGZ) {
        if ((rc = launch_inside(b, bk, st))) return rc;
        if ((rc = launch_exterior(b, bk, 1 | ((need & CF_BPP) ? 2 : 0), st))) return rc;
    } else if (need & CF_BPP) {
        // inside tables are resident; the outside still needs its exterior chain F5o
        if ((rc = launch_exterior(b, bk, 2, st))) return rc;
    }
    if (need & CF_BPP) {
        if ((rc = launch_outside(b, bk, st))) return rc;
        if ((rc = launch_post(b, bk, st))) return rc;
    }
    if ((need & CF_MEA) && (rc = launch_mea(b, bk, st))) return rc;
    if ((need & CF_MFE) && (rc = launch_viterbi(b, bk, false, st))) return rc;
    if ((need & CF_LOGZ64) && (rc = launch_viterbi(b, bk, true, st))) return rc;
    return CF_OK;
}

// The pipelines a forked run spreads over the lane streams: every length
// bucket, and a bucket of uneven lengths (longest >= 1.5x shortest) with at
// least 2 * kPartMin sequences cut into up to kLanes consecutive parts.  In
// such a bucket each stage ends with its last folds still running on a
// partly idle GPU (the packed sweeps' fold queue hands out the longest folds
// first, so these are short ones, but the next stage waits for all of them);
// parts on different streams let the next stage of the longest part start in
// that idle room (20,000 x U[30, 500] nt, MEA at 3 gammas + Viterbi, one
// packed bucket: 9,910 folds/s whole, 10,860 in four parts).  Equal lengths
// gain nothing from it (10,000 x 200 nt: 190.0 -> 194.4 ms per step with four
// parts), so such buckets stay whole.
constexpr int kPartMin = 1024;
std::vector<Bucket> fork_parts(const cf_batch* b) {
    std::vector<Bucket> parts;
    for (const auto& bk : b->buckets) {
        const int n = bk.end - bk.begin;
        const bool uneven = n > 0 && 2 * b->L[b->order[bk.begin]] >= 3 * b->L[b->order[bk.end - 1]];
        const int k = uneven ? std::max(1, std::min(kLanes, n / kPartMin)) : 1;
        for (int p = 0; p < k; ++p) {
            Bucket q = bk;
            q.begin = bk.begin + (int)((int64_t)n * p / k);
            q.end = bk.begin + (int)((int64_t)n * (p + 1) / k);
            q.Lmax = b->L[b->order[q.begin]];  // sorted longest first
            if (q.end > q.begin) parts.push_back(q);
        }
    }
    return parts;
}

// Streams and events for forking a run (created once per batch, never while
// a stream is being captured).
int ensure_lanes(cf_batch* b) {
    cf_ctx* c = b->ctx;
    const int nlanes = std::min((int)fork_parts(b).size(), kLanes);
    for (int l = (int)b->lanes.size(); l < nlanes; ++l) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        CF_HIP(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "lane stream");
        b->lanes.push_back(s);
        CF_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming), "lane event");
        b->lane_ev.push_back(e);
    }
    if (!b->fork_ev) CF_HIP(c, hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming), "fork event");
    if (!b->vlane) CF_HIP(c, hipStreamCreateWithFlags(&b->vlane, hipStreamNonBlocking), "viterbi lane");
    if (!b->vlane_ev) CF_HIP(c, hipEventCreateWithFlags(&b->vlane_ev, hipEventDisableTiming), "viterbi event");
    return CF_OK;
}

// Enqueue a run on the batch stream.  With `timing`, stages run one after the
// other across all buckets and HIP events mark each stage.  Otherwise, when
// the batch may fork and has several parts (length buckets, large buckets
// cut up), each part's pipeline runs on a lane stream: a bucket of long folds
// (few workgroups, long serial chains) then overlaps with the short buckets
// instead of running alone, and the parts of one bucket fill each other's
// launch tails.
int enqueue_run(cf_batch* b, unsigned need, uint64_t seed, bool timing, bool fork_ok) {
    cf_ctx* c = b->ctx;
    int rc;
    // The FP64 Viterbi / exact-logZ sweep reads only the codes and masks: when
    // the run also has FP32 stages, it runs on its own stream beside them from
    // the start (split teams excluded: their workgroups must be co-resident).
    bool split_any = false;
    for (const auto& bk : b->buckets) split_any |= bk.split > 1;
    const unsigned fp64 = need & (CF_MFE | CF_LOGZ64);
    const bool vit_lane = !timing && fork_ok && b->allow_fork && !split_any && fp64 != 0u &&
                          (need & (CF_LOGZ | CF_BPP | CF_MEA | CF_SAMPLE | CF_GRAD)) != 0u;
    if (vit_lane) {
        if ((rc = ensure_lanes(b))) return rc;
        CF_HIP(c, hipEventRecord(b->fork_ev, b->stream), "fork");
        CF_HIP(c, hipStreamWaitEvent(b->vlane, b->fork_ev, 0), "fork wait");
        for (const auto& bk : b->buckets) {
            if ((need & CF_MFE) && (rc = launch_viterbi(b, bk, false, b->vlane))) return rc;
            if ((need & CF_LOGZ64) && (rc = launch_viterbi(b, bk, true, b->vlane))) return rc;
        }
        CF_HIP(c, hipEventRecord(b->vlane_ev, b->vlane), "join");
        need &= ~fp64;
    }
    const std::vector<Bucket> parts =
        (!timing && fork_ok && b->allow_fork) ? fork_parts(b) : std::vector<Bucket>();
    const int np = (int)parts.size();
    // The sampler reads only the inside tables and F5; when the run also has
    // an outside sweep and a posterior, it runs on a lane stream beside them
    // (each fills the other's idle room: 10,000 x 200 nt, logZ + bpp + 100
    // samples, 190.7 -> 178.2 ms per step), samples unchanged.
    const bool sample_early = np <= 1 && fork_ok && b->allow_fork && (need & CF_SAMPLE) && (need & CF_LOGZ) &&
                              (need & CF_BPP);
    b->sample_lane = sample_early;
    if (np > 1) {
        const int nlanes = std::min(np, kLanes);
        if ((rc = ensure_lanes(b))) return rc;
        CF_HIP(c, hipEventRecord(b->fork_ev, b->stream), "fork");
        for (int l = 0; l < nlanes; ++l) CF_HIP(c, hipStreamWaitEvent(b->lanes[l], b->fork_ev, 0), "fork wait");
        // parts are longest first: the long ones start first, one per lane
        for (int k = 0; k < np; ++k)
            if ((rc = enqueue_bucket(b, parts[k], need, b->lanes[k % nlanes]))) return rc;
        for (int l = 0; l < nlanes; ++l) {
            CF_HIP(c, hipEventRecord(b->lane_ev[l], b->lanes[l]), "join");
            CF_HIP(c, hipStreamWaitEvent(b->stream, b->lane_ev[l], 0), "join wait");
        }
    } else {
        auto mark = [&](int k) {
            if (timing) (void)hipEventRecord(b->ev[k], b->stream);
        };
        mark(0);
        for (const auto& bk : b->buckets)
            if ((need & CF_LOGZ) && (rc = launch_inside(b, bk, b->stream))) return rc;
        mark(1);
        for (const auto& bk : b->buckets) {
            if (need & CF_LOGZ) {
                if ((rc = launch_exterior(b, bk, 1 | ((need & CF_BPP) ? 2 : 0), b->stream))) return rc;
            } else if ((need & CF_BPP) && (rc = launch_exterior(b, bk, 2, b->stream))) {
                return rc;
            }
        }
        if (sample_early) {
            // One marker between the exterior and the outside sweep on the batch
            // stream, timed or not (a timed run forks at the stage mark, which
            // also starts the sampler's time): with a second marker the sampler
            // was dispatched ahead of the outside sweep in about one timed step
            // of three, 148 -> 157 ms per step at 10,000 x 200 nt.
            if ((rc = ensure_lanes(b))) return rc;
            hipEvent_t fork = timing ? b->ev[2] : b->fork_ev;
            CF_HIP(c, hipEventRecord(fork, b->stream), "fork");
            CF_HIP(c, hipStreamWaitEvent(b->lanes[0], fork, 0), "fork wait");
            if ((rc = launch_sample(b, seed, b->lanes[0]))) return rc;
            if (timing) (void)hipEventRecord(b->ev[kNumKernels + 4], b->lanes[0]);
            CF_HIP(c, hipEventRecord(b->lane_ev[0], b->lanes[0]), "join");
        } else {
            mark(2);
        }
        for (const auto& bk : b->buckets)
            if ((need & CF_BPP) && (rc = launch_outside(b, bk, b->stream, sample_early))) return rc;
        mark(3);
        for (const auto& bk : b->buckets)
            if ((need & CF_BPP) && (rc = launch_post(b, bk, b->stream))) return rc;
        mark(4);
        for (const auto& bk : b->buckets)
            if ((need & CF_MEA) && (rc = launch_mea(b, bk, b->stream))) return rc;
        mark(5);
        for (const auto& bk : b->buckets) {
            if ((need & CF_MFE) && (rc = launch_viterbi(b, bk, false, b->stream))) return rc;
            if ((need & CF_LOGZ64) && (rc = launch_viterbi(b, bk, true, b->stream))) return rc;
        }
        if (sample_early) CF_HIP(c, hipStreamWaitEvent(b->stream, b->lane_ev[0], 0), "join wait");
    }
    if (timing) (void)hipEventRecord(b->ev[6], b->stream);
    if ((need & CF_SAMPLE) && !sample_early && (rc = launch_sample(b, seed, b->stream))) return rc;
    if (timing) (void)hipEventRecord(b->ev[7], b->stream);
    if ((need & CF_GRAD) && (rc = launch_grad(b, b->stream))) return rc;
    if (timing) (void)hipEventRecord(b->ev[8], b->stream);
    if (vit_lane) CF_HIP(c, hipStreamWaitEvent(b->stream, b->vlane_ev, 0), "join wait");
    return CF_OK;
}


size_t batch_staging_bytes(int n, int64_t C) {
    return measure([&](Carver& k) {
        k.take<CfSeq>(n);
        k.take<int>(n);
        k.take<uint8_t>(C);
        k.take<uint8_t>(C);
        k.take<uint64_t>(n);
        k.take<int64_t>(n);
        k.take<int64_t>(n);
        k.take<float>(kMaxGamma + n);
        k.take<int>(4);
        k.take<unsigned>((int64_t)16 * n);
        k.take<int>((int64_t)4 * n);
    });
}

// Workgroups per fold for a wide bucket of nfold folds of up to Lb nt.  The
// parts of a fold wait for each other, so a split launch is kept to at most
// two workgroups per CU in total: it fits the device beside other work.
int bucket_split(const cf_ctx* c, int Lb, int nfold, int split) {
    int K = split > 0 ? split : ((c->split_min > 0 && Lb >= c->split_min) ? c->split_k : 1);
    while (K > 1 && (int64_t)cf_split_grid(nfold, K) > 2 * (int64_t)c->n_cu) --K;
    return K;
}

int batch_create(cf_ctx* c, int n, const char* seqs, const int64_t* offsets, const uint8_t* masks,
                 const uint64_t* keys, hipStream_t stream, void** staging, size_t* staging_bytes, bool sync,
                 cf_batch** out, int width, int split) {
    if (!c || !out) return fail(c, CF_EINVAL, "null argument");
    *out = nullptr;
    if (n <= 0) return fail(c, CF_EINVAL, "empty batch");
    if (!offsets || (!seqs && offsets[n] > 0)) return fail(c, CF_EINVAL, "null sequence buffer");
    if (!c->dP32) return fail(c, CF_ENOPARAMS, "parameters not set (cf_set_params)");
    CF_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    cf_batch* b = new cf_batch();
    b->ctx = c;
    b->stream = stream;
    b->n = n;
    b->L.resize(n);
    b->charoff.resize(n);
    for (int i = 0; i < n; ++i) {
        const int64_t l = offsets[i + 1] - offsets[i];
        if (l < 0 || l > 32000) {
            delete b;
            return fail(c, CF_EINVAL, "sequence length out of range (0..32000)");
        }
        b->L[i] = (int)l;
        b->charoff[i] = b->total_chars;
        b->total_chars += l;
        b->Lmax = std::max(b->Lmax, (int)l);
    }
    launch_order(n, b->L, seqs, offsets, masks, b->order);
    {
        int pos = 0;
        const int nb = (int)c->edges.size();
        for (int k = nb - 1; k >= 0; --k) {
            const int lo = k ? c->edges[k - 1] : -1;
            int end = pos;
            while (end < n && b->L[b->order[end]] > lo) ++end;
            if (end > pos) {
                const int Lb = b->L[b->order[pos]];
                int w = width > 0 ? width : c->threads[k];
                // Built-in tuning, small packed bucket: one wave per fold would
                // leave most SIMDs idle, so each fold gets as many waves as keep
                // about two per SIMD over the bucket, at most four and no more
                // than one diagonal of the fold can use (logZ + MEA + Viterbi,
                // ms per run, built-in -> this rule: 1,000 x 128 nt 2 waves, as
                // before; 500 x 200 nt 33.6 -> 29.5 with 4 instead of 2 waves;
                // 100 x 200 nt 4 waves and 2,000 x 128 nt packed, as before).
                if (width <= 0 && !c->tuned && w == 64) {
                    const int nfold = end - pos;
                    const int by_gpu = (8 * c->n_cu) / std::max(1, nfold);
                    const int by_len = (Lb + 1 + 63) / 64;
                    const int waves = std::min(4, std::min(by_gpu, by_len));
                    if (waves >= 2) w = 64 * waves;
                }
                // wide buckets of long folds may split each fold over several
                // CUs (one 256-thread workgroup, one wave per SIMD, per part)
                const int K = w > 64 ? bucket_split(c, Lb, end - pos, split) : 1;
                if (K > 1) w = 256;
                b->buckets.push_back({pos, end, w, Lb, K});
            }
            pos = end;
        }
    }
    std::vector<CfSeq> sd(n);
    int64_t C = 0, T = 0, X = 0, M = 0;
    for (int i = 0; i < n; ++i) {
        sd[i].L = b->L[i];
        sd[i].coff = (int32_t)C;
        sd[i].toff = T;
        sd[i].xoff = X;
        sd[i].moff = M;
        C += b->L[i] + 2;
        T += cf_tri_size(b->L[i]);
        X += std::max<int64_t>(cf_x_size(b->L[i]), 1);
        M += cf_mea_cells(b->L[i]);
    }
    if (C > INT32_MAX) {
        delete b;
        return fail(c, CF_EINVAL, "batch too large (total residues)");
    }
    b->total_C = C;
    b->total_T = T;
    b->total_X = X;
    b->total_M = M;
    const size_t in_bytes = batch_staging_bytes(n, C);
    int rc = region_get(c, b->r_in, in_bytes);
    if (rc) { delete b; return rc; }
    auto lay_inside = [&](Carver& k) {
        for (int t = 0; t < 4; ++t) k.take<float>(T);
        k.take<float>(X);
        k.take<float>(C);
        k.take<float>(C);
        k.take<float>(n);
    };
    rc = region_get(c, b->r_inside, measure(lay_inside));
    if (rc) { region_put(c, b->r_in); delete b; return rc; }
    {
        Carver k{(char*)b->r_inside.p};
        b->FC = k.take<float>(T);
        b->FM = k.take<float>(T);
        b->FM1 = k.take<float>(T);
        b->FM2 = k.take<float>(T);
        b->FCX = k.take<float>(X);
        b->F5 = k.take<float>(C);
        b->F5o = k.take<float>(C);
        b->logz = k.take<float>(n);
    }
    if (*staging_bytes < in_bytes) {
        if (*staging) (void)hipHostFree(*staging);
        *staging = nullptr;
        *staging_bytes = 0;
        if (hipHostMalloc(staging, in_bytes) != hipSuccess) {
            *staging = nullptr;
            cf_batch_destroy(b);
            return fail(c, CF_ENOMEM, "pinned staging allocation failed");
        }
        *staging_bytes = in_bytes;
    }
    Carver h{(char*)*staging};
    CfSeq* h_seqs = h.take<CfSeq>(n);
    int* h_order = h.take<int>(n);
    uint8_t* h_codes = h.take<uint8_t>(C);
    uint8_t* h_masks = h.take<uint8_t>(C);
    uint64_t* h_keys = h.take<uint64_t>(n);
    int64_t* h_charoff = h.take<int64_t>(n);
    int64_t* h_soff = h.take<int64_t>(n);
    std::memcpy(h_seqs, sd.data(), sizeof(CfSeq) * n);
    std::memcpy(h_order, b->order.data(), sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        uint8_t* cs = h_codes + sd[i].coff;
        uint8_t* ms = h_masks + sd[i].coff;
        const int l = b->L[i];
        cs[0] = 4;
        cs[l + 1] = 4;
        ms[0] = 0;
        ms[l + 1] = 0;
        const char* src = seqs ? seqs + offsets[i] : nullptr;
        const uint8_t* msrc = masks ? masks + offsets[i] : nullptr;
        for (int q = 0; q < l; ++q) {
            cs[q + 1] = encode_base(src[q]);
            ms[q + 1] = (msrc && msrc[q]) ? 1 : 0;
        }
        h_keys[i] = keys ? keys[i] : (uint64_t)i;
        h_charoff[i] = b->charoff[i];
        h_soff[i] = 0;
    }
    Carver d{(char*)b->r_in.p};
    b->d_seqs = d.take<CfSeq>(n);
    b->d_order = d.take<int>(n);
    b->d_codes = d.take<uint8_t>(C);
    uint8_t* dm = d.take<uint8_t>(C);
    b->d_masks = masks ? dm : nullptr;
    b->d_keys = d.take<uint64_t>(n);
    b->d_charoff = d.take<int64_t>(n);
    b->d_soff = d.take<int64_t>(n);
    b->d_gin = d.take<float>(kMaxGamma + n);
    b->d_status = d.take<int>(4);
    b->d_xsync = d.take<unsigned>((int64_t)16 * n);
    b->d_queue = d.take<int>((int64_t)4 * n);
    for (auto& ev : b->ev) {
        if (hipEventCreate(&ev) != hipSuccess) {
            cf_batch_destroy(b);
            return fail(c, CF_EDEVICE, "hipEventCreate failed");
        }
    }
    // upload everything except the per-run tail (d_soff / d_gin are written by batch_enqueue)
    const size_t head = (size_t)((char*)b->d_soff - (char*)b->r_in.p);
    hipError_t e = hipMemcpyAsync(b->r_in.p, *staging, head, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && sync) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        cf_batch_destroy(b);
        return hipfail(c, e, "upload batch");
    }
    *out = b;
    return CF_OK;
}

int batch_enqueue(cf_batch* b, unsigned flags, const float* gammas, int n_gamma, int n_samples, uint64_t seed,
                  const float* weights, void* staging, size_t staging_bytes, bool timing, unsigned* ran) {
    cf_ctx* c = b->ctx;
    unsigned need = flags & (CF_LOGZ | CF_BPP | CF_MEA | CF_MFE | CF_SAMPLE | CF_GRAD | CF_LOGZ64);
    if (need & (CF_MEA | CF_GRAD)) need |= CF_BPP;
    if (need & (CF_BPP | CF_SAMPLE)) need |= CF_LOGZ;
    if (need & CF_MEA) {
        if (!gammas || n_gamma <= 0 || n_gamma > kMaxGamma) return fail(c, CF_EINVAL, "MEA needs 1..64 gammas");
        for (int g = 0; g < n_gamma; ++g)
            if (!(gammas[g] > 0.0f) || !std::isfinite(gammas[g])) return fail(c, CF_EINVAL, "gamma must be > 0");
    } else {
        n_gamma = 0;
    }
    if (need & CF_SAMPLE) {
        if (n_samples <= 0) return fail(c, CF_EINVAL, "n_samples must be > 0");
    } else {
        n_samples = 0;
    }
    int rc = ensure_buffers(b, need, n_gamma, n_samples);
    if (rc) return rc;
    // a stage that does not run keeps the results (and their counts) of its last run
    if (need & CF_MEA) b->n_gamma = n_gamma;
    if (need & CF_SAMPLE) b->n_samples = n_samples;
    if (b->params_version != c->params_version) b->done = 0;
    unsigned run = need;
    if (!(flags & CF_RECOMPUTE) && !(flags & CF_GRAPH)) {
        if (b->done & CF_LOGZ) run &= ~CF_LOGZ;
        if (b->done & CF_BPP) run &= ~CF_BPP;
    }
    // per-run inputs: sample offsets, gammas, gradient weights and a cleared
    // status word (pinned -> device, async; outside any captured graph)
    const size_t tail = (size_t)((char*)(b->d_status + 4) - (char*)b->d_soff);
    if (staging_bytes < tail) return fail(c, CF_EINVAL, "staging buffer too small");
    {
        char* base = (char*)staging;
        int64_t* so = (int64_t*)base;
        for (int i = 0; i < b->n; ++i) so[i] = b->charoff[i] * (int64_t)n_samples;
        const size_t gap = (size_t)((char*)b->d_gin - (char*)b->d_soff);
        float* gw = (float*)(base + gap);
        for (int g = 0; g < kMaxGamma; ++g) gw[g] = g < n_gamma ? gammas[g] : 1.0f;
        for (int i = 0; i < b->n; ++i) gw[kMaxGamma + i] = weights ? weights[i] : 1.0f;
        std::memset(base + ((char*)b->d_status - (char*)b->d_soff), 0, 4 * sizeof(int));
        CF_HIP(c, hipMemcpyAsync(b->d_soff, staging, tail, hipMemcpyHostToDevice, b->stream), "per-run inputs");
    }
    CF_HIP(c, hipEventRecord(b->ev[kNumKernels + 1], b->stream), "event");
    if (flags & CF_GRAPH) {
        // the graph replays the launches it captured, with the buffer addresses
        // of that moment: any reallocation since then forces a new capture
        const bool reuse = b->gexec && b->g_flags == run && b->g_ngamma == n_gamma && b->g_nsamp == n_samples &&
                           b->g_seed == seed && b->g_gen == b->buf_gen;
        if (!reuse) {
            if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
            if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
            // the per-bucket streams join the capture through the fork/join events
            if (b->allow_fork && (rc = ensure_lanes(b))) return rc;
            CF_HIP(c, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
            rc = enqueue_run(b, run, seed, false, true);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return hipfail(c, e, "end capture");
            b->graph = g;
            CF_HIP(c, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0), "graph instantiate");
            b->g_flags = run;
            b->g_ngamma = n_gamma;
            b->g_nsamp = n_samples;
            b->g_seed = seed;
            b->g_gen = b->buf_gen;
        }
        CF_HIP(c, hipGraphLaunch(b->gexec, b->stream), "graph launch");
    } else {
        rc = enqueue_run(b, run, seed, timing, true);
        if (rc) return rc;
    }
    CF_HIP(c, hipEventRecord(b->ev[kNumKernels + 2], b->stream), "event");
    b->done |= need;
    b->params_version = c->params_version;
    if (ran) *ran = run;
    return CF_OK;
}

// ------------------------------------------------------------ no-alloc path
// One fixed device array per GPU for cf_fold_noalloc: the inputs, the DP
// tables and the results of one call are carved out of it, so the call makes
// no device allocation.  Its device address is looked up once per context
// (cf_create, with the context's device current) and used from then on.
constexpr size_t kScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Layout of a no-alloc call in the scratch array (Carver over `base`; with
// base = nullptr it only measures).
struct NoAllocLayout {
    CfSeq* seqs;
    int* order;
    uint8_t *codes, *masks;
    int64_t* charoff;
    float* gin;
    int* status;
    unsigned* xsync;
    int* queue;
    float *FC, *FM, *FM1, *FM2, *FCX, *F5, *F5o, *logz, *FCo, *FMo, *FM1o, *FM2o, *POST, *SC;
    uint8_t* mea_out;
    size_t bytes;
};
NoAllocLayout noalloc_layout(char* base, int n, int64_t C, int64_t T, int64_t X, int64_t M, int64_t chars,
                            bool mea) {
    NoAllocLayout l;
    Carver k{base};
    l.seqs = k.take<CfSeq>(n);
    l.order = k.take<int>(n);
    l.codes = k.take<uint8_t>(C);
    l.masks = k.take<uint8_t>(C);
    l.charoff = k.take<int64_t>(n);
    l.gin = k.take<float>(kMaxGamma + n);
    l.status = k.take<int>(4);
    l.xsync = k.take<unsigned>((int64_t)16 * n);
    l.queue = k.take<int>((int64_t)4 * n);
    l.FC = k.take<float>(T);
    l.FM = k.take<float>(T);
    l.FM1 = k.take<float>(T);
    l.FM2 = k.take<float>(T);
    l.FCX = k.take<float>(X);
    l.F5 = k.take<float>(C);
    l.F5o = k.take<float>(C);
    l.logz = k.take<float>(n);
    l.FCo = l.FMo = l.FM1o = l.FM2o = l.POST = l.SC = nullptr;
    l.mea_out = nullptr;
    if (mea) {
        l.FCo = k.take<float>(T);
        l.FMo = k.take<float>(T);
        l.FM1o = k.take<float>(T);
        l.FM2o = k.take<float>(T);
        l.POST = k.take<float>(T);
        l.SC = k.take<float>(M);
        l.mea_out = k.take<uint8_t>(chars);
    }
    l.bytes = align_up(k.off, 256);
    return l;
}

}  // namespace cfi

// =================================================================== C ABI
using namespace cfi;

extern "C" {

int cf_version(void) { return CF_VERSION; }
int cf_table_size(void) { return CF_NTAB; }
int cf_kernel_count(void) { return cfi::kNumKernels; }
const char* cf_kernel_name(int k) { return (k >= 0 && k < cfi::kNumKernels) ? cfi::kKernelNames[k] : ""; }

int cf_create(int device, cf_ctx** out) {
    if (!out) return CF_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CF_EDEVICE;
    if (device < 0 || device >= ndev) return CF_EINVAL;
    cf_ctx* c = new cf_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return CF_EDEVICE; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return CF_EDEVICE; }
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
        c->lds_max = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0) c->n_cu = v;
    const void* dyn[] = {(const void*)cf_inside_kernel<false, false>, (const void*)cf_inside_kernel<true, false>,
                         (const void*)cf_inside_kernel<false, true>,
                         (const void*)cf_outside_kernel<true, false, false>,
                         (const void*)cf_outside_kernel<false, false, false>,
                         (const void*)cf_outside_kernel<false, true, false>,
                         (const void*)cf_outside_kernel<false, false, true>, (const void*)cf_post_kernel,
                         (const void*)cf_exterior_kernel,
                         (const void*)cf_mea_kernel<true>, (const void*)cf_mea_kernel<false>,
                         (const void*)cf_viterbi_kernel<false>,
                         (const void*)cf_viterbi_kernel<true>};
    for (const void* f : dyn) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_max);
    (void)hipGetLastError();
    // the no-alloc path's fixed array: its address on this device (a host-side
    // use of the symbol's name would not be a device address)
    void* sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(cfi::scratch)) == hipSuccess) c->scratch = sp;
    else (void)hipGetLastError();
    *out = c;
    return CF_OK;
}

void cf_destroy(cf_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->dP32) (void)hipFree(c->dP32);
    if (c->dP64) (void)hipFree(c->dP64);
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pool.trim();
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* cf_last_error(const cf_ctx* c) { return c ? c->err.c_str() : "null context"; }

int cf_set_params(cf_ctx* c, const double* tables) {
    if (!c || !tables) return fail(c, CF_EINVAL, "null argument");
    CF_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    CfParams<float> h32;
    CfParams<double> h64;
    for (int k = 0; k < CF_NTAB; ++k) {
        if (!std::isfinite(tables[k])) return fail(c, CF_EINVAL, "non-finite parameter value");
        h64.t[k] = tables[k];
        h32.t[k] = (float)tables[k];  // cpu._params_f32: FP64 tables cast to float32
    }
    for (int a = 0; a < 5; ++a)
        for (int bb = 0; bb < 5; ++bb)
            for (int cc = 0; cc < 5; ++cc)
                for (int d = 0; d < 5; ++d) {
                    const int q = cf_i4(a, bb, cc, d);
                    // the reference's ScoreJunctionB: hc[a][b] + tm[a][b][c][d], one add in each precision
                    volatile float hf = h32.t[CF_O_HC + cf_i2(a, bb)];
                    volatile float tf = h32.t[CF_O_TM + q];
                    h32.JB[q] = hf + tf;
                    h64.JB[q] = h64.t[CF_O_HC + cf_i2(a, bb)] + h64.t[CF_O_TM + q];
                }
    if (!c->dP32) CF_HIP(c, hipMalloc((void**)&c->dP32, sizeof(CfParams<float>)), "hipMalloc params");
    if (!c->dP64) CF_HIP(c, hipMalloc((void**)&c->dP64, sizeof(CfParams<double>)), "hipMalloc params");
    CF_HIP(c, hipStreamSynchronize(c->stream), "sync");
    CF_HIP(c, hipMemcpy(c->dP32, &h32, sizeof(h32), hipMemcpyHostToDevice), "upload params");
    CF_HIP(c, hipMemcpy(c->dP64, &h64, sizeof(h64), hipMemcpyHostToDevice), "upload params");
    ++c->params_version;
    return CF_OK;
}

int cf_copy_probe(cf_ctx* c, int64_t bytes) {
    if (!c || bytes <= 0 || bytes % 16) return fail(c, CF_EINVAL, "copy probe: bytes must be a positive multiple of 16");
    CF_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    void *src = nullptr, *dst = nullptr;
    CF_HIP(c, hipMalloc(&src, (size_t)bytes), "hipMalloc probe");
    hipError_t e = hipMalloc(&dst, (size_t)bytes);
    if (e != hipSuccess) {
        (void)hipFree(src);
        return hipfail(c, e, "hipMalloc probe");
    }
    const int64_t n = bytes / 16;
    hipLaunchKernelGGL(cf_copy_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)), dim3(256), 0,
                       c->stream, (const float4*)src, (float4*)dst, n);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return hipfail(c, e, "copy probe");
    return CF_OK;
}

int cf_set_split(cf_ctx* c, int min_length, int wgs_per_fold) {
    if (!c || min_length < 0 || wgs_per_fold < 1 || wgs_per_fold > 16)
        return fail(c, CF_EINVAL, "split: min_length >= 0 and 1..16 workgroups per fold");
    c->split_min = wgs_per_fold > 1 ? min_length : 0;
    c->split_k = wgs_per_fold;
    return CF_OK;
}

int cf_set_packing(cf_ctx* c, int teams_inside, int teams_outside) {
    if (!c || teams_inside < 1 || teams_inside > 16 || teams_outside < 1 || teams_outside > 16)
        return fail(c, CF_EINVAL, "packing: 1..16 teams per workgroup");
    c->pack_in = teams_inside;
    c->pack_out = teams_outside;
    return CF_OK;
}

int cf_set_tuning(cf_ctx* c, const int* edges, const int* threads, int nb) {
    if (!c || !edges || !threads || nb <= 0) return fail(c, CF_EINVAL, "bad tuning arguments");
    std::vector<int> e(edges, edges + nb), t(threads, threads + nb);
    for (int k = 0; k < nb; ++k) {
        if (t[k] < 64 || t[k] > 1024 || t[k] % 64) return fail(c, CF_EINVAL, "threads must be a multiple of 64 in 64..1024");
        if (k && e[k] <= e[k - 1]) return fail(c, CF_EINVAL, "bucket edges must increase");
    }
    e.back() = 1 << 30;
    c->edges = e;
    c->threads = t;
    c->tuned = true;
    return CF_OK;
}

int64_t cf_fold_bytes(int L, unsigned flags, int n_gamma, int n_samples) {
    if (L < 0) return -1;
    const int64_t T = cf_tri_size(L), X = cf_x_size(L);
    int64_t bytes = (L + 2) * 2 + sizeof(CfSeq) + 64;  // codes, masks, descriptor
    bytes += (4 * T + X + 2 * (L + 2) + 1) * 4;        // inside tables, FCX, F5, F5o, logZ
    if (flags & (CF_BPP | CF_MEA | CF_GRAD)) bytes += 5 * T * 4;
    if (flags & CF_MEA) bytes += (int64_t)std::max(1, n_gamma) * (cf_mea_cells(L) * 4 + L);
    if (flags & CF_MFE) bytes += 3 * T * 8 + 8 + L;
    if (flags & CF_SAMPLE) bytes += (int64_t)std::max(1, n_samples) * L;
    return bytes;
}

int cf_batch_create(cf_ctx* c, int n, const char* seqs, const int64_t* offsets, const uint8_t* masks,
                    const uint64_t* keys, cf_batch** out) {
    if (!c) return CF_EINVAL;
    const int rc = cfi::batch_create(c, n, seqs, offsets, masks, keys, c->stream, &c->pinned, &c->pinned_bytes, true, out);
    if (rc == CF_OK) (*out)->allow_fork = true;
    return rc;
}

int cf_batch_run(cf_batch* b, unsigned flags, const float* gammas, int n_gamma, int n_samples, uint64_t seed,
                 const float* grad_weights) {
    if (!b) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    CF_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    const size_t need_stage = sizeof(float) * (cfi::kMaxGamma + (size_t)b->n) + sizeof(int64_t) * (size_t)b->n + 4096;
    if (c->pinned_bytes < need_stage) {
        CF_HIP(c, hipStreamSynchronize(b->stream), "sync");
        if (c->pinned) (void)hipHostFree(c->pinned);
        c->pinned = nullptr;
        c->pinned_bytes = 0;
        CF_HIP(c, hipHostMalloc(&c->pinned, need_stage), "pinned staging");
        c->pinned_bytes = need_stage;
    }
    unsigned ran = 0;
    int rc = cfi::batch_enqueue(b, flags, gammas, n_gamma, n_samples, seed, grad_weights, c->pinned,
                                c->pinned_bytes, (flags & CF_TIMING) != 0, &ran);
    if (rc) {
        (void)hipStreamSynchronize(b->stream);
        return rc;
    }
    CF_HIP(c, hipStreamSynchronize(b->stream), "fold run");
    {
        int st = 0;
        CF_HIP(c, hipMemcpy(&st, b->d_status, sizeof(int), hipMemcpyDeviceToHost), "read status");
        if (st & 2) {
            b->done = 0;
            return fail(c, CF_EDEVICE, "split-team fold: its workgroups were not resident together (results "
                                       "discarded)");
        }
        if (st & 1) {
            b->done &= ~(unsigned)CF_SAMPLE;
            return fail(c, CF_EDEVICE, "sampler traceback stack overflow (samples discarded)");
        }
    }
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b->ev[cfi::kNumKernels + 1], b->ev[cfi::kNumKernels + 2]) == hipSuccess)
        b->last_ms = ms;
    if ((flags & CF_TIMING) && !(flags & CF_GRAPH)) {
        for (int k = 0; k < cfi::kNumKernels; ++k) {
            float t = 0.0f;
            if (hipEventElapsedTime(&t, b->ev[k], b->ev[k + 1]) == hipSuccess) b->kms[k] = t;
        }
        // a sampler on the lane stream runs from the outside sweep's start mark
        // to its own end mark
        float t = 0.0f;
        if (b->sample_lane && hipEventElapsedTime(&t, b->ev[2], b->ev[cfi::kNumKernels + 4]) == hipSuccess)
            b->kms[6] = t;
    }
    return CF_OK;
}

int cf_batch_logz(const cf_batch* b, float* out) {
    if (!b || !out) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    if (!(b->done & CF_LOGZ)) return fail(c, CF_ESTATE, "logZ not computed");
    CF_HIP(c, hipMemcpy(out, b->logz, sizeof(float) * b->n, hipMemcpyDeviceToHost), "read logz");
    return CF_OK;
}

int cf_batch_bpp(const cf_batch* b, int index, float* out) {
    if (!b || !out || index < 0 || index >= b->n) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    if (!(b->done & CF_BPP)) return fail(c, CF_ESTATE, "bpp not computed");
    const int L = b->L[index];
    int64_t toff = 0;
    for (int i = 0; i < index; ++i) toff += cf_tri_size(b->L[i]);
    std::vector<float> tri((size_t)cf_tri_size(L));
    CF_HIP(c, hipMemcpy(tri.data(), b->POST + toff, sizeof(float) * tri.size(), hipMemcpyDeviceToHost), "read bpp");
    std::fill(out, out + (size_t)L * L, 0.0f);
    for (int a = 1; a <= L; ++a)
        for (int cc = a + 1; cc <= L; ++cc) out[(size_t)(a - 1) * L + (cc - 1)] = tri[cf_tri(L, a, cc - 1)];
    return CF_OK;
}

int cf_batch_mea(const cf_batch* b, int g, char* out) {
    if (!b || !out) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    if (!(b->done & CF_MEA) || g < 0 || g >= b->n_gamma) return fail(c, CF_ESTATE, "MEA not computed for this gamma");
    if (b->total_chars)
        CF_HIP(c, hipMemcpy(out, b->mea_out + (int64_t)g * b->total_chars, b->total_chars, hipMemcpyDeviceToHost),
               "read mea");
    return CF_OK;
}

int cf_batch_mfe(const cf_batch* b, char* out, double* scores) {
    if (!b || !out) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    if (!(b->done & CF_MFE)) return fail(c, CF_ESTATE, "MFE not computed");
    if (b->total_chars) CF_HIP(c, hipMemcpy(out, b->mfe_out, b->total_chars, hipMemcpyDeviceToHost), "read mfe");
    if (scores) CF_HIP(c, hipMemcpy(scores, b->vscore, sizeof(double) * b->n, hipMemcpyDeviceToHost), "read scores");
    return CF_OK;
}

int cf_batch_samples(const cf_batch* b, char* out) {
    if (!b || !out) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    if (!(b->done & CF_SAMPLE)) return fail(c, CF_ESTATE, "samples not computed");
    const int64_t bytes = (int64_t)b->n_samples * b->total_chars;
    if (bytes) CF_HIP(c, hipMemcpy(out, b->samp_out, bytes, hipMemcpyDeviceToHost), "read samples");
    return CF_OK;
}

int cf_batch_table(const cf_batch* b, int index, int which, float* out) {
    if (!b || !out || index < 0 || index >= b->n || which < 0 || which > 10) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    const bool outside = which >= 4 && which != 9;
    if (!(b->done & CF_LOGZ) || (outside && !(b->done & CF_BPP))) return fail(c, CF_ESTATE, "table not computed");
    const int L = b->L[index];
    int64_t toff = 0, coff = 0;
    for (int i = 0; i < index; ++i) {
        toff += cf_tri_size(b->L[i]);
        coff += b->L[i] + 2;
    }
    const float* src = nullptr;
    int64_t cnt = cf_tri_size(L);
    switch (which) {
        case 0: src = b->FC + toff; break;
        case 1: src = b->FM + toff; break;
        case 2: src = b->FM1 + toff; break;
        case 3: src = b->FM2 + toff; break;
        case 4: src = b->FCo + toff; break;
        case 5: src = b->FMo + toff; break;
        case 6: src = b->FM1o + toff; break;
        case 7: src = b->FM2o + toff; break;
        case 8: src = b->POST + toff; break;
        case 9: src = b->F5 + coff; cnt = L + 2; break;
        default: src = b->F5o + coff; cnt = L + 2; break;
    }
    CF_HIP(c, hipMemcpy(out, src, sizeof(float) * cnt, hipMemcpyDeviceToHost), "read table");
    return CF_OK;
}

int cf_batch_logz64(const cf_batch* b, double* out) {
    if (!b || !out) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    if (!(b->done & CF_LOGZ64)) return fail(c, CF_ESTATE, "FP64 logZ not computed");
    CF_HIP(c, hipMemcpy(out, b->logz64, sizeof(double) * b->n, hipMemcpyDeviceToHost), "read logz64");
    return CF_OK;
}

int cf_batch_grad(const cf_batch* b, double* out) {
    if (!b || !out) return CF_EINVAL;
    cf_ctx* c = b->ctx;
    if (!(b->done & CF_GRAD)) return fail(c, CF_ESTATE, "gradient not computed");
    CF_HIP(c, hipMemcpy(out, b->grad, sizeof(double) * CF_NTAB, hipMemcpyDeviceToHost), "read grad");
    return CF_OK;
}

float cf_batch_last_ms(const cf_batch* b) { return b ? b->last_ms : 0.0f; }

int cf_batch_kernel_ms(const cf_batch* b, float* out) {
    if (!b || !out) return CF_EINVAL;
    std::copy(b->kms, b->kms + cfi::kNumKernels, out);
    return CF_OK;
}

int64_t cf_batch_device_bytes(const cf_batch* b) { return b ? b->device_bytes() : 0; }

int cf_fold_noalloc(cf_ctx* c, int n, const char* seqs, const int64_t* offsets, const uint8_t* masks,
                    unsigned flags, float gamma, float* out_logz, char* out_mea) {
    if (!c) return CF_EINVAL;
    if (n <= 0 || !offsets || !out_logz || (!seqs && offsets[n] > 0)) return fail(c, CF_EINVAL, "bad arguments");
    if (flags != CF_LOGZ && flags != CF_MEA) return fail(c, CF_EINVAL, "no-alloc fold: flags must be CF_LOGZ or CF_MEA");
    const bool mea = flags == CF_MEA;
    if (mea && (!out_mea || !(gamma > 0.0f) || !std::isfinite(gamma))) return fail(c, CF_EINVAL, "MEA needs gamma > 0");
    if (!c->dP32) return fail(c, CF_ENOPARAMS, "parameters not set (cf_set_params)");
    if (!c->scratch) return fail(c, CF_EDEVICE, "no-alloc fold: scratch array address unavailable on this device");
    CF_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    cf_batch b;  // host-side description only; every device buffer is in the scratch array
    b.ctx = c;
    b.stream = c->stream;
    b.n = n;
    b.L.resize(n);
    b.charoff.resize(n);
    for (int i = 0; i < n; ++i) {
        const int64_t l = offsets[i + 1] - offsets[i];
        if (l < 0 || l > 32000) return fail(c, CF_EINVAL, "sequence length out of range");
        b.L[i] = (int)l;
        b.charoff[i] = b.total_chars;
        b.total_chars += l;
        b.Lmax = std::max(b.Lmax, (int)l);
    }
    std::vector<CfSeq> sd(n);
    int64_t C = 0, T = 0, X = 0, M = 0;
    for (int i = 0; i < n; ++i) {
        sd[i].L = b.L[i];
        sd[i].coff = (int32_t)C;
        sd[i].toff = T;
        sd[i].xoff = X;
        sd[i].moff = M;
        C += b.L[i] + 2;
        T += cf_tri_size(b.L[i]);
        X += std::max<int64_t>(cf_x_size(b.L[i]), 1);
        M += cf_mea_cells(b.L[i]);
    }
    const NoAllocLayout need = noalloc_layout(nullptr, n, C, T, X, M, b.total_chars, mea);
    if (need.bytes > kScratchBytes)
        return fail(c, CF_EINVAL, "no-alloc fold: batch needs " + std::to_string(need.bytes) +
                                      " bytes of scratch, more than its " + std::to_string(kScratchBytes));
    const NoAllocLayout l = noalloc_layout((char*)c->scratch, n, C, T, X, M, b.total_chars, mea);
    b.total_C = C;
    b.total_T = T;
    b.total_X = X;
    b.total_M = M;
    // one bucket, longest first; a wide team per fold for small calls
    launch_order(n, b.L, seqs, offsets, masks, b.order);
    const int64_t work = (int64_t)n * b.Lmax * b.Lmax;
    const int width = work <= (int64_t)83000 * c->n_cu ? 256 : 64;
    b.buckets.push_back({0, n, width, b.Lmax, 1});
    // inputs: staged in host memory of this call, copied into the scratch array
    std::vector<uint8_t> codes((size_t)C), mk((size_t)C, 0);
    for (int i = 0; i < n; ++i) {
        uint8_t* cs = codes.data() + sd[i].coff;
        uint8_t* ms = mk.data() + sd[i].coff;
        const int L = b.L[i];
        cs[0] = cs[L + 1] = 4;
        for (int q = 0; q < L; ++q) {
            cs[q + 1] = encode_base(seqs[offsets[i] + q]);
            ms[q + 1] = (masks && masks[offsets[i] + q]) ? 1 : 0;
        }
    }
    std::vector<float> gin(kMaxGamma + n, 1.0f);
    gin[0] = mea ? gamma : 1.0f;
    const int zero[4] = {0, 0, 0, 0};
    hipStream_t st = c->stream;
    CF_HIP(c, hipMemcpyAsync(l.seqs, sd.data(), sizeof(CfSeq) * n, hipMemcpyHostToDevice, st), "upload");
    CF_HIP(c, hipMemcpyAsync(l.order, b.order.data(), sizeof(int) * n, hipMemcpyHostToDevice, st), "upload");
    CF_HIP(c, hipMemcpyAsync(l.codes, codes.data(), (size_t)C, hipMemcpyHostToDevice, st), "upload");
    if (masks) CF_HIP(c, hipMemcpyAsync(l.masks, mk.data(), (size_t)C, hipMemcpyHostToDevice, st), "upload");
    CF_HIP(c, hipMemcpyAsync(l.charoff, b.charoff.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st), "upload");
    CF_HIP(c, hipMemcpyAsync(l.gin, gin.data(), sizeof(float) * gin.size(), hipMemcpyHostToDevice, st), "upload");
    CF_HIP(c, hipMemcpyAsync(l.status, zero, sizeof(zero), hipMemcpyHostToDevice, st), "upload");
    b.d_seqs = l.seqs;
    b.d_order = l.order;
    b.d_codes = l.codes;
    b.d_masks = masks ? l.masks : nullptr;
    b.d_charoff = l.charoff;
    b.d_gin = l.gin;
    b.d_status = l.status;
    b.d_xsync = l.xsync;
    b.d_queue = l.queue;
    b.FC = l.FC; b.FM = l.FM; b.FM1 = l.FM1; b.FM2 = l.FM2; b.FCX = l.FCX;
    b.F5 = l.F5; b.F5o = l.F5o; b.logz = l.logz;
    b.FCo = l.FCo; b.FMo = l.FMo; b.FM1o = l.FM1o; b.FM2o = l.FM2o; b.POST = l.POST;
    b.SC = l.SC; b.mea_out = l.mea_out;
    b.n_gamma = mea ? 1 : 0;
    const unsigned need_flags = mea ? (CF_LOGZ | CF_BPP | CF_MEA) : CF_LOGZ;
    int rc = enqueue_bucket(&b, b.buckets[0], need_flags, st);
    if (rc) {
        (void)hipStreamSynchronize(st);
        return rc;
    }
    CF_HIP(c, hipMemcpyAsync(out_logz, l.logz, sizeof(float) * n, hipMemcpyDeviceToHost, st), "read logz");
    if (mea && b.total_chars)
        CF_HIP(c, hipMemcpyAsync(out_mea, l.mea_out, (size_t)b.total_chars, hipMemcpyDeviceToHost, st), "read mea");
    int status = 0;
    CF_HIP(c, hipMemcpyAsync(&status, l.status, sizeof(int), hipMemcpyDeviceToHost, st), "read status");
    CF_HIP(c, hipStreamSynchronize(st), "no-alloc fold");
    if (status) return fail(c, CF_EDEVICE, "no-alloc fold: device status " + std::to_string(status));
    return CF_OK;
}

void cf_batch_destroy(cf_batch* b) { cfi::batch_free(b, true); }

}  // extern "C"

namespace cfi {
void batch_free(cf_batch* b, bool sync) {
    if (!b) return;
    cf_ctx* c = b->ctx;
    (void)hipSetDevice(c->device);
    if (sync) (void)hipStreamSynchronize(b->stream);
    for (auto s : b->lanes)
        if (s) (void)hipStreamDestroy(s);
    for (auto e : b->lane_ev)
        if (e) (void)hipEventDestroy(e);
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    if (b->vlane) (void)hipStreamDestroy(b->vlane);
    if (b->vlane_ev) (void)hipEventDestroy(b->vlane_ev);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    region_put(c, b->r_in);
    region_put(c, b->r_inside);
    region_put(c, b->r_out);
    region_put(c, b->r_mea);
    region_put(c, b->r_vit);
    region_put(c, b->r_smp);
    region_put(c, b->r_grad);
    delete b;
}
}  // namespace cfi
