// cf_internal.h -- host-side state shared by the batch API (cf_api.hip) and
// the serving engine (cf_server.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "cf_device.h"
#include "contrafold_amd.h"

namespace cfi {

static const char* kKernelNames[] = {"inside", "exterior", "outside", "post", "mea", "viterbi", "sample", "grad"};
constexpr int kNumKernels = 8;
enum { K_INSIDE = 0, K_EXTERIOR, K_OUTSIDE, K_POST, K_MEA, K_VITERBI, K_SAMPLE, K_GRAD };

// Caching device allocator: block sizes are rounded up to a power of two
// (at least 2 MiB) and kept on a free list after release, so steady-state
// batches of varying size reuse blocks instead of calling hipMalloc.
struct DevPool {
    std::multimap<size_t, void*> free_;
    static size_t round(size_t n) {
        size_t r = size_t(2) << 20;
        while (r < n) r <<= 1;
        return r;
    }
    // A free block up to 4x the rounded size is reused (misses cost a
    // hipMalloc); *got is the block's size, which put() must be given back.
    hipError_t get(void** p, size_t n, size_t* got) {
        n = round(n ? n : 1);
        auto it = free_.lower_bound(n);
        if (it != free_.end() && it->first <= 4 * n) {
            *p = it->second;
            *got = it->first;
            free_.erase(it);
            return hipSuccess;
        }
        *got = n;
        hipError_t e = hipMalloc(p, n);
        if (e != hipSuccess) {  // drop the cache and retry once
            (void)hipGetLastError();
            trim();
            e = hipMalloc(p, n);
        }
        return e;
    }
    void put(void* p, size_t n) {
        if (p) free_.emplace(n, p);
    }
    void trim() {
        for (auto& kv : free_) (void)hipFree(kv.second);
        free_.clear();
    }
};

}  // namespace cfi

struct cf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    CfParams<float>* dP32 = nullptr;
    CfParams<double>* dP64 = nullptr;
    uint64_t params_version = 0;
    // bucket edges (max length) and team width per bucket; width 64 = packed
    std::vector<int> edges{128, 384, 1 << 30};
    std::vector<int> threads{64, 64, 512};
    bool tuned = false;  // edges / widths set by cf_set_tuning (else the built-in default)
    // split teams: folds of at least split_min nt in a wide bucket run on
    // split_k workgroups of 256 threads each (0 = off); see cf_set_split
    int split_min = 0, split_k = 1;
    // one-wave teams per workgroup of the packed inside and outside sweeps
    // (fewer when LDS is short); see cf_set_packing
    int pack_in = 8, pack_out = 4;
    int lds_max = 65536;
    int n_cu = 256;
    cfi::DevPool pool;
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    void* scratch = nullptr;  // device address of the no-alloc path's fixed array on this device
};

struct Region {
    void* p = nullptr;
    size_t n = 0;
};

struct cf_batch {
    cf_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;  // the context stream, or a serving stream
    int n = 0, Lmax = 0;
    std::vector<int> L;
    std::vector<int64_t> charoff;
    std::vector<int> order;
    struct Bucket {
        int begin, end, threads, Lmax;
        int split = 1;  // workgroups per fold of the inside / outside / posterior launches
    };
    std::vector<Bucket> buckets;
    int64_t total_C = 0, total_T = 0, total_X = 0, total_M = 0, total_chars = 0;
    // inputs + inside
    Region r_in, r_inside;
    CfSeq* d_seqs = nullptr;
    int* d_order = nullptr;
    uint8_t* d_codes = nullptr;
    uint8_t* d_masks = nullptr;
    uint64_t* d_keys = nullptr;
    int64_t* d_charoff = nullptr;
    int64_t* d_soff = nullptr;
    float* d_gin = nullptr;     // per-run small inputs: gammas (64) then grad weights (n)
    int* d_status = nullptr;    // per-run kernel status bits (1: sampler stack overflow, 2: split team not
                                // co-resident), zeroed by each run
    unsigned* d_xsync = nullptr;  // split-team barrier counters: 16 words per sequence (by launch order)
    int* d_queue = nullptr;       // fold queues of persistent packed launches: 4 per sequence (by launch order)
    float *FC = nullptr, *FM = nullptr, *FM1 = nullptr, *FM2 = nullptr, *FCX = nullptr, *F5 = nullptr,
          *F5o = nullptr, *logz = nullptr;
    // outside
    Region r_out;
    float *FCo = nullptr, *FMo = nullptr, *FM1o = nullptr, *FM2o = nullptr, *POST = nullptr;
    // MEA
    Region r_mea;
    int mea_cap = 0, n_gamma = 0;
    float* SC = nullptr;  // tiled MEA score tables, total_M floats per gamma
    uint8_t* mea_out = nullptr;
    // Viterbi
    Region r_vit;
    double *VFC = nullptr, *VFM = nullptr, *VFM1 = nullptr, *vscore = nullptr, *logz64 = nullptr;
    uint8_t* mfe_out = nullptr;
    // sampling
    Region r_smp;
    int smp_cap = 0, n_samples = 0, stack_cap = 0, sgrid = 0;
    uint8_t* samp_out = nullptr;
    uint32_t* stack = nullptr;
    // gradient
    Region r_grad;
    double* grad = nullptr;
    double* seqcnt = nullptr;  // per-sequence expected counts, n x 2577
    unsigned done = 0;
    uint64_t params_version = 0;
    // stage marks 0..8, run start, run end, an unused slot, and the end of a
    // sampler launched on a lane stream beside the outside sweep (it starts
    // at stage mark 2)
    hipEvent_t ev[cfi::kNumKernels + 5] = {};
    bool sample_lane = false;  // the last run sampled on lanes[0]
    float last_ms = 0.0f;
    float kms[cfi::kNumKernels] = {};
    // per-bucket streams of an eager multi-bucket run (see enqueue_run)
    bool allow_fork = false;
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> lane_ev;
    hipEvent_t fork_ev = nullptr;
    hipStream_t vlane = nullptr;  // the FP64 Viterbi sweep beside the FP32 stages
    hipEvent_t vlane_ev = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    unsigned g_flags = 0;
    int g_ngamma = -1, g_nsamp = -1;
    uint64_t g_seed = 0;
    uint64_t buf_gen = 0, g_gen = 0;  // region reallocations; the count the graph was captured at
    int64_t device_bytes() const {
        return (int64_t)(r_in.n + r_inside.n + r_out.n + r_mea.n + r_vit.n + r_smp.n + r_grad.n);
    }
};


namespace cfi {
constexpr int kMaxGamma = 64;
int fail(cf_ctx* c, int code, const std::string& msg);
// Pack and upload a batch on `stream` from the pinned buffer `staging`
// (which must stay untouched until the stream has passed the upload).  With
// sync = true the upload is waited for and `staging` may be reused at once.
// width > 0 overrides the tuned team width of every bucket.
// split > 0 overrides the split-team factor of every wide bucket.
int batch_create(cf_ctx* c, int n, const char* seqs, const int64_t* offsets, const uint8_t* masks,
                 const uint64_t* keys, hipStream_t stream, void** staging, size_t* staging_bytes, bool sync,
                 cf_batch** out, int width = 0, int split = 0);
// Release a batch.  sync = false skips waiting for the batch's stream: only
// for a batch whose work is known to be complete (its done event has
// passed) on a stream that may already hold later work.
void batch_free(cf_batch* b, bool sync);
// Bytes of pinned staging batch_create needs for these sizes.
size_t batch_staging_bytes(int n, int64_t total_residues);
// Prepare buffers and per-run inputs, then enqueue the stages of `flags` on
// the batch stream without waiting.  Host arrays are copied into `staging`
// (pinned) and uploaded asynchronously.
int batch_enqueue(cf_batch* b, unsigned flags, const float* gammas, int n_gamma, int n_samples, uint64_t seed,
                  const float* weights, void* staging, size_t staging_bytes, bool timing, unsigned* ran);
}  // namespace cfi
