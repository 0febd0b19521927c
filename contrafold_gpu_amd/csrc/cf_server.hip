// cf_server.hip -- continuous-batching fold server (no counterpart in the
// reference, which folds synchronously per call, gpu.py:480-583).
//
// Requests wait in a FIFO.  Whenever one of the server's streams is idle, the
// scheduler picks a head request (the oldest of its group with the highest
// response ratio, see launch) and takes every later pending request of the
// same kind (MEA / sample / bpp / logZ, same gamma and sample count) and the
// same length class, up to max_batch, as long as their device table bytes
// (cf_fold_bytes, O(L^2)) fit the HBM budget next to the batches already in
// flight.  The batch is packed
// and uploaded asynchronously and its kernels are enqueued on that stream, so
// batches admitted while others run overlap with them on the GPU.  Completion
// is detected with events; results are copied out and the memory returns to
// the context's pool.  A request's sampling stream is keyed by its own 64-bit
// key (cf_server_submit's `seed`), so its result does not depend on the batch
// it lands in.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "cf_internal.h"

namespace {

struct Req {
    int64_t id = 0;
    std::string seq;
    std::vector<uint8_t> mask;
    unsigned kind = CF_MEA;
    float gamma = 6.0f;
    int n_samples = 0;
    uint64_t seed = 0;
    double arrival_us = 0.0, launch_us = 0.0, done_us = 0.0;
    int state = 0;  // 0 pending, 1 in flight, 2 done, 3 failed
    int fail_code = CF_ENOMEM;  // status a failed request reports
    int64_t bytes = 0;
    std::vector<char> result;
    float logz = 0.0f;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    void* stage_in = nullptr;
    size_t stage_in_bytes = 0;
    void* stage_run = nullptr;
    size_t stage_run_bytes = 0;
    void* out = nullptr;  // pinned: logZ then the kind's payload, copied on the slot's stream
    size_t out_bytes = 0;
    double launched_us = 0.0;
    cf_batch* batch = nullptr;
    std::vector<int64_t> reqs;
    unsigned kind = 0;
    int64_t bytes = 0;
    bool long_class = false;
};

// Length classes: a batch retires when its slowest fold finishes, so a batch
// only takes requests of one class and short folds never wait on long ones.
int length_class(int L) {
    return L <= 96 ? 0 : L <= 192 ? 1 : L <= 320 ? 2 : L <= 512 ? 3 : L <= 768 ? 4 : 5;
}

unsigned kind_flags(unsigned kind) {
    switch (kind) {
        case CF_MEA: return CF_MEA;
        case CF_BPP: return CF_BPP;
        case CF_SAMPLE: return CF_SAMPLE;
        default: return CF_LOGZ;
    }
}

}  // namespace

struct cf_server {
    cf_ctx* ctx = nullptr;
    int64_t budget = 0, in_use = 0;
    int max_batch = 256;
    std::vector<Slot> slots;
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Req> reqs;
    std::deque<int64_t> completed;
    int64_t next_id = 0;
    // Pinned buffers that were outgrown.  hipHostFree waits for the whole
    // device (every batch in flight), so they are released only at destroy.
    std::vector<void*> retired_pinned;
    // statistics (cf_server_stats)
    double n_batches = 0, n_folds = 0, launch_us = 0, retire_us = 0, run_us = 0, max_pending = 0;
    std::chrono::steady_clock::time_point t0;
    double now_us() const {
        return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Results were copied to the slot's pinned buffer on the slot's stream before
// its done event, so retiring a batch makes no blocking device call (a
// synchronous copy would queue behind other batches' kernels on a shared
// hardware queue).
int retire(cf_server* s, Slot& sl) {
    cf_batch* b = sl.batch;
    const double t = s->now_us();
    s->run_us += t - sl.launched_us;
    const float* lz = (const float*)sl.out;
    const char* pay = (const char*)sl.out + sizeof(float) * (size_t)b->n;
    size_t pay_bytes = 0;
    if (sl.kind == CF_MEA) pay_bytes = (size_t)b->total_chars;
    else if (sl.kind == CF_SAMPLE) pay_bytes = (size_t)b->n_samples * (size_t)b->total_chars;
    else if (sl.kind == CF_BPP) pay_bytes = sizeof(float) * (size_t)b->total_T;
    int status = 0;
    std::memcpy(&status, pay + pay_bytes, sizeof(int));
    int64_t toff = 0;
    for (size_t k = 0; k < sl.reqs.size(); ++k) {
        Req& r = s->reqs[sl.reqs[k]];
        r.state = status ? 3 : 2;
        if (status) r.fail_code = CF_EDEVICE;
        r.done_us = t;
        r.logz = lz[k];
        const int L = b->L[k];
        if (sl.kind == CF_MEA) {
            r.result.assign(pay + b->charoff[k], pay + b->charoff[k] + L);
        } else if (sl.kind == CF_SAMPLE) {
            const int64_t off = b->charoff[k] * (int64_t)b->n_samples;
            r.result.assign(pay + off, pay + off + (int64_t)b->n_samples * L);
        } else if (sl.kind == CF_BPP) {
            // POST is indexed by FC cell: entry (a, c-1) holds P(a pairs c)
            const float* tri = (const float*)pay + toff;
            r.result.assign((size_t)L * L * sizeof(float), 0);
            float* m = (float*)r.result.data();
            for (int a = 1; a <= L; ++a)
                for (int c = a + 1; c <= L; ++c) m[(size_t)(a - 1) * L + (c - 1)] = tri[cf_tri(L, a, c - 1)];
        }
        toff += cf_tri_size(L);
        s->completed.push_back(r.id);
    }
    cf_batch_destroy(b);
    sl.batch = nullptr;
    sl.reqs.clear();
    s->in_use -= sl.bytes;
    sl.bytes = 0;
    s->retire_us += s->now_us() - t;
    return CF_OK;
}

// Grows a slot's pinned buffer to at least `need` bytes (powers of two from
// 1 MiB); the old buffer is parked until the server is destroyed.
bool grow_pinned(cf_server* s, void** p, size_t* bytes, size_t need) {
    if (*bytes >= need) return true;
    size_t cap = (size_t)1 << 20;
    while (cap < need) cap <<= 1;
    void* q = nullptr;
    if (hipHostMalloc(&q, cap) != hipSuccess) return false;
    if (*p) s->retired_pinned.push_back(*p);
    *p = q;
    *bytes = cap;
    return true;
}

// Workgroup width for a small batch whose longest fold has length Lmax: each
// fold gets a whole workgroup sized to its diagonals (latency over packing).
int serve_width(int Lmax) { return Lmax <= 100 ? 128 : Lmax <= 250 ? 256 : Lmax <= 500 ? 512 : 1024; }

// Folds of length class >= kLongClass take long; at most half of the slots
// may hold them, so short requests always find a slot.
constexpr int kLongClass = 3;
// Typical device time (us) of a small batch of each length class on a loaded
// GPU (measured run time p50 of the serving mix: <= 100 nt 10 ms, 101-200 nt
// 45 ms, 201-400 nt 60 ms, 401-700 nt 190 ms, longer 250 ms).
constexpr double kClassUs[6] = {10e3, 40e3, 60e3, 150e3, 200e3, 250e3};

int launch(cf_server* s, Slot& sl) {
    cf_ctx* c = s->ctx;
    const double t_start = s->now_us();
    s->max_pending = std::max(s->max_pending, (double)s->pending.size());
    while (!s->pending.empty() && s->reqs[s->pending.front()].state != 0) s->pending.pop_front();
    if (s->pending.empty()) return 0;
    // Which group (kind, gamma, sample count, length class) goes next: the one
    // whose oldest request has the highest response ratio, its wait over its
    // class's typical fold time, so short folds do not queue behind a long
    // batch's turn and a long request's priority still grows while it waits.
    // Long classes are skipped while they hold their share of the slots.
    int long_busy = 0;
    for (const auto& o : s->slots)
        if (o.batch && o.long_class) ++long_busy;
    const bool long_ok = long_busy < (int)s->slots.size() / 2;
    const double now = s->now_us();
    int64_t head_id = -1;
    double best = -1.0;
    std::vector<int64_t> seen;  // the oldest request of each group met so far
    for (int64_t id : s->pending) {
        const Req& r = s->reqs[id];
        if (r.state != 0) continue;
        const int lc = length_class((int)r.seq.size());
        if (!long_ok && lc >= kLongClass) continue;
        bool older = false;
        for (int64_t o : seen) {
            const Req& q = s->reqs[o];
            if (q.kind == r.kind && q.gamma == r.gamma && q.n_samples == r.n_samples &&
                length_class((int)q.seq.size()) == lc) {
                older = true;
                break;
            }
        }
        if (older) continue;
        seen.push_back(id);
        const double ratio = (now - r.arrival_us + kClassUs[lc]) / kClassUs[lc];
        if (ratio > best) {
            best = ratio;
            head_id = id;
        }
    }
    if (head_id < 0) return 0;
    const Req& head = s->reqs[head_id];
    const unsigned kind = head.kind;
    const float gamma = head.gamma;
    const int nsamp = head.n_samples;
    const int lclass = length_class((int)head.seq.size());
    if (head.bytes > s->budget) {  // can never fit: fail it
        Req& r = s->reqs[head.id];
        r.state = 3;
        r.done_us = s->now_us();
        s->completed.push_back(r.id);
        // leave the queue now: cf_server_result erases the request
        s->pending.erase(std::find(s->pending.begin(), s->pending.end(), head_id));
        return 0;
    }
    std::vector<int64_t> pick;
    int64_t bytes = 0;
    std::deque<int64_t> rest;
    bool blocked = false;
    for (int64_t id : s->pending) {
        Req& r = s->reqs[id];
        const bool same = r.kind == kind && r.gamma == gamma && r.n_samples == nsamp &&
                          length_class((int)r.seq.size()) == lclass;
        if (!blocked && same && (int)pick.size() < s->max_batch && s->in_use + bytes + r.bytes <= s->budget) {
            pick.push_back(id);
            bytes += r.bytes;
        } else {
            if (same && !pick.empty()) blocked = blocked || (s->in_use + bytes + r.bytes > s->budget);
            rest.push_back(id);
        }
    }
    if (pick.empty()) return 0;  // waiting for memory
    s->pending.swap(rest);
    // pack
    const int n = (int)pick.size();
    std::string seqs;
    std::vector<int64_t> offs(n + 1, 0);
    std::vector<uint8_t> masks;
    std::vector<uint64_t> keys(n);
    bool any_mask = false;
    for (int k = 0; k < n; ++k) {
        const Req& r = s->reqs[pick[k]];
        any_mask = any_mask || !r.mask.empty();
        offs[k + 1] = offs[k] + (int64_t)r.seq.size();
    }
    seqs.reserve(offs[n]);
    if (any_mask) masks.assign(offs[n], 0);
    for (int k = 0; k < n; ++k) {
        Req& r = s->reqs[pick[k]];
        seqs += r.seq;
        if (!r.mask.empty()) std::memcpy(masks.data() + offs[k], r.mask.data(), r.mask.size());
        keys[k] = r.seed;  // the request's own stream key: independent of the batch it joins
        r.state = 1;
        r.launch_us = s->now_us();
    }
    // A small batch cannot fill the GPU with one wave per fold; giving each
    // fold a whole workgroup cuts its latency instead.
    int lmax = 0;
    for (int k = 0; k < n; ++k) lmax = std::max(lmax, (int)(offs[k + 1] - offs[k]));
    const int width = (n <= 4 * c->n_cu) ? serve_width(lmax) : 0;
    cf_batch* b = nullptr;
    int rc = CF_OK;
    // the slot's previous batch has retired, so its pinned buffers are idle
    if (!grow_pinned(s, &sl.stage_in, &sl.stage_in_bytes, cfi::batch_staging_bytes(n, offs[n] + 2 * (int64_t)n)))
        rc = cfi::fail(c, CF_ENOMEM, "pinned staging");
    if (!rc)
        rc = cfi::batch_create(c, n, seqs.data(), offs.data(), any_mask ? masks.data() : nullptr, keys.data(),
                               sl.stream, &sl.stage_in, &sl.stage_in_bytes, false, &b, width);
    if (!rc) {
        const size_t need = sizeof(float) * (cfi::kMaxGamma + (size_t)n) + sizeof(int64_t) * (size_t)n + 4096;
        if (!grow_pinned(s, &sl.stage_run, &sl.stage_run_bytes, need)) rc = cfi::fail(c, CF_ENOMEM, "pinned staging");
    }
    if (!rc)
        rc = cfi::batch_enqueue(b, kind_flags(kind), &gamma, 1, nsamp > 0 ? nsamp : 1, 0ull, nullptr,
                                sl.stage_run, sl.stage_run_bytes, false, nullptr);
    if (!rc) {
        // results -> pinned host memory, in stream order after the kernels
        const void* src = nullptr;
        size_t pay = 0;
        if (kind == CF_MEA) src = b->mea_out, pay = (size_t)b->total_chars;
        else if (kind == CF_SAMPLE) src = b->samp_out, pay = (size_t)b->n_samples * b->total_chars;
        else if (kind == CF_BPP) src = b->POST, pay = sizeof(float) * (size_t)b->total_T;
        // every batch brings back its kernel status word
        const size_t stat = sizeof(int);
        const size_t need = sizeof(float) * (size_t)n + pay + stat;
        if (!grow_pinned(s, &sl.out, &sl.out_bytes, need)) rc = cfi::fail(c, CF_ENOMEM, "pinned result buffer");
        if (!rc && hipMemcpyAsync(sl.out, b->logz, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, sl.stream) !=
                       hipSuccess)
            rc = cfi::fail(c, CF_EDEVICE, "result copy");
        if (!rc && pay &&
            hipMemcpyAsync((char*)sl.out + sizeof(float) * (size_t)n, src, pay, hipMemcpyDeviceToHost, sl.stream) !=
                hipSuccess)
            rc = cfi::fail(c, CF_EDEVICE, "result copy");
        if (!rc &&
            hipMemcpyAsync((char*)sl.out + sizeof(float) * (size_t)n + pay, b->d_status, stat, hipMemcpyDeviceToHost,
                           sl.stream) != hipSuccess)
            rc = cfi::fail(c, CF_EDEVICE, "result copy");
    }
    if (!rc && hipEventRecord(sl.done, sl.stream) != hipSuccess) rc = cfi::fail(c, CF_EDEVICE, "event record");
    if (rc) {
        if (b) cf_batch_destroy(b);
        for (int64_t id : pick) {
            Req& r = s->reqs[id];
            r.state = 3;
            r.done_us = s->now_us();
            s->completed.push_back(id);
        }
        return rc;
    }
    sl.batch = b;
    sl.reqs = pick;
    sl.kind = kind;
    sl.long_class = lclass >= kLongClass;
    sl.bytes = bytes;
    s->in_use += bytes;
    sl.launched_us = s->now_us();
    s->n_batches += 1;
    s->n_folds += n;
    s->launch_us += sl.launched_us - t_start;
    return n;
}

}  // namespace

extern "C" {

int cf_server_create(cf_ctx* c, int64_t budget, int max_batch, int n_streams, cf_server** out) {
    if (!c || !out || budget <= 0 || max_batch <= 0 || n_streams <= 0 || n_streams > 32)
        return cfi::fail(c, CF_EINVAL, "bad server arguments");
    *out = nullptr;
    if (hipSetDevice(c->device) != hipSuccess) return cfi::fail(c, CF_EDEVICE, "hipSetDevice");
    cf_server* s = new cf_server();
    s->ctx = c;
    s->budget = budget;
    s->max_batch = max_batch;
    s->t0 = std::chrono::steady_clock::now();
    s->slots.resize(n_streams);
    for (auto& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            cf_server_destroy(s);
            return cfi::fail(c, CF_EDEVICE, "stream/event creation failed");
        }
    }
    *out = s;
    return CF_OK;
}

int64_t cf_server_submit(cf_server* s, const char* seq, int L, const uint8_t* mask, unsigned kind, float gamma,
                         int n_samples, uint64_t seed, double arrival_us) {
    if (!s || L < 0 || (L > 0 && !seq)) return CF_EINVAL;
    if (kind != CF_MEA && kind != CF_BPP && kind != CF_SAMPLE && kind != CF_LOGZ)
        return cfi::fail(s->ctx, CF_EINVAL, "kind must be CF_MEA, CF_BPP, CF_SAMPLE or CF_LOGZ");
    if (kind == CF_MEA && !(gamma > 0.0f)) return cfi::fail(s->ctx, CF_EINVAL, "gamma must be > 0");
    if (kind == CF_SAMPLE && n_samples <= 0) return cfi::fail(s->ctx, CF_EINVAL, "n_samples must be > 0");
    if (L > 32000) return cfi::fail(s->ctx, CF_EINVAL, "sequence too long");
    Req r;
    r.id = s->next_id++;
    r.seq.assign(seq ? seq : "", (size_t)L);
    if (mask) r.mask.assign(mask, mask + L);
    r.kind = kind;
    r.gamma = kind == CF_MEA ? gamma : 0.0f;
    r.n_samples = kind == CF_SAMPLE ? n_samples : 0;
    r.seed = seed;
    r.arrival_us = arrival_us > 0.0 ? arrival_us : s->now_us();
    r.bytes = cf_fold_bytes(L, kind_flags(kind), 1, r.n_samples);
    const int64_t id = r.id;
    s->reqs.emplace(id, std::move(r));
    s->pending.push_back(id);
    return id;
}

int cf_server_poll(cf_server* s, int64_t* ids, int cap) {
    if (!s) return CF_EINVAL;
    for (auto& sl : s->slots) {
        if (sl.batch && hipEventQuery(sl.done) == hipSuccess) retire(s, sl);
    }
    for (auto& sl : s->slots) {
        if (!sl.batch && !s->pending.empty()) launch(s, sl);
    }
    int k = 0;
    while (k < cap && !s->completed.empty()) {
        if (ids) ids[k] = s->completed.front();
        s->completed.pop_front();
        ++k;
    }
    return k;
}

int cf_server_result(cf_server* s, int64_t id, char* structure, float* logz, double* latency_us) {
    if (!s) return CF_EINVAL;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return cfi::fail(s->ctx, CF_EINVAL, "unknown request id");
    Req& r = it->second;
    if (r.state == 3) {
        const int code = r.fail_code;
        s->reqs.erase(it);
        return cfi::fail(s->ctx, code,
                         code == CF_ENOMEM ? "request failed (does not fit the HBM budget or device error)"
                                           : "request failed (device status: sampler stack overflow or "
                                             "split-team fold not co-resident)");
    }
    if (r.state != 2) return cfi::fail(s->ctx, CF_ESTATE, "request not complete");
    if (structure && !r.result.empty()) std::memcpy(structure, r.result.data(), r.result.size());
    if (logz) *logz = r.logz;
    if (latency_us) *latency_us = r.done_us - r.arrival_us;
    s->reqs.erase(it);
    return CF_OK;
}

int cf_server_stats(const cf_server* s, double* out, int n) {
    if (!s || !out || n < 0) return CF_EINVAL;
    const double v[6] = {s->n_batches, s->n_folds, s->launch_us, s->retire_us, s->run_us, s->max_pending};
    for (int k = 0; k < n && k < 6; ++k) out[k] = v[k];
    return std::min(n, 6);
}

int cf_server_pending(const cf_server* s) {
    if (!s) return 0;
    int n = (int)s->pending.size();
    for (const auto& sl : s->slots) n += (int)sl.reqs.size();
    return n;
}

int cf_server_replay(cf_server* s, int n, const char* seqs, const int64_t* offsets, const uint8_t* masks,
                     const unsigned* kinds, const float* gammas, const int* n_samples, const double* arrival_us,
                     uint64_t seed, double* latency_us, double* wait_us, double* wall_s) {
    if (!s || n <= 0 || !offsets || !kinds || !arrival_us || !latency_us) return CF_EINVAL;
    std::unordered_map<int64_t, int> which;
    const double start = s->now_us();
    int next = 0, done = 0;
    std::vector<int64_t> ids(1024);
    while (done < n) {
        const double t = s->now_us() - start;
        while (next < n && arrival_us[next] <= t) {
            const int L = (int)(offsets[next + 1] - offsets[next]);
            const uint8_t* m = masks ? masks + offsets[next] : nullptr;
            bool has = false;
            for (int q = 0; m && q < L; ++q) has = has || m[q];
            const int64_t id = cf_server_submit(s, seqs + offsets[next], L, has ? m : nullptr, kinds[next],
                                                gammas ? gammas[next] : 6.0f, n_samples ? n_samples[next] : 0,
                                                seed + (uint64_t)next, start + arrival_us[next]);
            if (id < 0) return (int)id;
            which[id] = next;
            ++next;
        }
        const int k = cf_server_poll(s, ids.data(), (int)ids.size());
        for (int q = 0; q < k; ++q) {
            double lat = -1.0;
            const auto it = s->reqs.find(ids[q]);
            const double wait = (it != s->reqs.end() && it->second.state == 2)
                                    ? it->second.launch_us - it->second.arrival_us : -1.0;
            const int rc = cf_server_result(s, ids[q], nullptr, nullptr, &lat);
            latency_us[which[ids[q]]] = rc ? -1.0 : lat;
            if (wait_us) wait_us[which[ids[q]]] = rc ? -1.0 : wait;
            ++done;
        }
        if (k == 0) std::this_thread::yield();
    }
    if (wall_s) *wall_s = (s->now_us() - start) * 1e-6;
    return CF_OK;
}

void cf_server_destroy(cf_server* s) {
    if (!s) return;
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.batch) cf_batch_destroy(sl.batch);
        if (sl.stage_in) (void)hipHostFree(sl.stage_in);
        if (sl.stage_run) (void)hipHostFree(sl.stage_run);
        if (sl.out) (void)hipHostFree(sl.out);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    for (void* p : s->retired_pinned) (void)hipHostFree(p);
    delete s;
}

}  // extern "C"
