// cf_kernels.hip -- gfx950 kernels of the CONTRAfold fold engine.
//
// Work decomposition
//   * one workgroup per sequence for the O(L^2)-cell sweeps (inside, outside +
//     posterior, MEA, Viterbi).  A sweep walks anti-diagonals s = j - i; every
//     cell of a diagonal depends only on earlier diagonals, so lanes take the
//     cells of one diagonal and a barrier separates diagonals.
//   * the FC cells of a diagonal that can pair (canonical, not forced) are
//     compacted into a list first, so waves that run the long single-branch
//     window (<= 496 candidates) carry pairable cells on every lane.
//   * inside a cell, the candidates of the single-branch window are walked as
//     set bits of per-base partner masks: a lane touches only pairable inner
//     pairs, in the reference's (p ascending, q descending) order.
//   * the exterior chains F5 / F5o are strictly serial per sequence; one lane
//     per sequence runs them, 64 sequences per wave.
//   * sampling: one lane per (sequence, sample) over a persistent grid.
//
// Bit-exactness: each cell folds its terms with cf_lse in exactly the order
// of the reference's float32 engine (gpu_contrafold/cpu.py:583-895), with the
// same left-to-right additions; nothing is reassociated and -ffp-contract=off
// keeps every multiply-add as two rounded operations.
#include "cf_device.h"

#define WAVE 64

// ------------------------------------------------------------------ LDS plan
// A workgroup runs G "teams", each folding one sequence.  G = 1 is the wide
// layout (the whole workgroup on one sequence, __syncthreads between
// diagonals); G = blockDim/64 is the packed layout (one wave per sequence,
// the waves of a workgroup share the staged parameter tables and never wait
// for each other after staging).  LDS = parameter tables once + G team areas.
struct CfLds {
    int W;  // 32-bit words per partner-mask row
    // shared area (offsets from the start of LDS)
    int o_lse, o_exp, o_cs, o_jb, o_bp, o_bj, shared;
    // team area (offsets from the team's base)
    int o_qinfo, o_dtab, o_qo, o_s8, o_fo8, o_rs8, o_pmask, o_dbuf, o_alist, o_cnt, o_ring, team;
    int G, total;
};
__host__ __device__ inline int cf_align(int x, int a) { return (x + a - 1) / a * a; }
__host__ __device__ inline CfLds cf_lds_plan(int Lmax, bool ring, int dbuf_elem, int G = 1) {
    CfLds p;
    p.W = ((Lmax + 2 + 31) >> 5) + 1;
    int o = 0;
    p.o_lse = o; o += CF_LSE_TABLE_BYTES;  // coefficients, then the segment bins
    p.o_exp = o; o += 8 * 16;
    p.o_cs = o; o += 961 * 4;
    p.o_jb = o; o += 625 * 4;
    p.o_bp = o; o += 25 * 4;
    o = cf_align(o, 16);
    p.o_bj = o; o += 625 * 8;
    p.shared = cf_align(o, 16);
    o = 0;
    p.o_qinfo = o; o += (Lmax + 3) * 4;
    p.o_dtab = o; o += (Lmax + 1) * 4;
    p.o_qo = o; o += (Lmax + 3) * 2;
    o = cf_align(o, 4);
    p.o_s8 = o; o += Lmax + 3;
    p.o_fo8 = o; o += Lmax + 3;
    p.o_rs8 = o; o += Lmax + 3;
    o = cf_align(o, 16);
    p.o_pmask = o; o += 5 * p.W * 4;
    o = cf_align(o, 16);
    p.o_dbuf = o; o += (Lmax + 1 > 128 ? Lmax + 1 : 128) * dbuf_elem;  // the posterior queue (0: none)
    o = cf_align(o, 16);
    p.o_alist = o; o += 2 * (Lmax + 1) * 2;  // two diagonals' pairable cells
    o = cf_align(o, 16);
    p.o_cnt = o; o += 16;
    p.o_ring = o;
    if (ring) o += CF_RING * (Lmax + 1) * 4;
    p.team = cf_align(o, 16);
    p.G = G;
    p.total = p.shared + G * p.team;
    return p;
}

struct CfShared {
    float4* lse;
    float4* ex;
    float* cs;
    float* JB;
    float* bp;
    float2* BJ;  // (JB[q][p+1][q+1][p], bp[p+1][q]) for one 8-byte read per candidate
    int* qinfo;
    int* dtab;  // dtab[d] = dbase(L, d), the offset of diagonal d in a triangular table
    int16_t* qo;  // qo[q] = 25*s[q+1] + s[q]: the (q, q+1) part of an outer-pair junction index
    int8_t* s8;
    int8_t* fo8;
    int8_t* rs8;
    uint32_t* pmask;
    int W;
    void* dbuf;
    uint16_t* alist;
    int* cnt;
    void* ring;
};

__device__ inline CfShared cf_carve(unsigned char* smem, const CfLds& p, int team = 0) {
    CfShared S;
    S.lse = (float4*)(smem + p.o_lse);
    S.ex = (float4*)(smem + p.o_exp);
    S.cs = (float*)(smem + p.o_cs);
    S.JB = (float*)(smem + p.o_jb);
    S.bp = (float*)(smem + p.o_bp);
    S.BJ = (float2*)(smem + p.o_bj);
    unsigned char* t = smem + p.shared + team * p.team;
    S.qinfo = (int*)(t + p.o_qinfo);
    S.dtab = (int*)(t + p.o_dtab);
    S.qo = (int16_t*)(t + p.o_qo);
    S.s8 = (int8_t*)(t + p.o_s8);
    S.fo8 = (int8_t*)(t + p.o_fo8);
    S.rs8 = (int8_t*)(t + p.o_rs8);
    S.pmask = (uint32_t*)(t + p.o_pmask);
    S.W = p.W;
    S.dbuf = (void*)(t + p.o_dbuf);
    S.alist = (uint16_t*)(t + p.o_alist);
    S.cnt = (int*)(t + p.o_cnt);
    S.ring = (void*)(t + p.o_ring);
    return S;
}

// Synchronise the threads of one team.  A packed team is a single wave: its
// lanes run in lockstep, so ordering its own LDS and global accesses needs
// only the fences (waits for outstanding LDS operations and keeps the
// compiler from moving memory accesses across); a wide team is the
// workgroup.
template <bool PACK>
__device__ __forceinline__ void cf_team_sync() {
    if (PACK) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __syncthreads();
    }
}

// Stage the hot parameter tables in LDS (every thread of the workgroup takes
// part; the caller synchronises the workgroup afterwards).
__device__ void cf_stage_params(const CfShared& S, const CfParams<float>* __restrict__ P) {
    const int tid = threadIdx.x, NT = blockDim.x;
    for (int t = tid; t < CF_LSE_BINS; t += NT) cf_lse_stage(S.lse, t);
    for (int t = tid; t < 6; t += NT) S.ex[t] = make_float4(cf_exp_coef[t][0], cf_exp_coef[t][1], cf_exp_coef[t][2], cf_exp_coef[t][3]);
    for (int t = tid; t < 961; t += NT) S.cs[t] = P->t[CF_O_CS + t];
    for (int t = tid; t < 625; t += NT) S.JB[t] = P->JB[t];
    for (int t = tid; t < 25; t += NT) S.bp[t] = P->t[CF_O_BP + t];
    for (int t = tid; t < 625; t += NT) {
        // t = ((sq*5 + sp1)*5 + sq1)*5 + sp
        const int sq = t / 125, sp1 = (t / 25) % 5;
        S.BJ[t] = make_float2(P->JB[t], P->t[CF_O_BP + cf_i2(sp1, sq)]);
    }
}

// Stage one sequence (codes, forced mask, partner masks) in the team's area.
template <bool PACK>
__device__ void cf_stage_seq(const CfShared& S, const uint8_t* __restrict__ codes, const uint8_t* __restrict__ masks,
                             const CfSeq& sd, int tid, int NT) {
    const int L = sd.L;
    for (int q = tid; q < L + 2; q += NT) {
        S.s8[q] = (int8_t)codes[sd.coff + q];
        S.fo8[q] = (masks != nullptr && masks[sd.coff + q] != 0) ? 1 : 0;
    }
    if (tid == 0) { S.s8[L + 2] = 4; S.fo8[L + 2] = 0; S.cnt[0] = 0; S.cnt[1] = 0; }
    cf_team_sync<PACK>();
    for (int d = tid; d <= L; d += NT) S.dtab[d] = (int)cf_dbase(L, d);
    for (int q = tid; q < L + 2; q += NT) {
        const int c = S.s8[q];
        S.rs8[q] = (S.fo8[q] || c == 4) ? 4 : c;
        S.qinfo[q] = (c << 16) | (c * 125 + S.s8[q + 1] * 5);
        S.qo[q] = (int16_t)(S.s8[q + 1] * 25 + c);
    }
    for (int w = tid; w < S.W; w += NT) {
        uint32_t m[5] = {0u, 0u, 0u, 0u, 0u};
        for (int bit = 0; bit < 32; ++bit) {
            const int q = w * 32 + bit;
            if (q > L + 1) break;
            const int c = S.s8[q];
            if (S.fo8[q] || c == 4) continue;
            for (int x = 0; x < 4; ++x)
                if (cf_canon(x, c)) m[x] |= 1u << bit;
        }
        for (int x = 0; x < 5; ++x) S.pmask[x * S.W + w] = m[x];
    }
    cf_team_sync<PACK>();
}

// Team geometry of a launch: which sequence this thread's team folds, and the
// thread's rank and count within the team.  The wave index is made
// wave-uniform explicitly so that everything derived from the sequence
// descriptor (table bases, lengths, loop bounds) lives in scalar registers.
//
// Split teams (K > 1, wide layout only): K workgroups fold one sequence
// together, each on its own CU, so a long fold's anti-diagonals are issued by
// 4K SIMDs instead of the 4 of one CU.  tid / NT index the cells of the whole
// team (K * blockDim threads); ltid / lNT index the workgroup, which stages
// its own copy of the sequence in LDS and runs its own pairable-cell list.
// Workgroups are dispatched to the 8 XCDs round-robin by launch index, so the
// K parts of a fold take launch indices that are equal modulo 8 -- one XCD,
// one L2 -- and a block of 8 K consecutive indices holds 8 folds.
struct CfTeam {
    int team, tid, NT, li, ltid, lNT, part, K;
};
template <bool PACK>
__device__ __forceinline__ CfTeam cf_team(int G, int K = 1) {
    CfTeam t;
    t.ltid = threadIdx.x;
    t.lNT = blockDim.x;
    t.part = 0;
    t.K = 1;
    if (PACK) {
        t.team = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
        t.tid = threadIdx.x & (WAVE - 1);
        t.NT = WAVE;
        t.li = blockIdx.x * G + t.team;
        t.ltid = t.tid;
        t.lNT = WAVE;
    } else if (K > 1) {
        const int g = blockIdx.x, x = g & 7, r = g >> 3;
        t.team = 0;
        t.K = K;
        t.part = r % K;
        t.li = (r / K) * 8 + x;
        t.tid = t.part * blockDim.x + threadIdx.x;
        t.NT = K * blockDim.x;
    } else {
        t.team = 0;
        t.tid = threadIdx.x;
        t.NT = blockDim.x;
        t.li = blockIdx.x;
    }
    return t;
}
// Launch size of a split launch over n folds: whole blocks of 8 folds.
__host__ __device__ inline int cf_split_grid(int n, int K) { return ((n + 7) / 8) * 8 * K; }

// Persistent packed launches: a grid sized to what the device holds at once,
// whose one-wave teams take folds from a queue in launch order (longest
// first) until it is empty -- a team that finishes early takes the next fold,
// so uneven lengths need no cut into stream parts.  Lane 0 takes the ticket.
__device__ __forceinline__ int cf_next_fold(int* queue) {
    int v = 0;
    if ((threadIdx.x & (WAVE - 1)) == 0) v = atomicAdd(queue, 1);
    return __builtin_amdgcn_readfirstlane(v);
}

// Barrier of a split team: every workgroup of the fold arrives on the fold's
// counter and waits until all K have arrived for this generation.  Each wave
// first waits for its own loads and stores to complete; thread 0 then
// publishes with an agent-scope release and spins with agent-scope acquires,
// so the other workgroups' table writes are visible after the barrier.  The
// wait is bounded (about 2 s of wall clock): if the parts of a fold are not
// resident together, the part that gives up raises the fold's abort word
// (ctr[1]) and status bit 2; every part then stops waiting (it still
// arrives), runs to its end, and the host reports the failure.
#define CF_XBAR_LIMIT 200000000ull  // wall-clock ticks (100 MHz): 2 s
struct CfXBar {
    unsigned* ctr;  // this fold's arrival counter, then its abort word (zeroed before the launch)
    int* status;
    int* abort;     // LDS flag shared by the workgroup
    unsigned gen;
    int K;
};
__device__ __forceinline__ void cf_xbar(CfXBar& x) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's loads and stores have completed
    __syncthreads();
    x.gen += 1;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(x.ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!*x.abort) {
            const unsigned target = x.gen * (unsigned)x.K;
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(x.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (__hip_atomic_load(x.ctr + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    *x.abort = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - t0 > CF_XBAR_LIMIT) {
                    *x.abort = 1;
                    __hip_atomic_store(x.ctr + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    atomicOr(x.status, 2);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

__device__ __forceinline__ bool cf_allow(const CfShared& S, int a, int b) {
    return cf_canon(S.s8[a], S.s8[b]) && !S.fo8[a] && !S.fo8[b];
}

// ==================================================================== inside
// FP32 inside sweep (cpu.py:595-665; gpu.py:79-168).  Writes FC, FM1, FM,
// FM2 (the multiloop split, kept for sampling and gradients) and the exterior
// copy FCX of FC.  F5 is produced by cf_exterior_kernel.
//
// Diagonals are taken two at a time.  FM2 and FC of diagonal d read FM1 / FM
// / FC of diagonals <= d-2 only (and FC reads FM2 of its own cell), so FM2 and
// FC of diagonals s and s+1 are computed together -- twice the cells per pass
// over the team's lanes -- while FM1 and FM, which read diagonal d-1, follow
// one diagonal at a time:
//   A: FM2 of every cell of s and s+1; pairable cells of both into a list
//   B: FC of the listed cells
//   C: FM1, FM of diagonal s       D: FM1, FM of diagonal s+1
// Each cell's terms are folded in the reference's order, so every table is
// bit-identical to a one-diagonal-at-a-time sweep.
// LAT: the split-team (latency) form: 256-thread workgroups, one wave per
// SIMD, so a wave has the register room to keep the next batch of loads in
// flight while it folds the current one (the 1,024-thread form would spill).
template <bool PACK, bool LAT>
__global__ void __launch_bounds__(LAT ? 256 : 1024) cf_inside_kernel(const CfSeq* __restrict__ seqs, const int* __restrict__ order,
                                                        int nseq, const uint8_t* __restrict__ codes,
                                                        const uint8_t* __restrict__ masks,
                                                        const CfParams<float>* __restrict__ P, float* __restrict__ FC,
                                                        float* __restrict__ FM, float* __restrict__ FM1,
                                                        float* __restrict__ FM2, float* __restrict__ FCX, int Lmax,
                                                        int K, unsigned* __restrict__ xsync,
                                                        int* __restrict__ status, int* __restrict__ queue) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int xabort;
    const int G = PACK ? (int)(blockDim.x / WAVE) : 1;
    const CfTeam tm = cf_team<PACK>(G, PACK ? 1 : K);
    const CfLds plan = cf_lds_plan(Lmax, false, 0, G);
    const CfShared S = cf_carve(smem, plan, tm.team);
    cf_stage_params(S, P);
    if (threadIdx.x == 0) xabort = 0;
    __syncthreads();
    // queue (packed teams only): take folds until the queue is empty
    const bool persistent = PACK && queue != nullptr;
    for (int li = persistent ? cf_next_fold(queue) : tm.li; li < nseq;
         li = persistent ? cf_next_fold(queue) : nseq) {
    if (persistent) cf_team_sync<PACK>();  // the team's LDS area is reused fold after fold
    const CfSeq sd = seqs[order[li]];
    const int L = sd.L;
    const int tid = tm.tid, NT = tm.NT;
    cf_stage_seq<PACK>(S, codes, masks, sd, tm.ltid, tm.lNT);
    // team-wide sync (every part of a split team) and workgroup-local sync
    CfXBar xb{tm.K > 1 ? xsync + (size_t)li * 16 : nullptr, status, &xabort, 0u, tm.K};
    auto sync_team = [&]() {
        if (tm.K > 1) cf_xbar(xb);
        else cf_team_sync<PACK>();
    };

    const float mb = P->t[CF_O_MB], mu = P->t[CF_O_MU], mp = P->t[CF_O_MP];
    const float* __restrict__ T = P->t;
    float* __restrict__ fc_t = FC + sd.toff;
    float* __restrict__ fm_t = FM + sd.toff;
    float* __restrict__ fm1_t = FM1 + sd.toff;
    float* __restrict__ fm2_t = FM2 + sd.toff;
    float* __restrict__ fcx_t = FCX + sd.xoff;
    const float4* __restrict__ C = S.lse;

    for (int s = 0; s <= L; s += 2) {
        const int step = s >> 1;
        const int n0 = L + 1 - s;
        const int n1 = (s + 1 <= L) ? L - s : 0;
        // ---- A: multiloop split FM2 of diagonals s and s+1; collect pairable cells.
        // A lane takes row i of both diagonals: cells (i, i+s) and (i, i+s+1)
        // fold the same FM1[i][i+t] against FM[i+t][i+s] and FM[i+t][i+s+1],
        // so the row is read once for two cells (the second column is the
        // neighbouring lane's first, one step earlier, and hits the cache).
        for (int i = tid; i < n0; i += NT) {
            const int j0 = i + s;
            const bool inner0 = (i > 0) && (j0 < L);
            const bool inner1 = (i < n1) && (i > 0) && (j0 + 1 < L);  // implies inner0
            float a0 = CF_NEG, a1 = CF_NEG;
            if (inner0 && s >= 3) {
                // FM2[i][j] = LSE_{k=i+2..j-2} FM1[i][k] + FM[k][j]  (k = i + t):
                // t = 2..s-2 for both cells, then t = s-1 for (i, i+s+1) alone.
                // Loads of 8 consecutive terms are issued before their folds, so
                // memory latency overlaps; each cell's fold order is unchanged.
                int ia = S.dtab[2] + i;
                int ib0 = S.dtab[s - 2] + i + 2;
                int ib1 = S.dtab[s - 1] + i + 2;
                int t = 2;
                int tl = 2;  // split offset of the next load
                auto load8 = [&](float* x, float* y0, float* y1) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        x[u] = cf_ld(fm1_t, ia);
                        y0[u] = cf_ld(fm_t, ib0);
                        y1[u] = cf_ld(fm_t, ib1);
                        ia += L + 1 - (tl + u);
                        ib0 -= L - (s - (tl + u) - 1);
                        ib1 -= L - (s - (tl + u));
                    }
                    tl += 8;
                };
                auto fold8 = [&](const float* x, const float* y0, const float* y1) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        a0 = cf_lse(a0, x[u] + y0[u], C);
                        a1 = cf_lse(a1, x[u] + y1[u], C);
                    }
                };
                const int nfull = (s - 3) >> 3;  // batches of 8 of the t = 2 .. s-2 terms
                if constexpr (LAT) {
                    // A long fold's last diagonals have few cells, each a chain of
                    // hundreds of terms: the next batch's loads are issued before
                    // the current batch is folded, so one wave hides its own load
                    // latency (two buffers, ping-pong; fold order unchanged).
                    float xa[8], ya0[8], ya1[8], xb[8], yb0[8], yb1[8];
                    if (nfull > 0) load8(xa, ya0, ya1);
                    for (int k = 0; k < nfull; k += 2) {
                        if (k + 1 < nfull) load8(xb, yb0, yb1);
                        fold8(xa, ya0, ya1);
                        if (k + 1 >= nfull) break;
                        if (k + 2 < nfull) load8(xa, ya0, ya1);
                        fold8(xb, yb0, yb1);
                    }
                } else {
                    for (int k = 0; k < nfull; ++k) {
                        float x[8], y0[8], y1[8];
                        load8(x, y0, y1);
                        fold8(x, y0, y1);
                    }
                }
                t += 8 * nfull;
                for (; t <= s - 2; ++t) {
                    const float x = cf_ld(fm1_t, ia);
                    const float y0 = cf_ld(fm_t, ib0);
                    const float y1 = cf_ld(fm_t, ib1);
                    a0 = cf_lse(a0, x + y0, C);
                    a1 = cf_lse(a1, x + y1, C);
                    ia += L + 1 - t;
                    ib0 -= L - (s - t - 1);
                    ib1 -= L - (s - t);
                }
                if (inner1) {
                    const float x = cf_ld(fm1_t, ia);
                    const float y1 = cf_ld(fm_t, ib1);
                    a1 = cf_lse(a1, x + y1, C);
                }
            }
#pragma unroll
            for (int hi = 0; hi < 2; ++hi) {
                if (hi == 1 && i >= n1) break;
                const int d = s + hi;
                const int j = j0 + hi;
                const bool inner = hi ? inner1 : inner0;
                const int db = S.dtab[d];
                fm2_t[db + i] = inner ? (hi ? a1 : a0) : CF_NEG;
                if (inner && cf_allow(S, i, j + 1)) {
                    const int slotn = atomicAdd(&S.cnt[step & 1], 1);
                    S.alist[slotn] = (uint16_t)(i | (hi << 15));
                } else {
                    fc_t[db + i] = CF_NEG;
                    if (inner) fcx_t[cf_xbase(j + 1) + i - 1] = CF_NEG;
                }
            }
        }
        cf_team_sync<PACK>();
        // ---- B: FC of the pairable cells (hairpin, single-branch window, multiloop);
        // a split team's workgroups each take the cells they listed in A
        const int nal = S.cnt[step & 1];
        if (tm.ltid == 0) S.cnt[(step + 1) & 1] = 0;
        for (int idx = tm.ltid; idx < nal; idx += tm.lNT) {
            const int e0 = S.alist[idx];
            const int i = e0 & 0x7fff;
            const int d = s + (e0 >> 15);
            const int j = i + d;
            const int db = S.dtab[d];
            const int si = S.s8[i], sj1 = S.s8[j + 1], si1 = S.s8[i + 1], sj = S.s8[j];
            const float jB = S.JB[cf_i4(si, sj1, si1, sj)];
            float acc = jB + T[CF_O_HP + (d < 30 ? d : 30)];
            const float e_stack = S.bp[cf_i2(si1, sj)] + T[CF_O_STACK + cf_i4(si, sj1, si1, sj)];
            const float sn01 = T[CF_O_B01 + sj], sn10 = T[CF_O_B01 + si1], sn11 = T[CF_O_I11 + cf_i2(si1, sj)];
            const int pend = min(i + CF_CMAX, j - 2);
            // Rows p = i, i+1 hold the stacking pair, the 0x1 / 1x0 bulges and
            // the 1x1 interior loop (the only candidates with extra terms);
            // they are walked one candidate at a time.
            for (int p = i; p <= min(i + 1, pend); ++p) {
                const int qlo = max(p + 2, j - CF_CMAX + (p - i));
                const int rowJ = S.s8[p + 1] * 25 + S.s8[p];
                uint32_t bits = cf_window(S.pmask + S.rs8[p + 1] * S.W, qlo, j);
                while (bits != 0u) {
                    const int k = 31 - __clz(bits);
                    bits &= ~(1u << k);
                    const int q = qlo + k;
                    const int l1 = p - i, l2 = j - q;
                    const float fc = cf_ld(fc_t, S.dtab[d - l1 - l2 - 2] + p + 1);
                    const float2 bj = S.BJ[(S.qinfo[q] & 0xffff) + rowJ];
                    float e = S.cs[l1 * 31 + l2] + bj.y;
                    e = e + jB;
                    e = e + bj.x;
                    const float sn = (l1 == 0 && l2 == 1) ? sn01
                                   : (l1 == 1 && l2 == 0) ? sn10
                                   : (l1 == 1 && l2 == 1) ? sn11 : 0.0f;
                    e = e + sn;
                    if (l1 + l2 == 0) e = e_stack;
                    acc = cf_lse(acc, e + fc, C);
                }
            }
            // Rows p >= i+2: candidates (p ascending, q descending) gathered 8
            // at a time -- their reads do not depend on acc -- then folded in
            // order; exhausted slots hold NEG, which leaves acc unchanged.
            // The walk holds two rows' partner bits: the current row (cb, row
            // cp, window start cq, junction index crj) and the next non-empty
            // one (nb, np, nq, nrj); pr is the next row to read.  A candidate
            // q = cq + k indexes dtab[q - cp - 2], qinfo[q] and
            // cs[l1*31 + j - q] (l1 = cp - i).
            bool more = (i + 2 <= pend);
            uint32_t cb = 0u, nb = 0u;
            int cp = 0, cq = 0, crj = 0, np = 0, nq = 0, nrj = 0, pr = i + 2;
            auto load_row = [&](int pp, uint32_t& b, int& q, int& rj) {
                q = max(pp + 2, j - CF_CMAX + (pp - i));
                rj = S.s8[pp + 1] * 25 + S.s8[pp];
                b = cf_window(S.pmask + S.rs8[pp + 1] * S.W, q, j);
            };
            // The FC loads of a batch are consumed only in its fold, so all 8
            // are in flight together.  An empty slot reads FC cell (0, 0),
            // which the first step set to NEG, so its term is NEG.
            auto gather8 = [&](float* fv, float* ev) {
                if constexpr (PACK) {
                    // Packed teams: two buffered rows as below, but each
                    // slot's LDS reads follow its walk step (many resident
                    // waves hide their latency, and the batched reads take 90
                    // instead of 71 registers: 5 instead of 7 waves per SIMD)
                    if (cb == 0u) {
                        cb = nb; cp = np; cq = nq; crj = nrj;
                        nb = 0u;
                    }
                    while (cb == 0u && pr <= pend) {
                        load_row(pr, cb, cq, crj);
                        cp = pr++;
                    }
                    while (nb == 0u && pr <= pend) {
                        load_row(pr, nb, nq, nrj);
                        np = pr++;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const bool sw = cb == 0u;
                        cb = sw ? nb : cb;
                        cp = sw ? np : cp;
                        cq = sw ? nq : cq;
                        crj = sw ? nrj : crj;
                        nb = sw ? 0u : nb;
                        int off = 0;
                        float e = 0.0f;
                        if (cb != 0u) {
                            const int k = 31 - __clz(cb);
                            cb &= ~(1u << k);
                            const int q = cq + k;
                            off = S.dtab[q - cp - 2] + cp + 1;
                            const float2 bj = S.BJ[(S.qinfo[q] & 0xffff) + crj];
                            e = S.cs[(cp - i) * 31 + j - q] + bj.y;
                            e = e + jB;
                            e = e + bj.x;
                            e = e + 0.0f;  // the reference adds its (zero) bulge/1x1 extra here too
                        }
                        fv[u] = cf_ld(fc_t, off);
                        ev[u] = e;
                    }
                    more = (cb != 0u) || (nb != 0u) || (pr <= pend);
                    return;
                }
                // Wide and split teams, where one fold's latency is what
                // counts.  Lanes run out of their rows at different slots, so
                // a row change inside the slot loop would cost the whole wave
                // a row's LDS reads at almost every slot.  Instead the two
                // buffered rows are refilled once per batch (skipping empty
                // rows), the 8 slots take their candidates from them with
                // selects only, and then the LDS reads of all 8 slots are
                // issued together.  A slot is empty (ps < 0, entry 0 of each
                // table, NEG term) only when both buffered rows run out within
                // the batch, which the short last rows of a window can do.
                if (cb == 0u) {
                    cb = nb; cp = np; cq = nq; crj = nrj;
                    nb = 0u;
                }
                while (cb == 0u && pr <= pend) {
                    load_row(pr, cb, cq, crj);
                    cp = pr++;
                }
                while (nb == 0u && pr <= pend) {
                    load_row(pr, nb, nq, nrj);
                    np = pr++;
                }
                int qs[8], ps[8], rj[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool sw = cb == 0u;
                    cb = sw ? nb : cb;
                    cp = sw ? np : cp;
                    cq = sw ? nq : cq;
                    crj = sw ? nrj : crj;
                    nb = sw ? 0u : nb;
                    const int k = 31 - __clz(cb | 1u);  // highest set bit (0 if none)
                    ps[u] = cb != 0u ? cp : -1;
                    cb &= ~(1u << k);
                    qs[u] = cq + k;
                    rj[u] = crj;
                }
                int dv[8], qv[8];
                float cv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool h = ps[u] >= 0;
                    dv[u] = S.dtab[h ? qs[u] - ps[u] - 2 : 0];
                    qv[u] = S.qinfo[h ? qs[u] : 0];
                    cv[u] = S.cs[h ? (ps[u] - i) * 31 + j - qs[u] : 0];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float2 bj = S.BJ[(qv[u] & 0xffff) + rj[u]];
                    float e = cv[u] + bj.y;
                    e = e + jB;
                    e = e + bj.x;
                    e = e + 0.0f;  // the reference adds its (zero) bulge/1x1 extra here too
                    fv[u] = cf_ld(fc_t, ps[u] < 0 ? 0 : dv[u] + ps[u] + 1);
                    ev[u] = e;
                }
                more = (cb != 0u) || (nb != 0u) || (pr <= pend);
            };
            auto fold8 = [&](const float* fv, const float* ev) {
#pragma unroll
                for (int u = 0; u < 8; ++u) acc = cf_lse(acc, ev[u] + fv[u], C);
            };
            if constexpr (LAT) {
                // wide teams: the next batch is gathered before the current one
                // is folded (ping-pong buffers), so a lone wave hides its own
                // load latency; batches are folded in the same order
                float fa[8], ea[8], fb[8], eb[8];
                bool ha = more;
                if (ha) gather8(fa, ea);
                while (ha) {
                    const bool hb = more;
                    if (hb) gather8(fb, eb);
                    fold8(fa, ea);
                    if (!hb) break;
                    ha = more;
                    if (ha) gather8(fa, ea);
                    fold8(fb, eb);
                }
            } else {
                while (more) {
                    float fv[8], ev[8];
                    gather8(fv, ev);
                    fold8(fv, ev);
                }
            }
            const float fm2 = fm2_t[db + i];
            if (fm2 > CF_HALF) {
                float jA = T[CF_O_HC + cf_i2(si, sj1)];
                if (i < L) jA = jA + T[CF_O_DL + cf_i3(si, sj1, si1)];
                if (j > 0) jA = jA + T[CF_O_DR + cf_i3(si, sj1, sj)];
                float v = fm2 + jA;
                v = v + mp;
                v = v + mb;
                acc = cf_lse(acc, v, C);
            }
            fc_t[db + i] = acc;
            fcx_t[cf_xbase(j + 1) + i - 1] = acc;
        }
        // C and D read FC of diagonals <= s-1 only (finished a step ago); this
        // sync orders the list counters
        cf_team_sync<PACK>();
        // ---- C, D: FM1 and FM of diagonal s, then of s+1 (each reads its predecessor)
        for (int d = s; d <= s + 1 && d <= L; ++d) {
            const int ncell = L + 1 - d;
            const int db = S.dtab[d];
            for (int i = tid; i < ncell; i += NT) {
                const int j = i + d;
                const bool inner = (i > 0) && (j < L);
                float fm1 = CF_NEG, fmv = CF_NEG;
                if (inner && d >= 2) {
                    const int si1 = S.s8[i + 1], sj = S.s8[j];
                    if (cf_allow(S, i + 1, j)) {
                        const float fcin = fc_t[S.dtab[d - 2] + i + 1];
                        if (fcin > CF_HALF) {
                            float jA = T[CF_O_HC + cf_i2(sj, si1)];
                            if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, si1, S.s8[j + 1])];
                            if (i > 0) jA = jA + T[CF_O_DR + cf_i3(sj, si1, S.s8[i])];
                            float v = fcin + jA;
                            v = v + mp;
                            v = v + S.bp[cf_i2(si1, sj)];
                            fm1 = v;
                        }
                    }
                    const float f1n = fm1_t[S.dtab[d - 1] + i + 1];
                    if (f1n > CF_HALF) fm1 = cf_lse(fm1, f1n + mu, C);
                    const float fm2 = fm2_t[db + i];
                    if (fm2 > CF_HALF) fmv = fm2;
                    const float fml = fm_t[S.dtab[d - 1] + i];
                    if (fml > CF_HALF) fmv = cf_lse(fmv, fml + mu, C);
                    if (fm1 > CF_HALF) fmv = cf_lse(fmv, fm1, C);
                }
                fm1_t[db + i] = fm1;
                fm_t[db + i] = fmv;
            }
            sync_team();
        }
    }
    }  // folds of this team
}

// ================================================================== exterior
// The exterior chains (cpu.py:666-677 and 684-699), one wave per (sequence,
// chain).  The sequence, its mask and the chain's values live in LDS.
//   F5:  F5[j] folds F5[j-1] + eu and then the terms k = 0 .. j-1 in order.
//        The lanes evaluate 64 terms at once (coalesced reads of the FCX
//        column); the fold itself is serial and runs on wave-uniform values:
//        only the terms that exist are visited (folding NEG is the identity,
//        so skipping them leaves every bit of the result unchanged).
//   F5o: step j (descending) adds one source to each F5o[k], k < j; those
//        updates are independent, so lanes take one k each and the chain is
//        L steps long instead of L^2/2.
// Log-sum-exp on wave-uniform values (the F5 chain's serial fold): the
// segment comes from a three-level compare tree on d and the cubic's
// coefficients from one LDS read of the staged table (every lane reads the
// same row), rather than a scalar load that waits on the segment (the
// exterior launch of a 10,000 x 200-nt batch: 3.7 -> 3.3 ms).
__device__ __forceinline__ float cf_lse_u(float a, float b, const float4* __restrict__ C) {
    const float hi = fmaxf(a, b);
    const float lo = fminf(a, b);
    const float d = hi - lo;
    const bool b2 = d >= 3.3792499610f;
    const bool b1 = d >= (b2 ? 5.7890710412f : 1.6320158198f);
    const float t0 = b2 ? (b1 ? 7.8162726752f : 4.4261691294f) : (b1 ? 2.4912588184f : 0.6615367791f);
    const bool b0 = d >= t0;
    const float4 c = C[(b2 ? 4 : 0) | (b1 ? 2 : 0) | (b0 ? 1 : 0)];
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

__host__ __device__ inline int cf_ext_wave_bytes(int Lmax) { return cf_align((Lmax + 2) * 4 + 2 * (Lmax + 3), 16); }

__global__ void __launch_bounds__(256) cf_exterior_kernel(const CfSeq* __restrict__ seqs, const int* __restrict__ order,
                                                          int n, const uint8_t* __restrict__ codes,
                                                          const uint8_t* __restrict__ masks,
                                                          const CfParams<float>* __restrict__ P,
                                                          const float* __restrict__ FCX, float* __restrict__ F5,
                                                          float* __restrict__ F5o, float* __restrict__ logz,
                                                          int mode, int Lmax) {
    __shared__ float4 C[CF_LSE_TABLE_BYTES / 16];  // cf_lse's coefficients and segment bins
    __shared__ float tabs[25 + 25 + 125 + 125];
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    for (int t = tid; t < CF_LSE_BINS; t += blockDim.x) cf_lse_stage(C, t);
    for (int t = tid; t < 300; t += blockDim.x) {
        float v;
        if (t < 25) v = P->t[CF_O_BP + t];
        else if (t < 50) v = P->t[CF_O_HC + t - 25];
        else if (t < 175) v = P->t[CF_O_DL + t - 50];
        else v = P->t[CF_O_DR + t - 175];
        tabs[t] = v;
    }
    __syncthreads();
    const float* bp = tabs;
    const float* hc = tabs + 25;
    const float* dl = tabs + 50;
    const float* dr = tabs + 175;
    const int nw = (int)(blockDim.x / WAVE);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int lane = tid & (WAVE - 1);
    const int nchain = (mode == 3) ? 2 : 1;
    const int task = blockIdx.x * nw + w;
    if (task >= n * nchain) return;
    const int li = task / nchain;
    const bool outside = (mode == 3) ? (task % nchain) == 1 : (mode == 2);
    const int b = order[li];
    const CfSeq sd = seqs[b];
    const int L = sd.L;
    unsigned char* base = smem + w * cf_ext_wave_bytes(Lmax);
    float* __restrict__ f = (float*)base;
    int8_t* __restrict__ s8 = (int8_t*)(base + (Lmax + 2) * 4);
    int8_t* __restrict__ fo8 = s8 + (Lmax + 3);
    for (int q = lane; q < L + 2; q += WAVE) {
        s8[q] = (int8_t)codes[sd.coff + q];
        fo8[q] = (masks != nullptr && masks[sd.coff + q] != 0) ? 1 : 0;
    }
    const float* __restrict__ fcx = FCX + sd.xoff;
    const float eu = P->t[CF_O_EU], ep = P->t[CF_O_EP];
    if (!outside) {
        if (lane == 0) f[0] = 0.0f;
        cf_team_sync<true>();
        for (int j = 1; j <= L; ++j) {
            float acc = f[j - 1] + eu;
            const int sj = s8[j];
            const int sj1 = s8[j + 1];
            if (!fo8[j] && sj != 4) {
                const float* __restrict__ col = fcx + cf_xbase(j);
                for (int k0 = 0; k0 < j; k0 += WAVE) {
                    const int k = k0 + lane;
                    float v = CF_NEG;
                    if (k < j) {
                        const int sk1 = s8[k + 1], sk = s8[k];
                        const float fcv = col[k];
                        const float f5k = f[k];
                        if (cf_canon(sk1, sj) && !fo8[k + 1] && fcv > CF_HALF && f5k > CF_HALF) {
                            float jA = hc[cf_i2(sj, sk1)];
                            if (j < L) jA = jA + dl[cf_i3(sj, sk1, sj1)];
                            if (k > 0) jA = jA + dr[cf_i3(sj, sk1, sk)];
                            v = f5k + fcv;
                            v = v + ep;
                            v = v + bp[cf_i2(sk1, sj)];
                            v = v + jA;
                        }
                    }
                    uint64_t m = __ballot(v > CF_HALF);
                    while (m) {
                        const int t = __ffsll((unsigned long long)m) - 1;
                        m &= m - 1;
                        acc = cf_lse_u(acc, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), t)), C);
                    }
                }
            }
            if (lane == 0) f[j] = acc;
            cf_team_sync<true>();
        }
        if (lane == 0) f[L + 1] = CF_NEG;
        cf_team_sync<true>();
        float* __restrict__ f5 = F5 + sd.coff;
        for (int q = lane; q < L + 2; q += WAVE) f5[q] = f[q];
        if (lane == 0) logz[b] = f[L];
    } else {
        for (int q = lane; q < L + 2; q += WAVE) f[q] = (q == L) ? 0.0f : CF_NEG;
        cf_team_sync<true>();
        for (int j = L; j >= 1; --j) {
            const float f5oj = f[j];
            const float prev = cf_lse_u(f[j - 1], f5oj + eu, C);
            if (lane == 0) f[j - 1] = prev;
            cf_team_sync<true>();
            const int sj = s8[j];
            const int sj1 = s8[j + 1];
            if (!fo8[j] && sj != 4) {
                const float* __restrict__ col = fcx + cf_xbase(j);
                for (int k = lane; k < j; k += WAVE) {
                    const int sk1 = s8[k + 1], sk = s8[k];
                    const float fcv = col[k];
                    if (cf_canon(sk1, sj) && !fo8[k + 1] && fcv > CF_HALF) {
                        float jA = hc[cf_i2(sj, sk1)];
                        if (j < L) jA = jA + dl[cf_i3(sj, sk1, sj1)];
                        if (k > 0) jA = jA + dr[cf_i3(sj, sk1, sk)];
                        float temp = f5oj + ep;
                        temp = temp + bp[cf_i2(sk1, sj)];
                        temp = temp + jA;
                        f[k] = cf_lse(f[k], temp + fcv, C);
                    }
                }
                cf_team_sync<true>();
            }
        }
        float* __restrict__ f5o = F5o + sd.coff;
        for (int q = lane; q < L + 2; q += WAVE) f5o[q] = f[q];
    }
}

// =================================================================== outside
// FP32 outside sweep (cpu.py:680-758) as per-cell gathers that replay the
// reference's scatter order: each target cell folds its sources (all on larger
// diagonals) in the order the reference's (i ascending, j descending) sweep
// delivers them.  The posterior is a separate kernel (cf_post_kernel).
//
// Diagonals are taken two at a time, s and s-1 (descending).  FCo of
// diagonal d reads FCo of diagonals >= d+2 and FM1o of d+2; the split part of
// FMo reads FM2o of diagonals >= d+2; so those are computed for both
// diagonals together (twice the cells per pass).  FMo's last term reads FMo of
// d+1, and FM2o / FM1o read FMo, FCo and FM1o of d+1 on the same row, so they
// follow one diagonal at a time:
//   A: split part of FMo for s and s-1; pairable cells of both into a list
//   B: FCo of the listed cells; FMo of s completed
//   C: FMo of s-1 completed; FM2o, FM1o of s       D: FM2o, FM1o of s-1
// Each cell's terms are folded in the reference's order, so every table is
// bit-identical to a one-diagonal-at-a-time sweep.
template <bool RING, bool PACK, bool LAT>
__global__ void __launch_bounds__(LAT ? 256 : 1024) cf_outside_kernel(
    const CfSeq* __restrict__ seqs, const int* __restrict__ order, int nseq, const uint8_t* __restrict__ codes,
    const uint8_t* __restrict__ masks, const CfParams<float>* __restrict__ P, const float* __restrict__ FC,
    const float* __restrict__ FM, const float* __restrict__ FM1, const float* __restrict__ F5,
    const float* __restrict__ F5o, float* __restrict__ FCo, float* __restrict__ FMo, float* __restrict__ FM1o,
    float* __restrict__ FM2o, int Lmax, int K, unsigned* __restrict__ xsync, int* __restrict__ status,
    int* __restrict__ queue) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int xabort;
    const int G = PACK ? (int)(blockDim.x / WAVE) : 1;
    // a split team keeps no FCo ring: every part reads every diagonal
    const CfTeam tm = cf_team<PACK>(G, (PACK || RING) ? 1 : K);
    const CfLds plan = cf_lds_plan(Lmax, RING, 0, G);
    const CfShared S = cf_carve(smem, plan, tm.team);
    cf_stage_params(S, P);
    if (threadIdx.x == 0) xabort = 0;
    __syncthreads();
    const bool persistent = PACK && queue != nullptr;
    for (int li = persistent ? cf_next_fold(queue) : tm.li; li < nseq;
         li = persistent ? cf_next_fold(queue) : nseq) {
    if (persistent) cf_team_sync<PACK>();  // the team's LDS area is reused fold after fold
    const CfSeq sd = seqs[order[li]];
    const int L = sd.L;
    const int tid = tm.tid, NT = tm.NT;
    const int lane = tid & (WAVE - 1);
    cf_stage_seq<PACK>(S, codes, masks, sd, tm.ltid, tm.lNT);
    CfXBar xb{tm.K > 1 ? xsync + (size_t)li * 16 : nullptr, status, &xabort, 0u, tm.K};
    auto sync_team = [&]() {
        if (tm.K > 1) cf_xbar(xb);
        else cf_team_sync<PACK>();
    };

    const float mb = P->t[CF_O_MB], mu = P->t[CF_O_MU], mp = P->t[CF_O_MP], ep = P->t[CF_O_EP];
    const float* __restrict__ T = P->t;
    float* __restrict__ ring = (float*)S.ring;
    const int RS = Lmax + 1;
    const float* __restrict__ fc_t = FC + sd.toff;
    const float* __restrict__ fm_t = FM + sd.toff;
    const float* __restrict__ fm1_t = FM1 + sd.toff;
    const float* __restrict__ f5 = F5 + sd.coff;
    const float* __restrict__ f5o = F5o + sd.coff;
    float* __restrict__ fco_t = FCo + sd.toff;
    float* __restrict__ fmo_t = FMo + sd.toff;
    float* __restrict__ fm1o_t = FM1o + sd.toff;
    float* __restrict__ fm2o_t = FM2o + sd.toff;
    const float4* __restrict__ C = S.lse;

    // FMo[a][b] += FMo[a][b+1] + mu: the last source of a completed FMo cell.
    // Each thread finishes the cells it took in phase A (the same top-down
    // strips of na columns), so a split team needs no cross-workgroup barrier
    // between phase A and finishing diagonal s in phase B.
    auto fmo_finish = [&](int d, int na) {
        const int ncell = L + 1 - d;
        const int db = S.dtab[d];
        for (int c0 = __builtin_amdgcn_readfirstlane(tid - lane); c0 < na; c0 += NT) {
            const int a = na - WAVE - c0 + lane;
            if (a < 0 || a >= ncell) continue;
            const int b = a + d;
            const bool dom = (a > 0) && (a + 2 <= b) && (b < L);
            if (dom && b + 1 < L) {
                const float x = fmo_t[S.dtab[d + 1] + a];
                if (x > CF_HALF) fmo_t[db + a] = cf_lse(fmo_t[db + a], x + mu, C);
            }
        }
    };
    // FM2o and FM1o of every multiloop cell of diagonal d
    auto fm12o = [&](int d) {
        const int ncell = L + 1 - d;
        const int db = S.dtab[d];
        for (int i0 = __builtin_amdgcn_readfirstlane(tid - lane); i0 < ncell; i0 += NT) {
            const int a = i0 + lane;
            const int b = a + d;
            const bool live = a < ncell;
            const bool dom = live && (a > 0) && (a + 2 <= b) && (b < L);
            float fm2o = CF_NEG, f1 = CF_NEG, fmo = CF_NEG;
            if (dom) {
                fmo = fmo_t[db + a];
                fm2o = fmo;
                if (cf_allow(S, a, b + 1)) {
                    const float fco = RING ? ring[(d % CF_RING) * RS + a] : fco_t[db + a];
                    if (fco > CF_HALF) {
                        const int sa = S.s8[a], sb1 = S.s8[b + 1];
                        float jA = T[CF_O_HC + cf_i2(sa, sb1)];
                        if (a < L) jA = jA + T[CF_O_DL + cf_i3(sa, sb1, S.s8[a + 1])];
                        if (b > 0) jA = jA + T[CF_O_DR + cf_i3(sa, sb1, S.s8[b])];
                        float v = fco + jA;
                        v = v + mp;
                        v = v + mb;
                        fm2o = cf_lse(fm2o, v, C);
                    }
                }
                // (i) FM1o[a-1][b] + mu
                if (a - 1 > 0 && a + 1 <= b) {
                    const float x = fm1o_t[S.dtab[d + 1] + a - 1];
                    if (x > CF_HALF) f1 = cf_lse(f1, x + mu, C);
                }
            }
            // (ii) FM2o(a, b+v) + FM[b][b+v] for v = L-1-b .. 2 descending (lanes step v together)
            const int vmax = L - 1 - (i0 + d);
            int v = vmax;
            int ox = (int)cf_dbase(L, d + v), oy = (int)cf_dbase(L, v);  // diagonals d + v (FM2o) and v (FM)
            {
                int vl = v;  // split offset of the next load
                auto load8 = [&](float* x, float* y, bool* g) {
#pragma unroll
                    for (int w = 0; w < 8; ++w) {
                        const int vv = vl - w;
                        g[w] = dom & (vv <= L - 1 - b);
                        x[w] = cf_ld(fm2o_t, ox + (g[w] ? a : 0));
                        y[w] = cf_ld(fm_t, oy + (g[w] ? b : 0));
                        ox -= L + 2 - (d + vv);
                        oy -= L + 2 - vv;
                    }
                    vl -= 8;
                };
                auto fold8 = [&](const float* x, const float* y, const bool* g) {
#pragma unroll
                    for (int w = 0; w < 8; ++w) {
                        f1 = cf_lse(f1, x[w] + y[w], C);
                    }
                };
                const int nfull = v >= 1 ? (v - 1) / 8 : 0;  // batches while v - 7 >= 2
                if constexpr (LAT) {
                    float xa[8], ya[8], xb[8], yb[8];
                    bool ga[8], gb[8];
                    if (nfull > 0) load8(xa, ya, ga);
                    for (int k = 0; k < nfull; k += 2) {
                        if (k + 1 < nfull) load8(xb, yb, gb);
                        fold8(xa, ya, ga);
                        if (k + 1 >= nfull) break;
                        if (k + 2 < nfull) load8(xa, ya, ga);
                        fold8(xb, yb, gb);
                    }
                } else {
                    for (int k = 0; k < nfull; ++k) {
                        float x[8], y[8];
                        bool g[8];
                        load8(x, y, g);
                        fold8(x, y, g);
                    }
                }
                v -= 8 * nfull;
            }
            for (; v >= 2; --v) {
                const bool g = dom & (v <= L - 1 - b);
                const float x = cf_ld(fm2o_t, ox + (g ? a : 0));
                const float y = cf_ld(fm_t, oy + (g ? b : 0));
                ox -= L + 2 - (d + v);
                oy -= L + 2 - v;
                f1 = cf_lse(f1, x + y, C);
            }
            if (dom && fmo > CF_HALF) f1 = cf_lse(f1, fmo, C);
            if (live) {
                fm2o_t[db + a] = fm2o;
                fm1o_t[db + a] = f1;
            }
        }
    };

    for (int s = L; s >= 0; s -= 2) {
        const int step = (L - s) >> 1;
        const int n0 = L + 1 - s;
        const int n1 = (s >= 1) ? L + 2 - s : 0;  // cells of diagonal s-1
        // ---- A: split part of FMo for diagonals s and s-1; collect pairable cells.
        // A lane takes column a of both diagonals: cells (a, a+s) and
        // (a, a+s-1) fold FM2o of their own rows against the same FM1[i'][a],
        // so that column is read once for two cells.
        // A cell's split chain is as long as its column a, and a wave loops as
        // long as its top lane: the waves take 64-column strips from the top
        // down, so the partial strip (na not a multiple of 64) is the one with
        // the shortest chains.
        const int na = max(n0, n1);
        for (int c0 = __builtin_amdgcn_readfirstlane(tid - lane); c0 < na; c0 += NT) {
            const int a = na - WAVE - c0 + lane;
            const bool live0 = a >= 0 && a < n0, live1 = a >= 0 && a < n1;
            const bool dom0 = live0 && (a > 0) && (s >= 2) && (a + s < L);
            const bool dom1 = live1 && (a > 0) && (s >= 3) && (a + s - 1 < L);
            float f0 = CF_NEG, f1 = CF_NEG;
            // FM2o(i', b) + FM1[i'][a] for i' = 1 .. a-2 ascending, i.e. u = a - i'
            // descending from the pass's largest a.  Wave-uniform offsets of the
            // diagonals s + u, s - 1 + u (FM2o) and u (FM1) step down with u;
            // lanes outside a domain read cell 0 of the diagonal and fold NEG,
            // which leaves the sum unchanged.
            const int amax = na - 1 - c0;
            int u = amax - 1;
            int ox0 = (int)cf_dbase(L, s + u), ox1 = (int)cf_dbase(L, s - 1 + u), oy = (int)cf_dbase(L, u);
            {
                int ul = u;  // split offset of the next load
                auto load8 = [&](float* x0, float* x1, float* y, bool* g0, bool* g1) {
#pragma unroll
                    for (int w = 0; w < 8; ++w) {
                        const int uu = ul - w;
                        g0[w] = dom0 & (uu <= a - 1);
                        g1[w] = dom1 & (uu <= a - 1);
                        x0[w] = cf_ld(fm2o_t, ox0 + (g0[w] ? a - uu : 0));
                        x1[w] = cf_ld(fm2o_t, ox1 + (g1[w] ? a - uu : 0));
                        y[w] = cf_ld(fm1_t, oy + ((g0[w] | g1[w]) ? a - uu : 0));
                        ox0 -= L + 2 - (s + uu);
                        ox1 -= L + 2 - (s - 1 + uu);
                        oy -= L + 2 - uu;
                    }
                    ul -= 8;
                };
                auto fold8 = [&](const float* x0, const float* x1, const float* y, const bool* g0, const bool* g1) {
#pragma unroll
                    for (int w = 0; w < 8; ++w) {
                        f0 = cf_lse(f0, x0[w] + y[w], C);
                        f1 = cf_lse(f1, x1[w] + y[w], C);
                    }
                };
                const int nfull = u >= 1 ? (u - 1) / 8 : 0;  // batches while u - 7 >= 2
                if constexpr (LAT) {
                    // wide teams: next batch loaded before the current is folded
                    float xa0[8], xa1[8], ya[8], xb0[8], xb1[8], yb[8];
                    bool ga0[8], ga1[8], gb0[8], gb1[8];
                    if (nfull > 0) load8(xa0, xa1, ya, ga0, ga1);
                    for (int k = 0; k < nfull; k += 2) {
                        if (k + 1 < nfull) load8(xb0, xb1, yb, gb0, gb1);
                        fold8(xa0, xa1, ya, ga0, ga1);
                        if (k + 1 >= nfull) break;
                        if (k + 2 < nfull) load8(xa0, xa1, ya, ga0, ga1);
                        fold8(xb0, xb1, yb, gb0, gb1);
                    }
                } else {
                    for (int k = 0; k < nfull; ++k) {
                        float x0[8], x1[8], y[8];
                        bool g0[8], g1[8];
                        load8(x0, x1, y, g0, g1);
                        fold8(x0, x1, y, g0, g1);
                    }
                }
                u -= 8 * nfull;
            }
            for (; u >= 2; --u) {
                const bool g0 = dom0 & (u <= a - 1);
                const bool g1 = dom1 & (u <= a - 1);
                const float x0 = cf_ld(fm2o_t, ox0 + (g0 ? a - u : 0));
                const float x1 = cf_ld(fm2o_t, ox1 + (g1 ? a - u : 0));
                const float y = cf_ld(fm1_t, oy + ((g0 | g1) ? a - u : 0));
                ox0 -= L + 2 - (s + u);
                ox1 -= L + 2 - (s - 1 + u);
                oy -= L + 2 - u;
                f0 = cf_lse(f0, x0 + y, C);
                f1 = cf_lse(f1, x1 + y, C);
            }
#pragma unroll
            for (int hi = 0; hi < 2; ++hi) {
                if (!(hi ? live1 : live0)) continue;
                const int d = s - hi;
                const int b = a + d;
                const int db = S.dtab[d];
                fmo_t[db + a] = hi ? f1 : f0;
                if ((a > 0) && (b < L) && cf_allow(S, a, b + 1)) {
                    const int slotn = atomicAdd(&S.cnt[step & 1], 1);
                    S.alist[slotn] = (uint16_t)(a | (hi << 15));
                } else {
                    fco_t[db + a] = CF_NEG;
                    if (RING) ring[(d % CF_RING) * RS + a] = CF_NEG;
                }
            }
        }
        cf_team_sync<PACK>();
        // ---- B: FCo of the pairable cells of both diagonals; FMo of s completed
        // (a split team's workgroups each take the cells they listed in A)
        const int nal = S.cnt[step & 1];
        if (tm.ltid == 0) S.cnt[(step + 1) & 1] = 0;
        for (int idx = tm.ltid; idx < nal; idx += tm.lNT) {
            const int e0 = S.alist[idx];
            const int a = e0 & 0x7fff;
            const int d = s - (e0 >> 15);
            const int b = a + d;
            const int db = S.dtab[d];
            const int sa = S.s8[a], sb1 = S.s8[b + 1], sam = S.s8[a - 1], sb2 = S.s8[b + 2];
            const float bpab = S.bp[cf_i2(sa, sb1)];
            // jA of branch (a, b+1) seen from the enclosing loop (exterior or multiloop)
            float jAx = T[CF_O_HC + cf_i2(sb1, sa)];
            if (b + 1 < L) jAx = jAx + T[CF_O_DL + cf_i3(sb1, sa, sb2)];
            if (a - 1 > 0) jAx = jAx + T[CF_O_DR + cf_i3(sb1, sa, sam)];
            const float jB2 = S.JB[cf_i4(sb1, sa, sb2, sam)];
            const float e_stack = bpab + T[CF_O_STACK + cf_i4(sam, sb2, sa, sb1)];
            // bulge 0x1 / 1x0 and interior 1x1 extras of the three nearest outer pairs
            const float sn01 = T[CF_O_B01 + sb2], sn10 = T[CF_O_B01 + sam], sn11 = T[CF_O_I11 + cf_i2(sam, sb2)];
            const bool fm1dom = (a - 1 > 0) && (a + 1 <= b + 1) && (b + 1 < L);
            const float fm1o_src = fm1dom ? fm1o_t[S.dtab[d + 2] + a - 1] : CF_NEG;
            const int lmax = min(CF_CMAX, a - 2);
            const int qbase = b + 2;
            float acc = CF_NEG;
            const float f5a = f5[a - 1];
            if (f5a > CF_HALF) {
                float temp = f5o[b + 1] + ep;
                temp = temp + bpab;
                temp = temp + jAx;
                acc = temp + f5a;
            }
            // Sources in the reference's order: outer rows i' = a-1-l1 ascending
            // (l1 descending), partners descending; in row l1 = 0 the FM1o source
            // of iteration (a-1, b+1) comes after every k >= 1 and before k = 0.
            // Sources stream 8 at a time -- their reads do not depend on acc;
            // exhausted slots fold NEG, which leaves acc unchanged.  Rows l1 >= 2
            // add the reference's zero extra; rows 1 and 0 (k >= 1) carry the
            // bulge 0x1 / 1x0 and 1x1 extras and are walked by their own
            // instance; the stacking source k = 0 of row 0 is read up front.
            // Without the ring the FCo loads of a batch are consumed only in
            // its fold, so all 8 are in flight together; an empty slot reads
            // cell (0, L), set to NEG by the first step, and its term is NEG.
            const int off_neg = S.dtab[L];
            const uint32_t bits_r0 = (lmax >= 0) ? cf_window(S.pmask + S.rs8[a - 1] * S.W, qbase, min(L, qbase + CF_CMAX))
                                                 : 0u;
            float fco_stk = CF_NEG;
            if (bits_r0 & 1u) {
                if constexpr (RING) fco_stk = ring[((d + 2) % CF_RING) * RS + a - 1];
                else fco_stk = cf_ld(fco_t, S.dtab[d + 2] + a - 1);
            }
            const int16_t* qor = S.qo + (qbase - 1);
            auto walk = [&](auto extras, int l1, int l1_last) {
                if (l1 < l1_last) return;
                bool more = true;
                // The walk holds two rows' partner bits: the current row (cb,
                // row cl, junction index crj) and the next non-empty one (nb,
                // nl, nrj); lr is the next row to read (rows run l1
                // descending).  A candidate l2 of row l1 indexes
                // dtab[d + l1 + 2 + l2] and cs[l1*31 + l2].
                uint32_t cb = 0u, nb = 0u;
                int cl = 0, crj = 0, nl = 0, nrj = 0, lr = l1;
                auto load_row = [&](int ll, uint32_t& b, int& rj) {
                    const int ipn = a - 1 - ll;
                    b = (ll == 0) ? (bits_r0 & ~1u)
                                  : cf_window(S.pmask + S.rs8[ipn] * S.W, qbase, min(L, qbase + CF_CMAX - ll));
                    rj = S.s8[ipn] * 125 + S.s8[ipn + 1] * 5;
                };
                auto gather8 = [&](float* fv, float* ev) {
                    if constexpr (!PACK) {
                        // Wide and split teams, where one fold's latency is
                        // what counts (the packed teams keep the one-slot-at-
                        // a-time form below, whose LDS latency other resident
                        // waves hide, in fewer registers).  The two buffered
                        // rows are refilled once per batch (skipping empty
                        // rows), the 8 slots take their candidates from them
                        // with selects only -- a row change inside the slot
                        // loop would cost the wave a row's LDS reads at almost
                        // every slot -- and then the slots' LDS reads are
                        // issued together.  A slot is empty (l1s < 0: entry 0
                        // of each table, NEG term) only when both buffered
                        // rows run out within the batch.
                        if (cb == 0u) {
                            cb = nb; cl = nl; crj = nrj;
                            nb = 0u;
                        }
                        while (cb == 0u && lr >= l1_last) {
                            load_row(lr, cb, crj);
                            cl = lr--;
                        }
                        while (nb == 0u && lr >= l1_last) {
                            load_row(lr, nb, nrj);
                            nl = lr--;
                        }
                        int l1s[8], l2s[8], rjs[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const bool sw = cb == 0u;
                            cb = sw ? nb : cb;
                            cl = sw ? nl : cl;
                            crj = sw ? nrj : crj;
                            nb = sw ? 0u : nb;
                            const int k = 31 - __clz(cb | 1u);  // highest set bit (0 if none)
                            l1s[u] = cb != 0u ? cl : -1;
                            cb &= ~(1u << k);
                            l2s[u] = k;
                            rjs[u] = crj;
                        }
                        int dv[8], qv[8];
                        float cv[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const bool h = l1s[u] >= 0;
                            dv[u] = S.dtab[h ? d + l1s[u] + 2 + l2s[u] : L];
                            cv[u] = S.cs[h ? l1s[u] * 31 + l2s[u] : 0];
                            qv[u] = qor[h ? l2s[u] : 0];
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const bool h = l1s[u] >= 0;
                            const int l1u = l1s[u], l2u = l2s[u];
                            float e = cv[u] + bpab;
                            e = e + S.JB[rjs[u] + qv[u]];
                            e = e + jB2;
                            if constexpr (decltype(extras)::value) {
                                const float sn = (l1u == 0) ? (l2u == 1 ? sn01 : 0.0f)
                                               : (l2u == 0 ? sn10 : (l2u == 1 ? sn11 : 0.0f));
                                e = e + sn;
                            } else {
                                e = e + 0.0f;  // the reference adds its (zero) bulge/1x1 extra here too
                            }
                            if constexpr (RING) {
                                const float fco = h ? ring[((d + l1u + l2u + 2) % CF_RING) * RS + a - 1 - l1u] : CF_NEG;
                                fv[u] = fco + e;
                            } else {
                                fv[u] = cf_ld(fco_t, h ? dv[u] + a - 1 - l1u : off_neg);
                                ev[u] = e;
                            }
                        }
                        more = (cb != 0u) || (nb != 0u) || (lr >= l1_last);
                        return;
                    }
                    // Packed teams: the same two buffered rows, but each
                    // slot's LDS reads follow its walk step (many resident
                    // waves hide their latency, in fewer registers)
                    if (cb == 0u) {
                        cb = nb; cl = nl; crj = nrj;
                        nb = 0u;
                    }
                    while (cb == 0u && lr >= l1_last) {
                        load_row(lr, cb, crj);
                        cl = lr--;
                    }
                    while (nb == 0u && lr >= l1_last) {
                        load_row(lr, nb, nrj);
                        nl = lr--;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const bool sw = cb == 0u;
                        cb = sw ? nb : cb;
                        cl = sw ? nl : cl;
                        crj = sw ? nrj : crj;
                        nb = sw ? 0u : nb;
                        float e = 0.0f;
                        int off = off_neg;
                        int l2 = 0;
                        const bool has = cb != 0u;
                        if (has) {
                            l2 = 31 - __clz(cb);
                            cb &= ~(1u << l2);
                            off = S.dtab[d + cl + 2 + l2] + a - 1 - cl;
                            e = S.cs[cl * 31 + l2] + bpab;
                            e = e + S.JB[crj + qor[l2]];
                            e = e + jB2;
                            if constexpr (decltype(extras)::value) {
                                const float sn = (cl == 0) ? (l2 == 1 ? sn01 : 0.0f)
                                               : (l2 == 0 ? sn10 : (l2 == 1 ? sn11 : 0.0f));
                                e = e + sn;
                            } else {
                                e = e + 0.0f;  // the reference adds its (zero) bulge/1x1 extra here too
                            }
                        }
                        if constexpr (RING) {
                            const float fco = has ? ring[((d + cl + l2 + 2) % CF_RING) * RS + a - 1 - cl] : CF_NEG;
                            fv[u] = fco + e;
                        } else {
                            fv[u] = cf_ld(fco_t, off);
                            ev[u] = e;
                        }
                    }
                    more = (cb != 0u) || (nb != 0u) || (lr >= l1_last);
                };
                auto fold8 = [&](const float* fv, const float* ev) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if constexpr (RING)
                            acc = cf_lse(acc, fv[u], C);
                        else
                            acc = cf_lse(acc, fv[u] + ev[u], C);
                    }
                };
                if constexpr (LAT) {
                    // wide teams without the ring: the next batch is gathered
                    // before the current one is folded (ping-pong buffers)
                    float fa[8], ea[8], fb[8], eb[8];
                    gather8(fa, ea);
                    bool ha = true;
                    while (ha) {
                        const bool hb = more;
                        if (hb) gather8(fb, eb);
                        fold8(fa, ea);
                        if (!hb) break;
                        ha = more;
                        if (ha) gather8(fa, ea);
                        fold8(fb, eb);
                    }
                } else {
                    while (more) {
                        float fv[8], ev[8];
                        gather8(fv, ev);
                        fold8(fv, ev);
                    }
                }
            };
            if (lmax >= 2) walk(std::false_type{}, lmax, 2);
            if (lmax >= 0) {
                walk(std::true_type{}, min(lmax, 1), 0);
                if (fm1o_src > CF_HALF) {
                    float v = fm1o_src + jAx;
                    v = v + mp;
                    v = v + bpab;
                    acc = cf_lse(acc, v, C);
                }
                if (bits_r0 & 1u) acc = cf_lse(acc, (fco_stk > CF_HALF) ? fco_stk + e_stack : CF_NEG, C);
            }
            fco_t[db + a] = acc;
            if (RING) ring[(d % CF_RING) * RS + a] = acc;
        }
        fmo_finish(s, na);
        sync_team();
        // ---- C: FMo of s-1 completed; FM2o, FM1o of s
        if (s >= 1) fmo_finish(s - 1, na);
        fm12o(s);
        sync_team();
        // ---- D: FM2o, FM1o of s-1
        if (s >= 1) {
            fm12o(s - 1);
            sync_team();
        }
    }
    }  // folds of this team
}

// ================================================================ posterior
// Base-pair posteriors (cpu.py:765-819) from the finished outside tables.
// POST(a, b+1) sums exp(outside + inside - logZ) over the outer pairs that
// close (a, b+1) as a single-branch loop, the multiloop (FM1) source and the
// exterior source, in the reference's order.  Unlike FCo, the cells are
// independent once the outside sweep is done.  A workgroup takes one
// sequence (staged once in LDS) and its waves take every G-th diagonal; each
// lane evaluates one cell, walking its single-branch window with the reads of
// 8 sources in flight at a time.
//
// A wave runs until its busiest lane is done, and window sizes differ a lot
// (boundary cells, base composition, masks), so each wave pools up to
// CF_POST_POOL pairable cells, counts every pooled cell's candidates
// (popcounts of its window rows) and takes its batches of 64 in order of that
// count: lanes of one batch then finish together.
// POSTF is indexed by FC cell: POSTF[tri(a, b)] = P(a pairs b+1).
#define CF_POST_POOL 256
// per-wave work area: the pool (cells as a | s << 13 | count class << 26)
// with room for one more diagonal chunk, and the pool sorted by class
#define CF_POST_WORK ((CF_POST_POOL + 64) * 4 + CF_POST_POOL * 4)

__global__ void __launch_bounds__(1024) cf_post_kernel(
    const CfSeq* __restrict__ seqs, const int* __restrict__ order, int nseq, const uint8_t* __restrict__ codes,
    const uint8_t* __restrict__ masks, const CfParams<float>* __restrict__ P, const float* __restrict__ FC,
    const float* __restrict__ F5, const float* __restrict__ F5o, const float* __restrict__ FCo,
    const float* __restrict__ FM1o, float* __restrict__ POSTF, int Lmax, int K) {
    extern __shared__ __align__(16) unsigned char smem[];
    // K > 1: K workgroups share a sequence (launch indices as in cf_team) and
    // split its diagonals; the cells are independent, so they never wait
    int li = (int)blockIdx.x, part = 0;
    if (K > 1) {
        const int g = (int)blockIdx.x, r = g >> 3;
        part = r % K;
        li = (r / K) * 8 + (g & 7);
    }
    if (li >= nseq) return;
    const int G = (int)(blockDim.x / WAVE);
    const CfLds plan = cf_lds_plan(Lmax, false, 0, 1);
    const CfShared S = cf_carve(smem, plan, 0);
    cf_stage_params(S, P);
    __syncthreads();
    const CfSeq sd = seqs[order[li]];
    const int L = sd.L;
    cf_stage_seq<false>(S, codes, masks, sd, threadIdx.x, blockDim.x);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int lane = threadIdx.x & (WAVE - 1);
    unsigned char* work = smem + plan.total + wv * CF_POST_WORK;
    int* __restrict__ pool = (int*)work;
    int* __restrict__ sorted = pool + CF_POST_POOL + 64;

    const float mp = P->t[CF_O_MP], ep = P->t[CF_O_EP];
    const float* __restrict__ T = P->t;
    const float* __restrict__ fc_t = FC + sd.toff;
    const float* __restrict__ fco_t = FCo + sd.toff;
    const float* __restrict__ fm1o_t = FM1o + sd.toff;
    const float* __restrict__ f5 = F5 + sd.coff;
    const float* __restrict__ f5o = F5o + sd.coff;
    float* __restrict__ post_t = POSTF + sd.toff;
    const float4* __restrict__ E = S.ex;
    const float Z = f5[L];
    const int off_neg = S.dtab[L];  // FCo(0, L): no pair (0, L+1), always NEG

    auto window = [&](int a, int qbase, int l1) -> uint32_t {
        return cf_window(S.pmask + S.rs8[a - 1 - l1] * S.W, qbase, min(L, qbase + CF_CMAX - l1));
    };

    auto cell = [&](int a, int s) {
        const int b = a + s;
        const int lmax = min(CF_CMAX, a - 2);
        const int qbase = b + 2;
        // the cell's own reads are issued together
        const uint32_t bits0 = (lmax >= 0) ? window(a, qbase, 0) : 0u;
        const bool fm1dom = (a - 1 > 0) && (a + 1 <= b + 1) && (b + 1 < L);
        const float fcab = fc_t[S.dtab[s] + a];
        const float fm1o_src = fm1dom ? fm1o_t[S.dtab[s + 2] + a - 1] : CF_NEG;
        const float fco_stk = (bits0 & 1u) ? cf_ld(fco_t, S.dtab[s + 2] + a - 1) : CF_NEG;
        const float outx = f5o[b + 1] - Z;
        const float f5a = f5[a - 1];
        float post = 0.0f;
        if (fcab > CF_HALF) {
            const int sa = S.s8[a], sb1 = S.s8[b + 1], sam = S.s8[a - 1], sb2 = S.s8[b + 2];
            const float bpab = S.bp[cf_i2(sa, sb1)];
            float jAx = T[CF_O_HC + cf_i2(sb1, sa)];
            if (b + 1 < L) jAx = jAx + T[CF_O_DL + cf_i3(sb1, sa, sb2)];
            if (a - 1 > 0) jAx = jAx + T[CF_O_DR + cf_i3(sb1, sa, sam)];
            const float jB2 = S.JB[cf_i4(sb1, sa, sb2, sam)];
            // rows l1 = 0 .. lmax, partners ascending (the reverse of the FCo
            // order); in row 0 the FM1o source follows the stacking term k = 0
            if (lmax >= 0) {
                if (bits0 & 1u) {
                    const float stk = T[CF_O_STACK + cf_i4(sam, sb2, sa, sb1)];
                    const float outside = fco_stk - Z;
                    float e = outside + bpab;
                    e = e + stk;
                    e = e + fcab;
                    post = post + ((outside > CF_HALF) ? cf_fexp(e, E) : 0.0f);
                }
                if (fm1o_src > CF_HALF) {
                    float v = fm1o_src + fcab;
                    v = v + jAx;
                    v = v + mp;
                    v = v + bpab;
                    v = v - Z;
                    post = post + cf_fexp(v, E);
                }
                // Sources gathered 8 at a time (their reads do not depend on
                // the running sum), then added in order; an empty slot reads
                // the NEG cell and adds nothing.  Rows 0 and 1 carry the
                // 0x1 / 1x0 bulge and 1x1 extras; rows >= 2 add nothing more.
                const float sn01 = T[CF_O_B01 + sb2], sn10 = T[CF_O_B01 + sam], sn11 = T[CF_O_I11 + cf_i2(sam, sb2)];
                // A slot holds its FCo index and, packed, its JB index (10
                // bits) and cs index l1*31 + kk (10 bits); the extras are at
                // cs indices 1 (0x1), 31 (1x0) and 32 (1x1).
                // The walk holds two rows' partner bits, the current row (cb,
                // row cl, junction index crj) and the next non-empty one (nb,
                // nl, nrj), refilled once per batch; the 8 slots take their
                // sources from them with selects only.  (A row change inside
                // the slot loop would cost the wave a row's LDS reads at
                // almost every slot, since lanes run out of their rows at
                // different slots.)  A slot is empty only when both buffered
                // rows run out within the batch.
                auto rows = [&](auto extras, int r0, int r1, uint32_t bits) {
                    auto load_row = [&](int ll, uint32_t& b, int& rj) {
                        b = window(a, qbase, ll);
                        rj = S.s8[a - 1 - ll] * 125 + S.s8[a - ll] * 5;
                    };
                    uint32_t cb = bits, nb = 0u;
                    int cl = r0, crj = S.s8[a - 1 - r0] * 125 + S.s8[a - r0] * 5, nl = 0, nrj = 0, lr = r0 + 1;
                    bool more = true;
                    while (more) {
                        if (cb == 0u) {
                            cb = nb; cl = nl; crj = nrj;
                            nb = 0u;
                        }
                        while (cb == 0u && lr <= r1) {
                            load_row(lr, cb, crj);
                            cl = lr++;
                        }
                        while (nb == 0u && lr <= r1) {
                            load_row(lr, nb, nrj);
                            nl = lr++;
                        }
                        int off[8], tag[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const bool sw = cb == 0u;
                            cb = sw ? nb : cb;
                            cl = sw ? nl : cl;
                            crj = sw ? nrj : crj;
                            nb = sw ? 0u : nb;
                            off[u] = off_neg;
                            tag[u] = 0;
                            if (cb != 0u) {
                                const int kk = __ffs(cb) - 1;
                                cb &= cb - 1u;
                                off[u] = S.dtab[s + cl + kk + 2] + a - 1 - cl;
                                tag[u] = (crj + S.qo[qbase + kk - 1]) | ((cl * 31 + kk) << 10);
                            }
                        }
                        more = (cb != 0u) || (nb != 0u) || (lr <= r1);
                        float o[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) o[u] = cf_ld(fco_t, off[u]);
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int csi = tag[u] >> 10;
                            const float outside = o[u] - Z;
                            float e = outside + S.JB[tag[u] & 1023];
                            e = e + S.cs[csi];
                            e = e + fcab;
                            e = e + bpab;
                            e = e + jB2;
                            if constexpr (decltype(extras)::value)
                                e = e + (csi == 1 ? sn01 : csi == 31 ? sn10 : csi == 32 ? sn11 : 0.0f);
                            post = post + ((outside > CF_HALF) ? cf_fexp(e, E) : 0.0f);
                        }
                    }
                };
                rows(std::true_type{}, 0, min(lmax, 1), bits0 & ~1u);
                if (lmax >= 2) rows(std::false_type{}, 2, lmax, window(a, qbase, 2));
            }
            if (outx > CF_HALF && f5a > CF_HALF) {
                float v = outx + f5a;
                v = v + fcab;
                v = v + ep;
                v = v + bpab;
                v = v + jAx;
                post = post + cf_fexp(v, E);
            }
            post = post < 0.0f ? 0.0f : (post > 1.0f ? 1.0f : post);
        }
        post_t[S.dtab[s] + a] = post;
    };

    // candidates of a pooled cell, as a count class (16 per class)
    auto count_class = [&](int a, int s) -> int {
        const int lmax = min(CF_CMAX, a - 2);
        const int qbase = a + s + 2;
        int n = 0;
        for (int l1 = 0; l1 <= lmax; ++l1) n += __popc(window(a, qbase, l1));
        return min(n >> 4, 31);
    };

    // Sort the first n pooled cells by class (counting sort over the wave:
    // per class, ballots give each member its slot) and evaluate them 64 at a
    // time; lane l takes the l-th cell of each batch.
    auto drain = [&](int n) {
        constexpr int R = CF_POST_POOL / WAVE;
        int ent[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int t = r * WAVE + lane;
            ent[r] = -1;
            if (t < n) {
                const int e = pool[t];
                ent[r] = e | (count_class(e & 0x1fff, (e >> 13) & 0x1fff) << 26);
            }
        }
        int base = 0;
        for (int c = 0; c < 32; ++c) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bool hit = ent[r] >= 0 && (ent[r] >> 26) == c;
                const uint64_t m = __ballot(hit);
                if (hit) sorted[base + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = ent[r];
                base += __popcll(m);
            }
        }
        cf_team_sync<true>();
        for (int t0 = 0; t0 < n; t0 += WAVE) {
            if (t0 + lane < n) {
                const int e = sorted[t0 + lane];
                cell(e & 0x1fff, (e >> 13) & 0x1fff);
            }
        }
        cf_team_sync<true>();
    };

    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int pn = 0;  // pooled cells (wave-uniform)
    for (int s = part * G + wv; s <= L; s += K * G) {
        const int ncell = L + 1 - s;
        for (int i0 = 0; i0 < ncell; i0 += WAVE) {
            const int a = i0 + lane;
            const int b = a + s;
            const bool live = a < ncell;
            const bool pr = live && (a > 0) && (b < L) && cf_allow(S, a, b + 1);
            if (live && !pr) post_t[S.dtab[s] + a] = 0.0f;
            const uint64_t m = __ballot(pr);
            if (pr) pool[pn + __popcll(m & below)] = a | (s << 13);
            pn += __popcll(m);
            if (pn >= CF_POST_POOL) {
                cf_team_sync<true>();
                drain(CF_POST_POOL);
                const int rest = pn - CF_POST_POOL;
                const int moved = (lane < rest) ? pool[CF_POST_POOL + lane] : 0;
                cf_team_sync<true>();
                if (lane < rest) pool[lane] = moved;
                pn = rest;
            }
        }
    }
    cf_team_sync<true>();
    if (pn > 0) drain(pn);
}

// ======================================================================= MEA
// Maximum-expected-accuracy decoding (cpu.py:823-895).
//
// The score table S (cpu.py:843-871) is a max-plus recurrence: each candidate
// is one FP32 addition of two table values, and a maximum does not depend on
// the order its operands are compared in.  So S is computed here in an order
// chosen for data reuse, bit-identical to the reference's.  The reference's
// strict '>' scan only decides the traceback: it keeps the first candidate,
// in its order (i+1 unpaired, j unpaired, pair (i+1, j), bifurcations k
// ascending), that reaches the cell's maximum; the traceback finds that
// candidate again by replaying the order against S.
//
// Layout: S of one fold is stored in 16 x 16 tiles, tile (I, J) holding rows
// 16I.. and columns 16J.., the tiles of block diagonal D = J - I stored
// together, D ascending (cf_mea_tile).  Inside a tile the cells are stored by
// in-tile diagonals d = c - r ascending (cf_mea_tix), so the cells a step
// touches -- one in-tile diagonal per operand -- are contiguous across lanes.  A cell's bifurcation operands
// S[i][k], S[k][j] lie in tiles (I, K) and (K, J):
//   far  (I < K < J): both tiles lie on block diagonals below D.  Phase 1
//        stages the pair in LDS and folds a max-plus tile product into the
//        cell (each loaded value serves 16 cells);
//   near (K = I or K = J): one operand lies in the cell's own tile.  Phase 2
//        walks the tile's cells by w = c - r (in-tile column minus row):
//        the near operands of a cell lie on smaller w, and so do the
//        unpaired / pair candidates' cells.  Every cell of one step has
//        15 + w near terms (w - 1 in a diagonal tile), so lanes stay
//        converged.
// Teams: PACK -- one wave per (fold, gamma), several per workgroup, taking
// items from a queue; otherwise the workgroup is one team (phase-1 tiles
// spread over its waves, phase-2 cells over its threads).
#define CF_MT 16
__host__ __device__ inline int cf_mea_nb(int L) { return (L + CF_MT) / CF_MT; }  // ceil((L + 1) / 16)
__host__ __device__ inline int64_t cf_mea_cells(int L) {
    const int64_t nb = cf_mea_nb(L);
    return nb * (nb + 1) / 2 * (CF_MT * CF_MT);
}
// tile (I, I + D) of a fold with nb tile rows
__device__ __forceinline__ int cf_mea_tile(int nb, int I, int D) { return D * nb - ((D * (D - 1)) >> 1) + I; }
// offset of in-tile diagonal d (-15 .. 16) within a tile: diagonal d holds
// 16 - |d| cells, ordered by their smaller coordinate
__host__ __device__ __forceinline__ int cf_mea_doff(int d) {
    return d <= 0 ? (((15 + d) * (16 + d)) >> 1) : 120 + 16 * d - ((d * (d - 1)) >> 1);
}
__host__ __device__ __forceinline__ int cf_mea_tix(int r, int c) {
    const int d = c - r;
    return cf_mea_doff(d) + (d <= 0 ? c : r);
}
// element offset of cell (i, j), i <= j
__device__ __forceinline__ int cf_mea_at(int nb, int i, int j) {
    const int I = i >> 4, J = j >> 4;
    return (cf_mea_tile(nb, I, J - I) << 8) + cf_mea_tix(i & 15, j & 15);
}
// LDS of one team: per-wave tile staging (A transposed, B row-major), then
// up[], the traceback stack, codes and masks.
struct CfMeaLds {
    int o_up, o_st, o_s8, o_fo8, team;
};
__host__ __device__ inline CfMeaLds cf_mea_plan(int Lmax, int nw) {
    CfMeaLds p;
    int o = nw * 2 * CF_MT * CF_MT * 4;
    p.o_up = o; o += cf_align((Lmax + 2) * 4, 16);
    p.o_st = o; o += cf_align(2 * (Lmax + 2) * 4, 16);
    p.o_s8 = o; o += cf_align(Lmax + 3, 16);
    p.o_fo8 = o; o += cf_align(Lmax + 3, 16);
    p.team = o;
    return p;
}
__device__ __forceinline__ void cf_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ __forceinline__ int cf_dbase32(int L, int d) { return d * (L + 1) - ((d * (d - 1)) >> 1); }

template <bool PACK>
__global__ void __launch_bounds__(512) cf_mea_kernel(const CfSeq* __restrict__ seqs, const int* __restrict__ order,
                                                     int nl, int n_gamma, const uint8_t* __restrict__ codes,
                                                     const uint8_t* __restrict__ masks,
                                                     const float* __restrict__ POSTF, const float* __restrict__ gammas,
                                                     float* __restrict__ SC, int64_t tab_stride,
                                                     uint8_t* __restrict__ out, const int64_t* __restrict__ out_off,
                                                     int64_t out_stride, int Lmax, int* __restrict__ queue) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & (WAVE - 1);
    const int nw = PACK ? 1 : (int)(blockDim.x / WAVE);
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int team = PACK ? wid : 0;
    const int wv = PACK ? 0 : wid;  // wave within the team
    const int tid = PACK ? lane : (int)threadIdx.x;
    const int NT = PACK ? WAVE : (int)blockDim.x;
    const int G = PACK ? (int)(blockDim.x / WAVE) : 1;
    const CfMeaLds lp = cf_mea_plan(Lmax, nw);
    unsigned char* tarea = smem + team * lp.team;
    float* As = (float*)tarea + wv * 2 * CF_MT * CF_MT;  // As[k * 16 + r] = A[r][k]
    float* Bs = As + CF_MT * CF_MT;                      // Bs[k * 16 + c] = B[k][c]
    float* up = (float*)(tarea + lp.o_up);
    uint32_t* st = (uint32_t*)(tarea + lp.o_st);
    int8_t* s8 = (int8_t*)(tarea + lp.o_s8);
    int8_t* fo8 = (int8_t*)(tarea + lp.o_fo8);
    // phase 1: the tile elements 4 lane .. 4 lane + 3 a lane loads, as (r, c)
    int er[4], ec[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const int e = 4 * lane + x;
        int d = -15;
        while (d < 15 && cf_mea_doff(d + 1) <= e) ++d;
        const int p = e - cf_mea_doff(d);
        er[x] = d <= 0 ? p - d : p;
        ec[x] = d <= 0 ? p : p + d;
    }
    const int nitem = nl * n_gamma;
    const bool persistent = PACK && queue != nullptr;
    for (int item = persistent ? cf_next_fold(queue) : (PACK ? (int)blockIdx.x * G + team : (int)blockIdx.x);
         item < nitem; item = persistent ? cf_next_fold(queue) : nitem) {
    if (persistent) cf_team_sync<PACK>();  // the team's LDS area is reused item after item
    const int li = item / n_gamma, g = item - li * n_gamma;
    const int bseq = order[li];
    const CfSeq sd = seqs[bseq];
    const int L = sd.L;
    const int nb = cf_mea_nb(L);
    const float* __restrict__ post = POSTF + sd.toff;  // POST(a, b+1) at triangle cell (a, b)
    float* __restrict__ S = SC + g * tab_stride + sd.moff;
    for (int q = tid; q < L + 2; q += NT) {
        s8[q] = (int8_t)codes[sd.coff + q];
        fo8[q] = (masks != nullptr && masks[sd.coff + q] != 0) ? 1 : 0;
    }
    // up[i] = (1 - column sum - row sum) / (2 gamma), subtracted in the
    // reference's order (cpu.py:831-838): POST(j, i) for j = 1 .. i-1, then
    // POST(i, j) for j = i+1 .. L.  The threads of a block of consecutive i
    // step through one triangle diagonal together, so each load is coalesced:
    // the column term j = i - d lies on diagonal d - 1 (d descending), the row
    // term j = i + 1 + m on diagonal m (m ascending).
    const float two_g = 2.0f * gammas[g];
    for (int i0 = 0; i0 < L; i0 += NT) {
        const int i = 1 + i0 + tid;
        const bool on = i <= L;
        float u = 1.0f;
        const int dtop = min(i0 + NT, L) - 1;
        for (int d0 = dtop; d0 >= 1; d0 -= 8) {
            float x[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int d = d0 - t;
                const bool act = on && d >= 1 && d <= i - 1;
                const float v = post[act ? cf_dbase32(L, d - 1) + i - d : 0];
                x[t] = act ? v : 0.0f;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) u -= x[t];  // an inactive slot subtracts +0: u unchanged
        }
        const int mtop = L - 2 - i0;  // L - 1 - i at the block's first i
        for (int m0 = 0; m0 <= mtop; m0 += 8) {
            float x[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int m = m0 + t;
                const bool act = on && m <= L - 1 - i;
                const float v = post[act ? cf_dbase32(L, m) + i : 0];
                x[t] = act ? v : 0.0f;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) u -= x[t];
        }
        if (on) up[i] = u / two_g;
    }
    cf_team_sync<PACK>();

    for (int D = 0; D < nb; ++D) {
        const int ntile = nb - D;
        const int t0D = cf_mea_tile(nb, 0, D) << 8;  // tile (I, I + D) starts at t0D + 256 I
        // ---- phase 1: far bifurcations of the tiles (I, I+D), D >= 2
        if (D >= 2) {
            const int r = lane >> 2, c0 = (lane & 3) << 2;
            for (int I = wv; I < ntile; I += nw) {
                const int J = I + D;
                float4 acc = make_float4(CF_NEG, CF_NEG, CF_NEG, CF_NEG);
                // tile (I, K) of block diagonal K - I, tile (K, J) of J - K
                auto ldA = [&](int K) { return *(const float4*)(S + (cf_mea_tile(nb, I, K - I) << 8) + 4 * lane); };
                auto ldB = [&](int K) { return *(const float4*)(S + (cf_mea_tile(nb, K, J - K) << 8) + 4 * lane); };
                float4 an = ldA(I + 1), bn = ldB(I + 1);
                for (int K = I + 1; K < J; ++K) {
                    cf_wave_sync();  // the previous product has read the staged tiles
                    As[ec[0] * CF_MT + er[0]] = an.x;
                    As[ec[1] * CF_MT + er[1]] = an.y;
                    As[ec[2] * CF_MT + er[2]] = an.z;
                    As[ec[3] * CF_MT + er[3]] = an.w;
                    Bs[er[0] * CF_MT + ec[0]] = bn.x;
                    Bs[er[1] * CF_MT + ec[1]] = bn.y;
                    Bs[er[2] * CF_MT + ec[2]] = bn.z;
                    Bs[er[3] * CF_MT + ec[3]] = bn.w;
                    cf_wave_sync();
                    if (K + 1 < J) {
                        an = ldA(K + 1);
                        bn = ldB(K + 1);
                    }
#pragma unroll
                    for (int k = 0; k < CF_MT; ++k) {
                        const float a = As[k * CF_MT + r];
                        const float4 bv = *(const float4*)(Bs + k * CF_MT + c0);
                        acc.x = fmaxf(acc.x, a + bv.x);
                        acc.y = fmaxf(acc.y, a + bv.y);
                        acc.z = fmaxf(acc.z, a + bv.z);
                        acc.w = fmaxf(acc.w, a + bv.w);
                    }
                }
                float* __restrict__ dst = S + (cf_mea_tile(nb, I, D) << 8);
                dst[cf_mea_tix(r, c0 + 0)] = acc.x;
                dst[cf_mea_tix(r, c0 + 1)] = acc.y;
                dst[cf_mea_tix(r, c0 + 2)] = acc.z;
                dst[cf_mea_tix(r, c0 + 3)] = acc.w;
            }
            cf_team_sync<PACK>();
        }
        // ---- phase 2: the in-tile wavefront
        for (int w = (D == 0 ? 0 : 1 - CF_MT); w < CF_MT; ++w) {
            const int aw = w < 0 ? -w : w;
            const int nr = CF_MT - aw, rlo = w < 0 ? -w : 0;
            const int ncell = ntile * nr;
            const int nt = D == 0 ? w - 1 : CF_MT - 1 + w;  // near terms of every cell of this step
            for (int q = tid; q < ncell; q += NT) {
                const int I = q / nr;
                const int r = rlo + (q - I * nr), c = r + w;
                const int J = I + D;
                const int i = CF_MT * I + r, j = CF_MT * J + c;
                if (j > L) continue;
                const int self = (cf_mea_tile(nb, I, D) << 8) + cf_mea_doff(w) + (w <= 0 ? c : r);
                float best = 0.0f;
                if (j > i) {
                    best = -1.0f;
                    if (D >= 2) best = fmaxf(best, S[self]);
                    best = fmaxf(best, up[i + 1] + S[cf_mea_at(nb, i + 1, j)]);
                    best = fmaxf(best, up[j] + S[cf_mea_at(nb, i, j - 1)]);
                    if (j - i >= 2 && cf_canon(s8[i + 1], s8[j]) && !fo8[i + 1] && !fo8[j])
                        best = fmaxf(best, post[cf_dbase32(L, j - i - 2) + i + 1] + S[cf_mea_at(nb, i + 1, j - 1)]);
                    // near bifurcations: k in block I below row i (S[i][k] in tile (I, I),
                    // S[k][j] in this tile), then k in block J left of column j (S[i][k]
                    // in this tile, S[k][j] in tile (J, J)); past the last term the
                    // index repeats it, which leaves a maximum unchanged.  Slot m of
                    // either kind reads one in-tile diagonal per operand at position
                    // r + a uniform offset, so the lanes of a tile read contiguous cells.
                    // address = lb (per lane: 256 I + r) + a wave-uniform offset per slot
                    const int lb = (I << 8) + r;
                    const int n1 = D == 0 ? CF_MT : CF_MT - 1 - r;  // slots m < n1 are lo
                    for (int mb = 0; mb < nt; mb += 16) {
                        float xa[16], xb[16];
#pragma unroll
                        for (int t = 0; t < 16; ++t) {
                            const int m = min(mb + t, nt - 1);
                            // lo: S[i][i+1+m] (diagonal m+1 of tile (I, I)), S[i+1+m][j]
                            // (diagonal w-1-m of this tile)
                            const int alo = cf_mea_doff(m + 1);
                            const int blo = t0D + cf_mea_doff(w - 1 - m) + min(1 + m, w);
                            // hi, k = 16J + m - 15 + r: S[i][k] (diagonal m-15 of this tile),
                            // S[k][j] (diagonal w-m+15 of tile (J, J))
                            const int ahi = t0D + cf_mea_doff(m - 15) + min(m - 15, 0);
                            const int bhi = (D << 8) + cf_mea_doff(w - m + 15) + m - 15;
                            const bool lo = m < n1;
                            xa[t] = cf_ld(S, lb + (lo ? alo : ahi));
                            xb[t] = cf_ld(S, lb + (lo ? blo : bhi));
                        }
#pragma unroll
                        for (int t = 0; t < 16; ++t) best = fmaxf(best, xa[t] + xb[t]);
                    }
                }
                S[self] = best;
            }
            cf_team_sync<PACK>();
        }
    }

    // ---- traceback (cpu.py:873-895): replay of the reference's candidate order
    uint8_t* __restrict__ o = out + g * out_stride + out_off[bseq];
    for (int q = tid; q < L; q += NT) o[q] = '.';
    cf_team_sync<PACK>();
    if (wv == 0) {
        auto at = [&](int a, int b) { return S[cf_mea_at(nb, a, b)]; };
        if (lane == 0) st[0] = (uint32_t)L << 16;
        int sp = 1;
        cf_wave_sync();
        while (sp > 0) {
            --sp;
            const uint32_t e0 = st[sp];
            const int i = (int)(e0 & 0xffffu), j = (int)(e0 >> 16);
            if (j == i) continue;
            // the cell and its unpaired / pair candidates' operands, loaded together
            const bool two = j - i >= 2;
            const float best = at(i, j);
            const float x1 = at(i + 1, j), x2 = at(i, j - 1);
            const float x3 = at(i + 1, two ? j - 1 : j);
            const float p3 = post[two ? cf_dbase32(L, j - i - 2) + i + 1 : 0];
            if (!(best > -1.0f)) continue;  // the reference's tb stays -1
            int push = 0;
            uint32_t p0 = 0u, p1 = 0u;
            if (up[i + 1] + x1 == best) {
                push = 1; p0 = (uint32_t)(i + 1) | ((uint32_t)j << 16);
            } else if (up[j] + x2 == best) {
                push = 1; p0 = (uint32_t)i | ((uint32_t)(j - 1) << 16);
            } else if (two && cf_canon(s8[i + 1], s8[j]) && !fo8[i + 1] && !fo8[j] && p3 + x3 == best) {
                if (lane == 0) {
                    o[i] = '(';
                    o[j - 1] = ')';
                }
                push = 1; p0 = (uint32_t)(i + 1) | ((uint32_t)(j - 1) << 16);
            } else {
                // the first bifurcation k (ascending) that reaches the maximum
                for (int k0 = i + 1; k0 < j; k0 += WAVE) {
                    const int k = k0 + lane;
                    const bool hit = k < j && at(i, k) + at(k, j) == best;
                    const uint64_t m = __ballot(hit);
                    if (m != 0ull) {
                        const int ks = k0 + (int)__builtin_ctzll(m);
                        push = 2;
                        p0 = (uint32_t)i | ((uint32_t)ks << 16);
                        p1 = (uint32_t)ks | ((uint32_t)j << 16);
                        break;
                    }
                }
            }
            if (lane == 0) {
                if (push >= 1) st[sp] = p0;
                if (push == 2) st[sp + 1] = p1;
            }
            sp += push;
            cf_wave_sync();
        }
    }
    }  // items of this team
}

// =================================================================== FP64
// One FP64 sweep kernel, two semirings:
//  * LSE = false: Viterbi max-plus recurrence (cpu.py:296-394) and the
//    EPS = 1e-6 first-match traceback (cpu.py:396-502).  Max is exact, so the
//    F5 max over k is reduced in parallel; the traceback is one lane replaying
//    the reference's candidate order.
//  * LSE = true: the FP64 inside pass with exact log1p/exp (cpu.py:154-275),
//    terms folded in the reference's order; writes logZ = F5[L].
__device__ __forceinline__ double cf_lse64(double a, double b) {
    if (a < CF_HALF64) return b;
    if (b < CF_HALF64) return a;
    if (a > b) return a + log1p(exp(b - a));
    return b + log1p(exp(a - b));
}
template <bool LSE>
__device__ __forceinline__ void cf_acc(double& acc, double v) {
    if (LSE) acc = cf_lse64(acc, v);
    else if (v > acc) acc = v;
}
// LDS of the FP64 sweep: F5 chain, reduction partials, codes, traceback stack,
// the diagonal's FM2 values, partner masks, the pairable-cell list and the
// FP64 cs / JB / bp tables.
struct CfVitLds {
    int W, o_red, o_s8, o_fo8, o_rs8, o_st, o_fm2, o_pmask, o_alist, o_cnt, o_cs, o_jb, o_bp, total;
};
__host__ __device__ inline CfVitLds cf_vit_plan(int Lmax) {
    CfVitLds p;
    p.W = ((Lmax + 2 + 31) >> 5) + 1;
    int o = (Lmax + 2) * 8;  // f5 at 0
    p.o_red = o; o += 8 * 8;
    p.o_s8 = o; o += Lmax + 3;
    p.o_fo8 = o; o += Lmax + 3;
    p.o_rs8 = o; o += Lmax + 3;
    o = cf_align(o, 16);
    p.o_st = o; o += (4 * Lmax + 16) * 4;
    o = cf_align(o, 16);
    p.o_fm2 = o; o += 2 * (Lmax + 1) * 8;  // FM2 of the two diagonals of a step
    p.o_pmask = o; o += 5 * p.W * 4;
    o = cf_align(o, 16);
    p.o_alist = o; o += 2 * (Lmax + 1) * 2;
    o = cf_align(o, 16);
    p.o_cnt = o; o += 16;
    p.o_cs = o; o += 961 * 8;
    p.o_jb = o; o += 625 * 8;
    p.o_bp = o; o += 25 * 8;
    p.total = cf_align(o, 16);
    return p;
}

template <bool LSE>
__global__ void __launch_bounds__(256) cf_viterbi_kernel(const CfSeq* __restrict__ seqs, const int* __restrict__ order,
                                                         const uint8_t* __restrict__ codes,
                                                         const uint8_t* __restrict__ masks,
                                                         const CfParams<double>* __restrict__ P,
                                                         double* __restrict__ FC, double* __restrict__ FM,
                                                         double* __restrict__ FM1, double* __restrict__ score_out,
                                                         uint8_t* __restrict__ out, const int64_t* __restrict__ out_off,
                                                         int Lmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const CfVitLds lp = cf_vit_plan(Lmax);
    double* f5 = (double*)smem;                          // L+2
    double* red = (double*)(smem + lp.o_red);            // 8 wave partials
    int8_t* s8 = (int8_t*)(smem + lp.o_s8);              // L+3
    int8_t* fo8 = (int8_t*)(smem + lp.o_fo8);            // L+3
    int8_t* rs8 = (int8_t*)(smem + lp.o_rs8);            // L+3: pmask row of each base (4 = pairs with nothing)
    uint32_t* st = (uint32_t*)(smem + lp.o_st);          // 4L+16
    double* fm2b = (double*)(smem + lp.o_fm2);           // FM2 of diagonals s, s+1: [hi * (Lmax+1) + i]
    uint32_t* pmask = (uint32_t*)(smem + lp.o_pmask);    // 5 x W
    uint16_t* alist = (uint16_t*)(smem + lp.o_alist);    // pairable cells of both diagonals: i | hi << 15
    int* cnt = (int*)(smem + lp.o_cnt);
    double* cs64 = (double*)(smem + lp.o_cs);
    double* jb64 = (double*)(smem + lp.o_jb);
    double* bp64 = (double*)(smem + lp.o_bp);
    const int W = lp.W;
    const int bseq = order[blockIdx.x];
    const CfSeq sd = seqs[bseq];
    const int L = sd.L;
    const int tid = threadIdx.x, NT = blockDim.x;
    for (int q = tid; q < L + 2; q += NT) {
        s8[q] = (int8_t)codes[sd.coff + q];
        fo8[q] = (masks != nullptr && masks[sd.coff + q] != 0) ? 1 : 0;
    }
    for (int k = tid; k < 961; k += NT) cs64[k] = P->t[CF_O_CS + k];
    for (int k = tid; k < 625; k += NT) jb64[k] = P->JB[k];
    for (int k = tid; k < 25; k += NT) bp64[k] = P->t[CF_O_BP + k];
    if (tid == 0) { s8[L + 2] = 4; fo8[L + 2] = 0; cnt[0] = 0; cnt[1] = 0; }
    __syncthreads();
    for (int q = tid; q < L + 3; q += NT) rs8[q] = (fo8[q] || s8[q] == 4) ? 4 : s8[q];
    for (int w = tid; w < W; w += NT) {
        uint32_t m[5] = {0u, 0u, 0u, 0u, 0u};
        for (int bit = 0; bit < 32; ++bit) {
            const int q = w * 32 + bit;
            if (q > L + 1) break;
            const int c = s8[q];
            if (fo8[q] || c == 4) continue;
            for (int x = 0; x < 4; ++x)
                if (cf_canon(x, c)) m[x] |= 1u << bit;
        }
        for (int x = 0; x < 5; ++x) pmask[x * W + w] = m[x];
    }
    __syncthreads();
    const double* __restrict__ T = P->t;
    const double mb = T[CF_O_MB], mu = T[CF_O_MU], mp = T[CF_O_MP], eu = T[CF_O_EU], ep = T[CF_O_EP];
    double* __restrict__ fc_t = FC + sd.toff;
    double* __restrict__ fm_t = FM + sd.toff;
    double* __restrict__ fm1_t = FM1 + sd.toff;
    auto allow = [&](int a, int b) -> bool { return cf_canon(s8[a], s8[b]) && !fo8[a] && !fo8[b]; };

    // Diagonals two at a time, as in the FP32 inside sweep: FM2 and FC of
    // diagonal d read FM1 / FM / FC of diagonals <= d-2 only, FM1 and FM read
    // diagonal d-1.
    //   A: FM2 of s and s+1 (a lane takes row i of both, sharing FM1[i][.]);
    //      the pairable cells of both into a list
    //   B: FC of the listed cells: the single-branch window over the set bits
    //      of the partner masks, in the reference's (p ascending, q descending)
    //      order, 8 candidates per batch
    //   C, D: FM1 and FM of s, then of s+1
    // Every cell folds its terms in the reference's order.
    const int FS = Lmax + 1;
    auto fm_fm1 = [&](int d) {
        const int64_t db = cf_dbase(L, d);
        for (int i = tid; i <= L - d; i += NT) {
            const int j = i + d;
            const bool inner = (i > 0) && (j < L);
            double fm1 = CF_NEG64, fmv = CF_NEG64;
            if (inner && d >= 2) {
                const int si1 = s8[i + 1], sj = s8[j];
                if (allow(i + 1, j)) {
                    const double fcin = fc_t[cf_dbase(L, d - 2) + i + 1];
                    if (fcin > CF_HALF64) {
                        double jA = T[CF_O_HC + cf_i2(sj, si1)];
                        if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, si1, s8[j + 1])];
                        if (i > 0) jA = jA + T[CF_O_DR + cf_i3(sj, si1, s8[i])];
                        double v = fcin + jA;
                        v = v + mp;
                        v = v + bp64[cf_i2(si1, sj)];
                        cf_acc<LSE>(fm1, v);
                    }
                }
                const double fm2 = fm2b[(d & 1) * FS + i];
                const double f1n = fm1_t[cf_dbase(L, d - 1) + i + 1];
                if (f1n > CF_HALF64) cf_acc<LSE>(fm1, f1n + mu);
                if (fm2 > CF_HALF64) cf_acc<LSE>(fmv, fm2);
                const double fml = fm_t[cf_dbase(L, d - 1) + i];
                if (fml > CF_HALF64) cf_acc<LSE>(fmv, fml + mu);
                if (fm1 > CF_HALF64) cf_acc<LSE>(fmv, fm1);
            }
            fm1_t[db + i] = fm1;
            fm_t[db + i] = fmv;
        }
    };
    for (int s = 0; s <= L; s += 2) {
        const int step = s >> 1;
        const int n0 = L + 1 - s;
        const int n1 = (s + 1 <= L) ? L - s : 0;
        // ---- A
        // Max-plus (LSE = false): a maximum does not depend on the order of
        // its terms, so when a diagonal has fewer rows than threads, a group
        // of g lanes (a power of two, within one wave) shares a row's split
        // terms (t = 2 + sub, 2 + sub + g, ...) and the group takes the
        // maximum of its partial results.  The FP64 log-sum-exp keeps g = 1.
        int ga = 1;
        if (!LSE)
            while (ga < 8 && n0 * ga * 2 <= NT) ga <<= 1;
        const int suba = tid & (ga - 1);
        for (int i = tid / ga; i < n0; i += NT / ga) {
            const int j0 = i + s;
            const bool inner0 = (i > 0) && (j0 < L);
            const bool inner1 = (i < n1) && (i > 0) && (j0 + 1 < L);
            double a0 = CF_NEG64, a1 = CF_NEG64;
            if (inner0 && s >= 3) {
                // t = 2..s-2 for both cells (FM1[i][i+t] shared), t = s-1 for (i, i+s+1)
                int t = 2 + suba;
                int64_t ia = cf_dbase(L, t) + i;
                int64_t ib0 = cf_dbase(L, s - t) + i + t;
                int64_t ib1 = cf_dbase(L, s + 1 - t) + i + t;
                constexpr int VB = 4;  // FP64 terms per batch (3 x 4 doubles in flight)
                for (; t + (VB - 1) * ga <= s - 2; t += VB * ga) {
                    double x[VB], y0[VB], y1[VB];
#pragma unroll
                    for (int u = 0; u < VB; ++u) {
                        x[u] = cf_ld(fm1_t, ia);
                        y0[u] = cf_ld(fm_t, ib0);
                        y1[u] = cf_ld(fm_t, ib1);
                        const int tu = t + u * ga;
                        ia = cf_dbase(L, tu + ga) + i;
                        ib0 = cf_dbase(L, s - tu - ga) + i + tu + ga;
                        ib1 = cf_dbase(L, s + 1 - tu - ga) + i + tu + ga;
                    }
#pragma unroll
                    for (int u = 0; u < VB; ++u) {
                        if (x[u] > CF_HALF64 && y0[u] > CF_HALF64) cf_acc<LSE>(a0, x[u] + y0[u]);
                        if (x[u] > CF_HALF64 && y1[u] > CF_HALF64) cf_acc<LSE>(a1, x[u] + y1[u]);
                    }
                }
                for (; t <= s - 2; t += ga) {
                    const double x = cf_ld(fm1_t, ia), y0 = cf_ld(fm_t, ib0), y1 = cf_ld(fm_t, ib1);
                    if (x > CF_HALF64 && y0 > CF_HALF64) cf_acc<LSE>(a0, x + y0);
                    if (x > CF_HALF64 && y1 > CF_HALF64) cf_acc<LSE>(a1, x + y1);
                    ia = cf_dbase(L, t + ga) + i;
                    ib0 = cf_dbase(L, s - t - ga) + i + t + ga;
                    ib1 = cf_dbase(L, s + 1 - t - ga) + i + t + ga;
                }
                // t = s - 1 for (i, i+s+1) alone, by the lane of the group it falls to
                if (inner1 && t == s - 1) {
                    const double x = cf_ld(fm1_t, ia), y1 = cf_ld(fm_t, ib1);
                    if (x > CF_HALF64 && y1 > CF_HALF64) cf_acc<LSE>(a1, x + y1);
                }
            }
            if (!LSE) {
                for (int off = 1; off < ga; off <<= 1) {
                    const double o0 = __shfl_xor(a0, off), o1 = __shfl_xor(a1, off);
                    a0 = o0 > a0 ? o0 : a0;
                    a1 = o1 > a1 ? o1 : a1;
                }
                if (suba != 0) continue;
            }
#pragma unroll
            for (int hi = 0; hi < 2; ++hi) {
                if (hi == 1 && i >= n1) break;
                const int d = s + hi;
                const bool inner = hi ? inner1 : inner0;
                fm2b[(d & 1) * FS + i] = inner ? (hi ? a1 : a0) : CF_NEG64;
                if (inner && allow(i, i + d + 1)) {
                    alist[atomicAdd(&cnt[step & 1], 1)] = (uint16_t)(i | (hi << 15));
                } else {
                    fc_t[cf_dbase(L, d) + i] = CF_NEG64;
                }
            }
        }
        __syncthreads();
        // ---- B
        // Max-plus: a group of gb lanes shares a cell's window by rows
        // (p = i + sub, i + sub + gb, ...) when the diagonal pair has fewer
        // pairable cells than threads, and takes the maximum of its parts.
        const int nal = cnt[step & 1];
        if (tid == 0) cnt[(step + 1) & 1] = 0;
        int gb = 1;
        if (!LSE)
            while (gb < 8 && nal * gb * 2 <= NT) gb <<= 1;
        const int subb = tid & (gb - 1);
        for (int idx = tid / gb; idx < nal; idx += NT / gb) {
            const int e0 = alist[idx];
            const int i = e0 & 0x7fff;
            const int d = s + (e0 >> 15);
            const int j = i + d;
            const int si = s8[i], sj1 = s8[j + 1], si1 = s8[i + 1], sj = s8[j];
            const double jB = jb64[cf_i4(si, sj1, si1, sj)];
            double best = jB + T[CF_O_HP + (d < 30 ? d : 30)];
            // Candidates in the reference's order (p ascending, q descending),
            // 8 per batch: the FC loads of a batch are consumed only in its
            // fold; an empty slot reads cell (0, 0), NEG since the first
            // step, and is skipped.
            const int pmax = min(i + CF_CMAX, j);
            const double e_stack = bp64[cf_i2(si1, sj)] + T[CF_O_STACK + cf_i4(si, sj1, si1, sj)];
            const double sn01 = T[CF_O_B01 + sj], sn10 = T[CF_O_B01 + si1], sn11 = T[CF_O_I11 + cf_i2(si1, sj)];
            // The walk holds two rows' partner bits, the current row (cb, row
            // cp, window start cq) and the next non-empty one (nb, np, nq),
            // refilled once per batch; the 8 slots take their candidates from
            // them with selects only (a row change inside the slot loop would
            // cost the wave a row's LDS reads at almost every slot).  A slot is
            // empty only when both buffered rows run out within the batch.
            auto load_row = [&](int pp, uint32_t& b, int& q) {
                q = max(pp + 2, pp - i + j - CF_CMAX);
                b = cf_window(pmask + rs8[pp + 1] * W, q, j);
            };
            uint32_t cb = 0u, nb = 0u;
            int cp = 0, cq = 0, np = 0, nq = 0, pr = i + subb;
            bool more = pr <= pmax;
            while (more) {
                if (cb == 0u) {
                    cb = nb; cp = np; cq = nq;
                    nb = 0u;
                }
                while (cb == 0u && pr <= pmax) {
                    load_row(pr, cb, cq);
                    cp = pr;
                    pr += gb;
                }
                while (nb == 0u && pr <= pmax) {
                    load_row(pr, nb, nq);
                    np = pr;
                    pr += gb;
                }
                double fv[8], ev[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool sw = cb == 0u;
                    cb = sw ? nb : cb;
                    cp = sw ? np : cp;
                    cq = sw ? nq : cq;
                    nb = sw ? 0u : nb;
                    int64_t off = 0;
                    double e = 0.0;
                    if (cb != 0u) {
                        const int p = cp;
                        const int k = 31 - __clz(cb);
                        cb &= ~(1u << k);
                        const int q = cq + k;
                        const int l1 = p - i, l2 = j - q;
                        off = cf_tri(L, p + 1, q - 1);
                        if (l1 == 0 && l2 == 0) {
                            e = e_stack;
                        } else {
                            const int sp1 = s8[p + 1];
                            const double jB2 = jb64[cf_i4(s8[q], sp1, s8[q + 1], s8[p])];
                            const double sn = (l1 == 0 && l2 == 1) ? sn01
                                            : (l1 == 1 && l2 == 0) ? sn10
                                            : (l1 == 1 && l2 == 1) ? sn11 : 0.0;
                            e = cs64[l1 * 31 + l2] + bp64[cf_i2(sp1, s8[q])];
                            e = e + jB;
                            e = e + jB2;
                            e = e + sn;
                        }
                    }
                    fv[u] = fc_t[off];
                    ev[u] = e;
                }
                more = (cb != 0u) || (nb != 0u) || (pr <= pmax);
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (fv[u] > CF_HALF64) cf_acc<LSE>(best, ev[u] + fv[u]);
            }
            if (!LSE) {
                for (int off = 1; off < gb; off <<= 1) {
                    const double o = __shfl_xor(best, off);
                    best = o > best ? o : best;
                }
                if (subb != 0) continue;
            }
            const double fm2 = fm2b[(d & 1) * FS + i];
            if (fm2 > CF_HALF64) {
                double jA = T[CF_O_HC + cf_i2(si, sj1)];
                if (i < L) jA = jA + T[CF_O_DL + cf_i3(si, sj1, si1)];
                if (j > 0) jA = jA + T[CF_O_DR + cf_i3(si, sj1, sj)];
                double v = fm2 + jA;
                v = v + mp;
                v = v + mb;
                cf_acc<LSE>(best, v);
            }
            fc_t[cf_dbase(L, d) + i] = best;
        }
        __syncthreads();
        // ---- C, D
        fm_fm1(s);
        __syncthreads();
        if (s + 1 <= L) {
            fm_fm1(s + 1);
            __syncthreads();
        }
    }
    if (LSE) {
        // F5 chain (cpu.py:259-275), one lane, reference order
        if (tid == 0) {
            f5[0] = 0.0;
            for (int j = 1; j <= L; ++j) {
                double acc = f5[j - 1] + eu;
                const int sj = s8[j];
                for (int k = 0; k < j; ++k) {
                    if (!allow(k + 1, j)) continue;
                    const double fc = fc_t[cf_tri(L, k + 1, j - 1)];
                    if (!(fc > CF_HALF64) || !(f5[k] > CF_HALF64)) continue;
                    const int sk1 = s8[k + 1];
                    double jA = T[CF_O_HC + cf_i2(sj, sk1)];
                    if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, sk1, s8[j + 1])];
                    if (k > 0) jA = jA + T[CF_O_DR + cf_i3(sj, sk1, s8[k])];
                    double v = f5[k] + fc;
                    v = v + ep;
                    v = v + T[CF_O_BP + cf_i2(sk1, sj)];
                    v = v + jA;
                    acc = cf_lse64(acc, v);
                }
                f5[j] = acc;
            }
            score_out[bseq] = f5[L];
        }
        return;
    }
    // F5 (cpu.py:381-394): max over k reduced across the block
    if (tid == 0) f5[0] = 0.0;
    __syncthreads();
    const int nw = (NT + WAVE - 1) / WAVE;
    for (int j = 1; j <= L; ++j) {
        double best = CF_NEG64;
        const int sj = s8[j];
        for (int k = tid; k < j; k += NT) {
            if (allow(k + 1, j) && f5[k] > CF_HALF64) {
                const double fc = fc_t[cf_tri(L, k + 1, j - 1)];
                if (fc > CF_HALF64) {
                    const int sk1 = s8[k + 1];
                    double jA = T[CF_O_HC + cf_i2(sj, sk1)];
                    if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, sk1, s8[j + 1])];
                    if (k > 0) jA = jA + T[CF_O_DR + cf_i3(sj, sk1, s8[k])];
                    double v = f5[k] + fc;
                    v = v + ep;
                    v = v + T[CF_O_BP + cf_i2(sk1, sj)];
                    v = v + jA;
                    if (v > best) best = v;
                }
            }
        }
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            const double o = __shfl_xor(best, off);
            best = o > best ? o : best;
        }
        if ((tid & (WAVE - 1)) == 0) red[tid / WAVE] = best;
        __syncthreads();
        if (tid == 0) {
            double m = f5[j - 1] + eu;
            for (int w = 0; w < nw; ++w)
                if (red[w] > m) m = red[w];
            f5[j] = m;
        }
        __syncthreads();
    }
    // traceback
    uint8_t* __restrict__ o = out + out_off[bseq];
    for (int q = tid; q < L; q += NT) o[q] = '.';
    __syncthreads();
    if (tid != 0) return;
    if (score_out) score_out[bseq] = f5[L];
    const double EPS = 1e-6;
    auto push = [&](int& sp, int typ, int i, int j) { st[sp++] = (uint32_t)typ | ((uint32_t)i << 2) | ((uint32_t)j << 17); };
    auto hc = [&](int a, int b) { return T[CF_O_HC + cf_i2(a, b)]; };
    int sp = 0;
    push(sp, 0, 0, L);
    while (sp > 0) {
        --sp;
        const uint32_t e0 = st[sp];
        const int typ = e0 & 3, i = (e0 >> 2) & 0x7fff, j = (e0 >> 17) & 0x7fff;
        if (typ == 0) {
            if (j == 0) continue;
            const double val = f5[j];
            bool done = false;
            if (f5[j - 1] > CF_HALF64 && fabs(f5[j - 1] + eu - val) <= EPS) { push(sp, 0, 0, j - 1); done = true; }
            if (!done) {
                for (int k = 0; k < j; ++k) {
                    if (allow(k + 1, j) && f5[k] > CF_HALF64) {
                        const double fc = fc_t[cf_tri(L, k + 1, j - 1)];
                        if (!(fc > CF_HALF64)) continue;
                        const int sj = s8[j], sk1 = s8[k + 1];
                        double jA = hc(sj, sk1);
                        if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, sk1, s8[j + 1])];
                        if (k > 0) jA = jA + T[CF_O_DR + cf_i3(sj, sk1, s8[k])];
                        double v = f5[k] + fc;
                        v = v + ep;
                        v = v + T[CF_O_BP + cf_i2(sk1, sj)];
                        v = v + jA;
                        if (fabs(v - val) <= EPS) {
                            o[k] = '(';
                            o[j - 1] = ')';
                            push(sp, 1, k + 1, j - 1);
                            push(sp, 0, 0, k);
                            break;
                        }
                    }
                }
            }
        } else if (typ == 1) {
            const double val = fc_t[cf_tri(L, i, j)];
            const int si = s8[i], sj1 = s8[j + 1], si1 = s8[i + 1], sj = s8[j];
            const double jB = P->JB[cf_i4(si, sj1, si1, sj)];
            const int d = j - i;
            bool done = fabs(jB + T[CF_O_HP + (d <= 30 ? d : 30)] - val) <= EPS;
            if (!done) {
                const int pmax = min(i + CF_CMAX, j);
                for (int p = i; p <= pmax && !done; ++p) {
                    const int l1 = p - i;
                    const int qmin = max(p + 2, p - i + j - CF_CMAX);
                    for (int q = j; q >= qmin; --q) {
                        if (!allow(p + 1, q)) continue;
                        const double fc = fc_t[cf_tri(L, p + 1, q - 1)];
                        if (!(fc > CF_HALF64)) continue;
                        const int l2 = j - q;
                        double e;
                        if (l1 == 0 && l2 == 0) {
                            e = T[CF_O_BP + cf_i2(si1, sj)] + T[CF_O_STACK + cf_i4(si, sj1, si1, sj)];
                        } else {
                            const double jB2 = P->JB[cf_i4(s8[q], s8[p + 1], s8[q + 1], s8[p])];
                            double sn = 0.0;
                            if (l1 == 0 && l2 == 1) sn = T[CF_O_B01 + sj];
                            else if (l1 == 1 && l2 == 0) sn = T[CF_O_B01 + si1];
                            else if (l1 == 1 && l2 == 1) sn = T[CF_O_I11 + cf_i2(si1, sj)];
                            e = T[CF_O_CS + l1 * 31 + l2] + T[CF_O_BP + cf_i2(s8[p + 1], s8[q])];
                            e = e + jB;
                            e = e + jB2;
                            e = e + sn;
                        }
                        if (fabs(e + fc - val) <= EPS) {
                            o[p] = '(';
                            o[q - 1] = ')';
                            push(sp, 1, p + 1, q - 1);
                            done = true;
                            break;
                        }
                    }
                }
            }
            if (!done) {
                double jA = hc(si, sj1);
                if (i < L) jA = jA + T[CF_O_DL + cf_i3(si, sj1, si1)];
                if (j > 0) jA = jA + T[CF_O_DR + cf_i3(si, sj1, sj)];
                for (int k = i + 1; k < j; ++k) {
                    const double x = fm1_t[cf_tri(L, i, k)], y = fm_t[cf_tri(L, k, j)];
                    if (x > CF_HALF64 && y > CF_HALF64) {
                        double v = x + y;
                        v = v + jA;
                        v = v + mp;
                        v = v + mb;
                        if (fabs(v - val) <= EPS) {
                            push(sp, 3, i, k);
                            push(sp, 2, k, j);
                            break;
                        }
                    }
                }
            }
        } else if (typ == 2) {
            const double val = fm_t[cf_tri(L, i, j)];
            bool done = false;
            for (int k = i + 1; k < j; ++k) {
                const double x = fm1_t[cf_tri(L, i, k)], y = fm_t[cf_tri(L, k, j)];
                if (x > CF_HALF64 && y > CF_HALF64 && fabs(x + y - val) <= EPS) {
                    push(sp, 3, i, k);
                    push(sp, 2, k, j);
                    done = true;
                    break;
                }
            }
            if (!done) {
                const double x = fm_t[cf_tri(L, i, j - 1)];
                if (x > CF_HALF64 && fabs(x + mu - val) <= EPS) { push(sp, 2, i, j - 1); done = true; }
            }
            if (!done) {
                const double x = fm1_t[cf_tri(L, i, j)];
                if (x > CF_HALF64 && fabs(x - val) <= EPS) push(sp, 3, i, j);
            }
        } else {
            const double val = fm1_t[cf_tri(L, i, j)];
            bool done = false;
            if (allow(i + 1, j)) {
                const double fc = fc_t[cf_tri(L, i + 1, j - 1)];
                if (fc > CF_HALF64) {
                    const int si1 = s8[i + 1], sj = s8[j];
                    double jA = hc(sj, si1);
                    if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, si1, s8[j + 1])];
                    if (i > 0) jA = jA + T[CF_O_DR + cf_i3(sj, si1, s8[i])];
                    double v = fc + jA;
                    v = v + mp;
                    v = v + T[CF_O_BP + cf_i2(si1, sj)];
                    if (fabs(v - val) <= EPS) {
                        o[i] = '(';
                        o[j - 1] = ')';
                        push(sp, 1, i + 1, j - 1);
                        done = true;
                    }
                }
            }
            if (!done) {
                const double x = fm1_t[cf_tri(L, i + 1, j)];
                if (x > CF_HALF64 && fabs(x + mu - val) <= EPS) push(sp, 3, i + 1, j);
            }
        }
    }
}

// ================================================================== sampling
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based, so each
// (sequence key, sample, draw) has its own stream, independent of batch order.
__device__ __forceinline__ uint4 cf_philox(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}
// Draws of one (sequence, sample) stream: draw n is word n % 4 (in the order
// x, y, z, w) of the Philox block with counter n / 4.  The wave evaluates 64
// blocks at once, lane l holding block base + l, so 256 draws cost one vector
// Philox evaluation and a draw is read from its lane with readlane.  (On the
// wave-uniform counter alone, each block's ten rounds ran on the scalar unit,
// which the traceback's own control already keeps busy.)
struct CfRng {
    uint2 key;
    uint32_t sample, seqk, base, n;
    uint4 blk;
    __device__ void fill() {
        const uint32_t l = threadIdx.x & (WAVE - 1);
        blk = cf_philox(make_uint4(sample, base + l, seqk, 0x5eedu), key);
    }
    __device__ void start() {
        base = 0;
        n = 0;
        fill();
    }
    __device__ float next() {
        if (n == 4 * WAVE) {
            base += WAVE;
            n = 0;
            fill();
        }
        // the four words of the draw's lane into scalar registers, then a
        // scalar select (a select among the vector words would be compiled as
        // a dynamic index into a private-memory copy of the block)
        const int src = (int)(n >> 2);
        const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)blk.x, src);
        const uint32_t w1 = (uint32_t)__builtin_amdgcn_readlane((int)blk.y, src);
        const uint32_t w2 = (uint32_t)__builtin_amdgcn_readlane((int)blk.z, src);
        const uint32_t w3 = (uint32_t)__builtin_amdgcn_readlane((int)blk.w, src);
        const uint32_t c = n & 3u;
        const uint32_t x = c == 0u ? w0 : c == 1u ? w1 : c == 2u ? w2 : w3;
        ++n;
        return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f);  // (0, 1)
    }
};

// Inclusive prefix sum over the wave with DPP moves (no LDS round trips):
// shifts by 1, 2, 4 and 8 within each 16-lane row, then lane 15's row total
// is added to rows 1 and 3 and lane 31's running total to rows 2 and 3.  A
// lane a move does not write keeps the old value, zero.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float cf_dpp_take(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, kRowMask, 0xf, false));
}
__device__ __forceinline__ float cf_wave_incl_sum(float v) {
    v = v + cf_dpp_take<0x111, 0xf>(v);  // row_shr:1
    v = v + cf_dpp_take<0x112, 0xf>(v);  // row_shr:2
    v = v + cf_dpp_take<0x114, 0xf>(v);  // row_shr:4
    v = v + cf_dpp_take<0x118, 0xf>(v);  // row_shr:8
    v = v + cf_dpp_take<0x142, 0xa>(v);  // row_bcast:15
    v = v + cf_dpp_take<0x143, 0xc>(v);  // row_bcast:31
    return v;
}

// Stochastic traceback, one wave per (sequence, sample): the reference's state
// machine (gpu.py:220-477) with its candidate order, walked cooperatively.
// Each state draws one uniform r (Philox4x32-10, keyed by the sequence's key
// and the sample index, so a sample does not depend on its batch); the lanes
// evaluate 64 candidates' probabilities exp(w - total) at a time, a wave
// prefix sum gives the running sums, and the first candidate whose running
// sum reaches r is taken (the last existing one if round-off leaves r above
// the total).  The multiloop branch is resolved through the stored split FM2
// with a second draw.  Sums are formed in a different order than the
// reference's serial loop and the hardware exp2 is used, so sampling is
// checked statistically, not bit for bit.
// One chunk of a candidate scan.  `c` is the running sum before the chunk.
// Returns the lane of the picked candidate, or -1 (then `c` has the chunk's
// sum added and `last` the highest lane holding a candidate, if any).
__device__ __forceinline__ int cf_scan_pick(float p, bool valid, float r, float& c, int& last, int lane) {
    const float cum = cf_wave_incl_sum(valid ? p : 0.0f);
    const float run = c + cum;
    const uint64_t hit = __ballot(valid && r <= run);
    if (hit) return __ffsll((unsigned long long)hit) - 1;
    c = c + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cum), WAVE - 1));
    const uint64_t any = __ballot(valid);
    if (any) last = 63 - __clzll((long long)any);
    return -1;
}

// Per-wave LDS slice of the sampler: the first CF_SAMP_STK entries of the
// traceback stack (deeper entries go to the wave's HBM stack) and, in the
// local form, the sequence's F5 row, base codes and mask.  A traceback is a
// chain of a few hundred dependent reads; with the stack, the exterior row
// and every pairing test in LDS, only the table cells a state scores come
// from L2/HBM, and the next candidate chunk's cells are requested before
// the current chunk is scanned.  The local form is used when the slices
// still let four 512-thread workgroups share a CU.
#define CF_SAMP_STK 128
#define CF_SAMP_WAVES 8
__host__ __device__ inline int cf_samp_row(int lmax) { return (lmax + 2 + 3) & ~3; }
// Per-wave LDS slice: the stack top, and in the local form the F5 row (4n),
// the codes (n), the structure being drawn (n) and the mask (n, if any).
__host__ __device__ inline int cf_samp_pitch(int lmax, bool local, bool masks) {
    const int n = cf_samp_row(lmax);
    return CF_SAMP_STK * 4 + (local ? 6 * n + (masks ? n : 0) : 0);
}

// Eight waves per SIMD (the register cap spills a little cold state) beat
// five unspilled ones: 31 vs 41 ms for 10,000 x 200 nt x 100.
template <bool kLocal>
__global__ void __launch_bounds__(CF_SAMP_WAVES * WAVE, 8) cf_sample_kernel(
    const CfSeq* __restrict__ seqs, const int* __restrict__ sidx, int nseq, int nsamp, uint64_t seed,
    const uint64_t* __restrict__ keys, const uint8_t* __restrict__ codes, const uint8_t* __restrict__ masks,
    const CfParams<float>* __restrict__ P, const float* __restrict__ FC, const float* __restrict__ FM,
    const float* __restrict__ FM1, const float* __restrict__ FM2, const float* __restrict__ F5,
    uint32_t* __restrict__ stack, int stack_cap, uint8_t* __restrict__ out, const int64_t* __restrict__ out_off,
    int lmax, int* __restrict__ status, int* __restrict__ queue, int chunk) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ CfParams<float> sp;
    __shared__ uint8_t wl1[496], wl2[496];  // single-branch window order: l1 ascending, l2 ascending
    for (int t = threadIdx.x; t < CF_NTAB; t += blockDim.x) sp.t[t] = P->t[t];
    for (int t = threadIdx.x; t < 625; t += blockDim.x) sp.JB[t] = P->JB[t];
    for (int t = threadIdx.x; t < 496; t += blockDim.x) {
        int l1 = 0, rem = t;
        while (rem >= 31 - l1) {
            rem -= 31 - l1;
            ++l1;
        }
        wl1[t] = (uint8_t)l1;
        wl2[t] = (uint8_t)rem;
    }
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const int nw = (int)(blockDim.x / WAVE);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int64_t total = (int64_t)nseq * nsamp;
    const int64_t gw = (int64_t)blockIdx.x * nw + wv;
    uint32_t* __restrict__ st = stack + gw * stack_cap;
    const bool has_mask = masks != nullptr;
    unsigned char* slice = smem + (size_t)wv * cf_samp_pitch(lmax, kLocal, has_mask);
    uint32_t* lstk = (uint32_t*)slice;
    float* lf5 = (float*)(slice + CF_SAMP_STK * 4);
    uint8_t* lcod = slice + CF_SAMP_STK * 4 + 4 * cf_samp_row(lmax);
    uint8_t* lstr = lcod + cf_samp_row(lmax);
    uint8_t* lmsk = lstr + cf_samp_row(lmax);
    const float* __restrict__ T = sp.t;
    const float* __restrict__ JBt = sp.JB;
    const float mu = T[CF_O_MU], mp = T[CF_O_MP], eu = T[CF_O_EU], ep = T[CF_O_EP];
    int prev_b = -1;
    // Tasks are (sequence, sample) pairs in sequence-major order, taken
    // `chunk` (1-4) at a time from a counter: waves that are resident keep
    // taking work while other workgroups wait for room beside the outside
    // sweep, and a wave mostly stays on one sequence, refilling its LDS rows
    // once per chunk (10,000 x 200 nt x 100 beside the outside sweep: 80.5 ->
    // 71.0 ms, bench step 146.5 -> 144.9 ms; samples unchanged: each task's
    // stream depends only on the sequence key and the sample index).
    auto next_chunk = [&]() -> int64_t {
        int v = 0;
        if (lane == 0) v = atomicAdd(queue, chunk);
        return (int64_t)__builtin_amdgcn_readfirstlane(v);
    };
    for (int64_t c0 = next_chunk(); c0 < total; c0 = next_chunk())
    for (int64_t task = c0, tend = c0 + chunk < total ? c0 + chunk : total; task < tend; ++task) {
        const int li = (int)(task / nsamp);
        const int smp = (int)(task % nsamp);
        const int b = sidx[li];
        const CfSeq sd = seqs[b];
        const int L = sd.L;
        const uint8_t* __restrict__ s;
        const uint8_t* __restrict__ fo;
        const float* __restrict__ f5;
        if constexpr (kLocal) {
            if (b != prev_b) {
                __builtin_amdgcn_wave_barrier();  // the previous sequence's reads precede the refill
                for (int q = lane; q < L + 2; q += WAVE) {
                    lf5[q] = F5[sd.coff + q];
                    lcod[q] = codes[sd.coff + q];
                    if (has_mask) lmsk[q] = masks[sd.coff + q];
                }
                __builtin_amdgcn_wave_barrier();
                prev_b = b;
            }
            s = lcod;
            fo = has_mask ? lmsk : nullptr;
            f5 = lf5;
        } else {
            s = codes + sd.coff;
            fo = has_mask ? masks + sd.coff : nullptr;
            f5 = F5 + sd.coff;
        }
        auto allow = [&](int a, int c) -> bool {
            return cf_canon(s[a], s[c]) && (fo == nullptr || (fo[a] == 0 && fo[c] == 0));
        };
        const float* __restrict__ fc_t = FC + sd.toff;
        const float* __restrict__ fm_t = FM + sd.toff;
        const float* __restrict__ fm1_t = FM1 + sd.toff;
        const float* __restrict__ fm2_t = FM2 + sd.toff;
        // The structure is drawn into the wave's LDS slice (local form) and
        // written out once, 64 consecutive bytes per store; without the slice
        // one lane records each pair in global memory.  Either way a pair
        // costs one byte store per bracket, not one per lane.
        uint8_t* __restrict__ o = out + out_off[b] + (int64_t)smp * L;
        uint8_t* __restrict__ ostr = kLocal ? lstr : o;
        for (int q = lane; q < L; q += WAVE) ostr[q] = '.';
        if constexpr (kLocal) cf_team_sync<true>();
        auto put_pair = [&](int a, int c) {  // brackets at 0-based a and c
            if (lane == 0) {
                ostr[a] = '(';
                ostr[c] = ')';
            }
        };
        CfRng rng;
        const uint64_t key = keys ? keys[b] : (uint64_t)b;
        rng.key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32) ^ (uint32_t)(key >> 32));
        rng.seqk = (uint32_t)key;
        rng.sample = (uint32_t)smp;
        rng.start();
        // the stack is wave-uniform: every lane stores the same entry, and a
        // popped entry is read into scalar registers.  Depth bound: a popped
        // state is replaced by states over disjoint, non-empty sub-intervals of
        // its own interval (only the exterior state F5(0) is empty), so at most
        // L + 1 entries are ever held; stack_cap = 2 Lmax + 8.  A push beyond
        // the cap is still reported (status bit 1) rather than dropped silently.
        int sp_ = 0;
        bool overflow = false;
        auto push = [&](int typ, int i, int j) {
            const uint32_t e = (uint32_t)typ | ((uint32_t)i << 2) | ((uint32_t)j << 17);
            if (sp_ < CF_SAMP_STK) lstk[sp_++] = e;
            else if (sp_ < stack_cap) st[sp_++] = e;
            else overflow = true;
        };
        // multiloop split of FM2[i][j]: k with weight FM1[i][k] + FM[k][j];
        // the next chunk's cells are requested before the current one is scanned
        auto split = [&](int i, int j, float fm2) -> int {
            const float r2 = rng.next();
            float c2 = 0.0f;
            int last = -1;
            auto cells = [&](int k, float& x, float& y) {
                x = CF_NEG;
                y = CF_NEG;
                if (k < j) {
                    x = fm1_t[cf_tri32(L, i, k)];
                    y = fm_t[cf_tri32(L, k, j)];
                }
            };
            float xn, yn;
            cells(i + 1 + lane, xn, yn);
            for (int k0 = i + 1; k0 < j; k0 += WAVE) {
                const float x = xn, y = yn;
                if (k0 + WAVE < j) cells(k0 + WAVE + lane, xn, yn);
                const bool valid = x > CF_HALF && y > CF_HALF;
                const float p = valid ? __expf(x + y - fm2) : 0.0f;
                int lastc = -1;
                const int t = cf_scan_pick(p, valid, r2, c2, lastc, lane);
                if (t >= 0) return k0 + t;
                if (lastc >= 0) last = k0 + lastc;
            }
            return last;
        };
        push(0, 0, L);
        while (sp_ > 0) {
            --sp_;
            uint32_t e0;
            if (sp_ < CF_SAMP_STK) e0 = lstk[sp_];
            else e0 = st[sp_];
            e0 = __builtin_amdgcn_readfirstlane(e0);
            const int typ = e0 & 3, i = (e0 >> 2) & 0x7fff, j = (e0 >> 17) & 0x7fff;
            if (typ == 0) {  // F5(j): j unpaired, or pair (k+1, j) closing an exterior helix
                if (j == 0) continue;
                const float tot = f5[j];
                const float r = rng.next();
                float c = __expf(f5[j - 1] + eu - tot);
                int pick = -1;
                if (!(r <= c)) {
                    const int sj = s[j];
                    int last = -1;
                    // FC cell of the pair (k+1, j), NEG when k+1 cannot pair with j
                    auto cell = [&](int k) -> float {
                        return (k < j && allow(k + 1, j)) ? fc_t[cf_tri32(L, k + 1, j - 1)] : CF_NEG;
                    };
                    float fcn = cell(lane);
                    for (int k0 = 0; k0 < j; k0 += WAVE) {
                        const int k = k0 + lane;
                        const float fc = fcn;
                        if (k0 + WAVE < j) fcn = cell(k + WAVE);
                        float p = 0.0f;
                        bool valid = false;
                        if (fc > CF_HALF && f5[k] > CF_HALF) {
                            const int sk1 = s[k + 1];
                            float jA = T[CF_O_HC + cf_i2(sj, sk1)];
                            if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, sk1, s[j + 1])];
                            if (k > 0) jA = jA + T[CF_O_DR + cf_i3(sj, sk1, s[k])];
                            const float w = f5[k] + fc + ep + T[CF_O_BP + cf_i2(sk1, sj)] + jA;
                            p = __expf(w - tot);
                            valid = true;
                        }
                        int lastc = -1;
                        const int t = cf_scan_pick(p, valid, r, c, lastc, lane);
                        if (t >= 0) {
                            pick = k0 + t;
                            break;
                        }
                        if (lastc >= 0) last = k0 + lastc;
                    }
                    if (pick < 0) pick = last;
                }
                if (pick < 0) {
                    push(0, 0, j - 1);
                } else {
                    put_pair(pick, j - 1);
                    push(1, pick + 1, j - 1);
                    push(0, 0, pick);
                }
            } else if (typ == 1) {  // FC(i, j): pair (i, j+1) already recorded
                // single-branch window: rows l1 = 0 .. min(30, d-2), q = j - l2 >= p + 2
                const int d = j - i;
                const int nrows = min(CF_CMAX, d - 2) + 1;
                const int tend = nrows > 0 ? 31 * nrows - nrows * (nrows - 1) / 2 : 0;
                // FC cell of candidate t's inner pair (i+l1+1, j-l2), NEG when absent
                auto cell = [&](int t) -> float {
                    if (t >= tend) return CF_NEG;
                    const int pp = i + wl1[t], q = j - wl2[t];
                    return (q >= pp + 2 && allow(pp + 1, q)) ? fc_t[cf_tri32(L, pp + 1, q - 1)] : CF_NEG;
                };
                // the first chunk's cells and the multiloop total are requested with the state's own
                float fcn = cell(lane);
                const float tot = fc_t[cf_tri32(L, i, j)];
                const float fm2 = fm2_t[cf_tri32(L, i, j)];
                const int si = s[i], sj1 = s[j + 1], si1 = s[i + 1], sj = s[j];
                const float jB = JBt[cf_i4(si, sj1, si1, sj)];
                const float r = rng.next();
                float c = __expf(jB + T[CF_O_HP + (d <= 30 ? d : 30)] - tot);
                if (r <= c) continue;  // hairpin
                int pick = -1, last = -1;
                for (int t0 = 0; t0 < tend; t0 += WAVE) {
                    const int t = t0 + lane;
                    const float fc = fcn;
                    if (t0 + WAVE < tend) fcn = cell(t + WAVE);
                    float p = 0.0f;
                    bool valid = false;
                    if (fc > CF_HALF) {
                        const int l1 = wl1[t], l2 = wl2[t];
                        const int pp = i + l1, q = j - l2;
                        float e;
                        if (l1 == 0 && l2 == 0) {
                            e = T[CF_O_BP + cf_i2(si1, sj)] + T[CF_O_STACK + cf_i4(si, sj1, si1, sj)];
                        } else {
                            float sn = 0.0f;
                            if (l1 == 0 && l2 == 1) sn = T[CF_O_B01 + sj];
                            else if (l1 == 1 && l2 == 0) sn = T[CF_O_B01 + si1];
                            else if (l1 == 1 && l2 == 1) sn = T[CF_O_I11 + cf_i2(si1, sj)];
                            e = T[CF_O_CS + l1 * 31 + l2] + T[CF_O_BP + cf_i2(s[pp + 1], s[q])] + jB +
                                JBt[cf_i4(s[q], s[pp + 1], s[q + 1], s[pp])] + sn;
                        }
                        p = __expf(e + fc - tot);
                        valid = true;
                    }
                    int lastc = -1;
                    const int tp = cf_scan_pick(p, valid, r, c, lastc, lane);
                    if (tp >= 0) {
                        pick = t0 + tp;
                        break;
                    }
                    if (lastc >= 0) last = t0 + lastc;
                }
                if (pick < 0) {
                    if (fm2 > CF_HALF) {
                        // multiloop: taken when r passed every single-branch candidate
                        const int kk = split(i, j, fm2);
                        if (kk >= 0) {
                            push(3, i, kk);
                            push(2, kk, j);
                            continue;
                        }
                    }
                    pick = last;
                }
                if (pick >= 0) {
                    const int pp = i + wl1[pick], q = j - wl2[pick];
                    put_pair(pp, q - 1);
                    push(1, pp + 1, q - 1);
                }
            } else if (typ == 2) {  // FM(i, j): split, j unpaired, or one branch (in that order)
                const float tot = fm_t[cf_tri32(L, i, j)];
                const float r = rng.next();
                const float fm2 = fm2_t[cf_tri32(L, i, j)];
                const float wun = fm_t[cf_tri32(L, i, j - 1)];
                const float w1 = fm1_t[cf_tri32(L, i, j)];
                const bool h2 = fm2 > CF_HALF, hu = wun > CF_HALF, h1 = w1 > CF_HALF;
                float c = 0.0f;
                int opt = -1;
                if (h2) { c += __expf(fm2 - tot); if (r <= c || !(hu || h1)) opt = 0; }
                if (opt < 0 && hu) { c += __expf(wun + mu - tot); if (r <= c || !h1) opt = 1; }
                if (opt < 0 && h1) opt = 2;
                if (opt == 0) {
                    const int kk = split(i, j, fm2);
                    if (kk >= 0) { push(3, i, kk); push(2, kk, j); }
                } else if (opt == 1) {
                    push(2, i, j - 1);
                } else if (opt == 2) {
                    push(3, i, j);
                }
            } else {  // FM1(i, j), i < j: pair (i+1, j) or i+1 unpaired
                const bool pc = allow(i + 1, j);
                const float tot = fm1_t[cf_tri32(L, i, j)];
                const float fc = pc ? fc_t[cf_tri32(L, i + 1, j - 1)] : CF_NEG;
                const float wun = fm1_t[cf_tri32(L, i + 1, j)];
                const float r = rng.next();
                bool paired = false;
                if (fc > CF_HALF) {
                    const int si1 = s[i + 1], sj = s[j];
                    float jA = T[CF_O_HC + cf_i2(sj, si1)];
                    if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, si1, s[j + 1])];
                    if (i > 0) jA = jA + T[CF_O_DR + cf_i3(sj, si1, s[i])];
                    const float c = __expf(fc + jA + mp + T[CF_O_BP + cf_i2(si1, sj)] - tot);
                    paired = (r <= c) || !(wun > CF_HALF);
                }
                if (paired) {
                    put_pair(i, j - 1);
                    push(1, i + 1, j - 1);
                } else if (wun > CF_HALF) {
                    push(3, i + 1, j);
                }
            }
        }
        if constexpr (kLocal) {
            cf_team_sync<true>();  // lane 0's brackets precede the copy
            for (int q = lane; q < L; q += WAVE) o[q] = lstr[q];
            cf_team_sync<true>();  // the copy's reads precede the next structure
        }
        if (overflow && lane == 0) atomicOr(status, 1);
    }
}

// ================================================================== gradient
// Expected feature counts: every decomposition event's probability
// exp(outside + event + inside - Z) is added to each parameter-table entry the
// event's score uses.  The counts equal d logZ / d table (the outside pass is
// the adjoint of the inside pass).
//
// Determinism: a sequence's counts are accumulated in LDS as 64-bit fixed
// point (units of 2^-40): integer addition is associative, so the result does
// not depend on which thread adds an event first.  Each workgroup writes its
// sequence's counts (FP64) to its own row of `seqcnt`; cf_grad_reduce_kernel
// then sums the rows in sequence order.  A probability is at most ~1 and a
// sequence's count of one entry at most a few times L, so 2^40 units leave
// more than 2^8 of headroom below 2^63 at L = 32,000; the rounding step
// (2^-41 per event) is far below the FP32 probabilities' own precision.
#define CF_GRAD_FIX 1099511627776.0f  // 2^40
__device__ __forceinline__ void cf_cnt(unsigned long long* cnt, int idx, float p) {
    atomicAdd(&cnt[idx], (unsigned long long)__float2ull_rn(p * CF_GRAD_FIX));
}

__global__ void __launch_bounds__(256) cf_grad_kernel(
    const CfSeq* __restrict__ seqs, const int* __restrict__ order, const uint8_t* __restrict__ codes,
    const uint8_t* __restrict__ masks, const CfParams<float>* __restrict__ P, const float* __restrict__ FC,
    const float* __restrict__ FM, const float* __restrict__ FM1, const float* __restrict__ FM2,
    const float* __restrict__ F5, const float* __restrict__ FCo, const float* __restrict__ FMo,
    const float* __restrict__ FM1o, const float* __restrict__ F5o, double* __restrict__ seqcnt) {
    __shared__ unsigned long long cnt[CF_NTAB];
    const int bseq = order[blockIdx.x];
    const CfSeq sd = seqs[bseq];
    const int L = sd.L;
    const int tid = threadIdx.x, NT = blockDim.x;
    for (int t = tid; t < CF_NTAB; t += NT) cnt[t] = 0ull;
    __syncthreads();
    const uint8_t* __restrict__ s = codes + sd.coff;
    const uint8_t* __restrict__ fo = masks ? masks + sd.coff : nullptr;
    auto allow = [&](int a, int c) -> bool {
        return cf_canon(s[a], s[c]) && (fo == nullptr || (fo[a] == 0 && fo[c] == 0));
    };
    const float* __restrict__ T = P->t;
    const float mb = T[CF_O_MB], mu = T[CF_O_MU], mp = T[CF_O_MP], eu = T[CF_O_EU], ep = T[CF_O_EP];
    const float* __restrict__ fc_t = FC + sd.toff;
    const float* __restrict__ fm_t = FM + sd.toff;
    const float* __restrict__ fm1_t = FM1 + sd.toff;
    const float* __restrict__ fm2_t = FM2 + sd.toff;
    const float* __restrict__ fco_t = FCo + sd.toff;
    const float* __restrict__ fmo_t = FMo + sd.toff;
    const float* __restrict__ fm1o_t = FM1o + sd.toff;
    const float* __restrict__ f5 = F5 + sd.coff;
    const float* __restrict__ f5o = F5o + sd.coff;
    const float Z = f5[L];
    for (int d = 0; d <= L; ++d)
    for (int i = tid; i <= L - d; i += NT) {
        const int j = i + d;
        const int64_t c = cf_dbase(L, d) + i;
        if (i > 0 && j < L && allow(i, j + 1)) {
            const float fco = fco_t[c];
            if (fco > CF_HALF) {
                const int si = s[i], sj1 = s[j + 1], si1 = s[i + 1], sj = s[j];
                const int jbi = cf_i4(si, sj1, si1, sj);
                const float jB = P->JB[jbi];
                const int hp = d <= 30 ? d : 30;
                float pr = expf(fco + jB + T[CF_O_HP + hp] - Z);
                cf_cnt(cnt, CF_O_HC + cf_i2(si, sj1), pr);
                cf_cnt(cnt, CF_O_TM + jbi, pr);
                cf_cnt(cnt, CF_O_HP + hp, pr);
                const int pmax = min(i + CF_CMAX, j);
                for (int p = i; p <= pmax; ++p) {
                    const int l1 = p - i;
                    const int qmin = max(p + 2, p - i + j - CF_CMAX);
                    for (int q = j; q >= qmin; --q) {
                        if (!allow(p + 1, q)) continue;
                        const float fc = fc_t[cf_tri(L, p + 1, q - 1)];
                        if (!(fc > CF_HALF)) continue;
                        const int l2 = j - q;
                        if (l1 == 0 && l2 == 0) {
                            const float e = T[CF_O_BP + cf_i2(si1, sj)] + T[CF_O_STACK + jbi];
                            pr = expf(fco + e + fc - Z);
                            cf_cnt(cnt, CF_O_BP + cf_i2(si1, sj), pr);
                            cf_cnt(cnt, CF_O_STACK + jbi, pr);
                        } else {
                            const int jb2 = cf_i4(s[q], s[p + 1], s[q + 1], s[p]);
                            int sni = -1;
                            if (l1 == 0 && l2 == 1) sni = CF_O_B01 + sj;
                            else if (l1 == 1 && l2 == 0) sni = CF_O_B01 + si1;
                            else if (l1 == 1 && l2 == 1) sni = CF_O_I11 + cf_i2(si1, sj);
                            const float e = T[CF_O_CS + l1 * 31 + l2] + T[CF_O_BP + cf_i2(s[p + 1], s[q])] + jB +
                                            P->JB[jb2] + (sni >= 0 ? T[sni] : 0.0f);
                            pr = expf(fco + e + fc - Z);
                            cf_cnt(cnt, CF_O_CS + l1 * 31 + l2, pr);
                            cf_cnt(cnt, CF_O_BP + cf_i2(s[p + 1], s[q]), pr);
                            cf_cnt(cnt, CF_O_HC + cf_i2(si, sj1), pr);
                            cf_cnt(cnt, CF_O_TM + jbi, pr);
                            cf_cnt(cnt, CF_O_HC + cf_i2(s[q], s[p + 1]), pr);
                            cf_cnt(cnt, CF_O_TM + jb2, pr);
                            if (sni >= 0) cf_cnt(cnt, sni, pr);
                        }
                    }
                }
                const float fm2 = fm2_t[c];
                if (fm2 > CF_HALF) {
                    float jA = T[CF_O_HC + cf_i2(si, sj1)];
                    if (i < L) jA = jA + T[CF_O_DL + cf_i3(si, sj1, si1)];
                    if (j > 0) jA = jA + T[CF_O_DR + cf_i3(si, sj1, sj)];
                    pr = expf(fco + fm2 + jA + mp + mb - Z);
                    cf_cnt(cnt, CF_O_HC + cf_i2(si, sj1), pr);
                    if (i < L) cf_cnt(cnt, CF_O_DL + cf_i3(si, sj1, si1), pr);
                    if (j > 0) cf_cnt(cnt, CF_O_DR + cf_i3(si, sj1, sj), pr);
                    cf_cnt(cnt, CF_O_MP, pr);
                    cf_cnt(cnt, CF_O_MB, pr);
                }
            }
        }
        if (i > 0 && i + 2 <= j && j < L) {
            const float f1o = fm1o_t[c];
            if (f1o > CF_HALF) {
                if (allow(i + 1, j)) {
                    const float fc = fc_t[cf_tri(L, i + 1, j - 1)];
                    if (fc > CF_HALF) {
                        const int si1 = s[i + 1], sj = s[j];
                        float jA = T[CF_O_HC + cf_i2(sj, si1)];
                        if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, si1, s[j + 1])];
                        if (i > 0) jA = jA + T[CF_O_DR + cf_i3(sj, si1, s[i])];
                        const float pr = expf(f1o + fc + jA + mp + T[CF_O_BP + cf_i2(si1, sj)] - Z);
                        cf_cnt(cnt, CF_O_HC + cf_i2(sj, si1), pr);
                        if (j < L) cf_cnt(cnt, CF_O_DL + cf_i3(sj, si1, s[j + 1]), pr);
                        if (i > 0) cf_cnt(cnt, CF_O_DR + cf_i3(sj, si1, s[i]), pr);
                        cf_cnt(cnt, CF_O_MP, pr);
                        cf_cnt(cnt, CF_O_BP + cf_i2(si1, sj), pr);
                    }
                }
                const float f1n = fm1_t[cf_tri(L, i + 1, j)];
                if (f1n > CF_HALF) cf_cnt(cnt, CF_O_MU, expf(f1o + f1n + mu - Z));
            }
            const float fmo = fmo_t[c];
            if (fmo > CF_HALF) {
                const float fml = fm_t[cf_tri(L, i, j - 1)];
                if (fml > CF_HALF) cf_cnt(cnt, CF_O_MU, expf(fmo + fml + mu - Z));
            }
        }
    }
    // exterior loop events
    for (int j = 1 + tid; j <= L; j += NT) {
        cf_cnt(cnt, CF_O_EU, expf(f5o[j] + f5[j - 1] + eu - Z));
        const int sj = s[j];
        for (int k = 0; k < j; ++k) {
            if (!allow(k + 1, j)) continue;
            const float fc = fc_t[cf_tri(L, k + 1, j - 1)];
            if (!(fc > CF_HALF) || !(f5[k] > CF_HALF)) continue;
            const int sk1 = s[k + 1];
            float jA = T[CF_O_HC + cf_i2(sj, sk1)];
            if (j < L) jA = jA + T[CF_O_DL + cf_i3(sj, sk1, s[j + 1])];
            if (k > 0) jA = jA + T[CF_O_DR + cf_i3(sj, sk1, s[k])];
            const float pr = expf(f5o[j] + f5[k] + fc + ep + T[CF_O_BP + cf_i2(sk1, sj)] + jA - Z);
            cf_cnt(cnt, CF_O_EP, pr);
            cf_cnt(cnt, CF_O_BP + cf_i2(sk1, sj), pr);
            cf_cnt(cnt, CF_O_HC + cf_i2(sj, sk1), pr);
            if (j < L) cf_cnt(cnt, CF_O_DL + cf_i3(sj, sk1, s[j + 1]), pr);
            if (k > 0) cf_cnt(cnt, CF_O_DR + cf_i3(sj, sk1, s[k]), pr);
        }
    }
    __syncthreads();
    double* __restrict__ row = seqcnt + (int64_t)bseq * CF_NTAB;
    for (int t = tid; t < CF_NTAB; t += NT) row[t] = (double)cnt[t] * (1.0 / (double)CF_GRAD_FIX);
}

// grad[t] = sum over sequences b = 0, 1, ..., n-1 (in that order) of
// w_b * seqcnt[b][t]: one thread per table entry, so the sum's order is fixed.
__global__ void __launch_bounds__(256) cf_grad_reduce_kernel(const double* __restrict__ seqcnt, int n,
                                                             const float* __restrict__ weights,
                                                             double* __restrict__ grad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= CF_NTAB) return;
    double acc = 0.0;
    for (int b = 0; b < n; ++b) acc += (weights ? (double)weights[b] : 1.0) * seqcnt[(int64_t)b * CF_NTAB + t];
    grad[t] = acc;
}

// ------------------------------------------------------------ explicit inits
template __global__ void cf_inside_kernel<false, false>(const CfSeq*, const int*, int, const uint8_t*, const uint8_t*,
                                                 const CfParams<float>*, float*, float*, float*, float*, float*, int,
                                                 int, unsigned*, int*, int*);
template __global__ void cf_inside_kernel<true, false>(const CfSeq*, const int*, int, const uint8_t*, const uint8_t*,
                                                const CfParams<float>*, float*, float*, float*, float*, float*, int,
                                                int, unsigned*, int*, int*);
template __global__ void cf_outside_kernel<true, false, false>(const CfSeq*, const int*, int, const uint8_t*, const uint8_t*,
                                                        const CfParams<float>*, const float*, const float*, const float*,
                                                        const float*, const float*, float*, float*, float*, float*,
                                                        int, int, unsigned*, int*, int*);
template __global__ void cf_outside_kernel<false, false, false>(const CfSeq*, const int*, int, const uint8_t*, const uint8_t*,
                                                         const CfParams<float>*, const float*, const float*,
                                                         const float*, const float*, const float*, float*, float*,
                                                         float*, float*, int, int, unsigned*, int*, int*);
template __global__ void cf_outside_kernel<false, true, false>(const CfSeq*, const int*, int, const uint8_t*, const uint8_t*,
                                                        const CfParams<float>*, const float*, const float*, const float*,
                                                        const float*, const float*, float*, float*, float*, float*,
                                                        int, int, unsigned*, int*, int*);
template __global__ void cf_inside_kernel<false, true>(const CfSeq*, const int*, int, const uint8_t*, const uint8_t*,
                                                       const CfParams<float>*, float*, float*, float*, float*, float*,
                                                       int, int, unsigned*, int*, int*);
template __global__ void cf_outside_kernel<false, false, true>(const CfSeq*, const int*, int, const uint8_t*,
                                                               const uint8_t*, const CfParams<float>*, const float*,
                                                               const float*, const float*, const float*, const float*,
                                                               float*, float*, float*, float*, int, int, unsigned*,
                                                               int*, int*);
template __global__ void cf_viterbi_kernel<true>(const CfSeq*, const int*, const uint8_t*, const uint8_t*,
                                                 const CfParams<double>*, double*, double*, double*, double*, uint8_t*,
                                                 const int64_t*, int);
template __global__ void cf_viterbi_kernel<false>(const CfSeq*, const int*, const uint8_t*, const uint8_t*,
                                                  const CfParams<double>*, double*, double*, double*, double*,
                                                  uint8_t*, const int64_t*, int);
