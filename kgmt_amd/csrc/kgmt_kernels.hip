// kgmt_kernels.hip — HIP kernels of the KGMT per-iteration expansion path (gfx950).
//
// One iteration of the planner is three launches over a table of queries, two
// for a single query (its last k_update job does k_finish's work):
//   k_propagate  persistent over tiles of children: Philox draws, per-wave
//                sort by step count, RK4 rollout, AABB test against the
//                obstacles staged in LDS (candidate boxes per wave), region
//                and next-frontier flag of every child, the row of each
//                flagged child (dense per 32-child segment), region counters
//                and new subregion bits summed per tile in LDS, per-chunk
//                flag counts.
//                (KGMT.cu:218-299, statePropagator.cu, collisionCheck.cu)
//   k_update     one workgroup per 1024-element job, old-node chunks then
//                child chunks: region scores with a fixed pairwise summation
//                order (each re-activation job for itself), re-activation of
//                old nodes, compaction of the flagged children, tree append
//                and goal test, and the next expansion list (segmented
//                repeat) at offsets found by a decoupled look-back.
//                (Graph.cu:207-358, KGMT.cu:170-181, 305-403, Planner.cu:52-80)
//   k_finish     (batches) one workgroup: closes the iteration of every query (tree
//                size, stop rules, next launch regime; KGMT.cu:84-95, 183-201,
//                399-402) and builds the next iteration's work lists.
// Per plan: k_init (scene build and first iteration's state) before the first
// iteration, and k_paths
// (goal path walk, KGMT.cu:342-361) after the last.
// All sizes stay on the device: the host only polls a completion word that
// the kernels write into host memory.
#include "kgmt_launch.hpp"

namespace kgmt
{

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------
// total[key]++ for lanes with `has`, valid[key]++ for lanes with `has && ok`:
// one atomic per distinct key held by several lanes (leader loop over the
// ballot; children of one parent share regions, and atomics of every lane on
// a shared counter serialise in the L2). The leader rounds stop at the first key that only one lane holds: the
// remaining lanes then add their own counts (their keys are spread, so
// contention is low). Integer sums: the result is the same either way.
__device__ __forceinline__ void wave_count(int32_t* total, int32_t* validc, int key, bool has, bool ok)
{
    uint64_t pending = __ballot(has);
    while(pending)
        {
            const int leader    = __ffsll((unsigned long long)pending) - 1;
            const int k         = __builtin_amdgcn_readlane(key, leader);
            const uint64_t same = __ballot(has && key == k);
            const uint64_t good = __ballot(has && ok && key == k);
            if((int)lane_id() == leader)
                {
                    atomicAdd(&total[k], (int32_t)__popcll(same));
                    if(good) atomicAdd(&validc[k], (int32_t)__popcll(good));
                }
            pending &= ~same;
            if(__popcll(same) < 2) break;
        }
    if((pending >> lane_id()) & 1ull)
        {
            atomicAdd(&total[key], 1);
            if(ok) atomicAdd(&validc[key], 1);
        }
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_or(uint32_t v)
{
    return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

// OR of v over the 64 lanes (all active), as a wave-uniform value: DPP within
// each row of 16 lanes, then the four rows' lane 0 read into scalars.
__device__ __forceinline__ uint32_t wave_or(uint32_t v)
{
    v = dpp_or<0xB1>(v);   // quad_perm [1,0,3,2]
    v = dpp_or<0x4E>(v);   // quad_perm [2,3,0,1]
    v = dpp_or<0x141>(v);  // row_half_mirror
    v = dpp_or<0x140>(v);  // row_mirror
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) | (uint32_t)__builtin_amdgcn_readlane((int)v, 16) |
           (uint32_t)__builtin_amdgcn_readlane((int)v, 32) | (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}

// Sets bit `bit` of `words` for every lane with `want`: one atomicOr per
// distinct word per wave, with the OR of that word's lane masks (a wave's
// children come from few parents and mostly fall in the same few words; one
// atomic per lane on a shared word serialises in the L2). Call with every
// lane of the wave.
__device__ __forceinline__ void wave_set_bits(uint32_t* words, uint32_t bit, bool want)
{
    const uint32_t w = bit >> 5, m = 1u << (bit & 31u);
    uint64_t pending = __ballot(want);
    while(pending)
        {
            const int leader    = __ffsll((unsigned long long)pending) - 1;
            const uint32_t k    = (uint32_t)__builtin_amdgcn_readlane((int)w, leader);
            const bool mine     = want && w == k;
            const uint64_t same = __ballot(mine);
            const uint32_t v    = wave_or(mine ? m : 0u);
            if((int)lane_id() == leader) atomicOr(&words[k], v);
            pending &= ~same;
            if(__popcll(same) < 2) break;  // spread words: one atomic per lane below
        }
    if((pending >> lane_id()) & 1ull) atomicOr(&words[w], m);
}

constexpr int kScoreThreads = 1024;
constexpr int kMaxRegions   = 8192;

// v of lane + C (C <= 8: row_shl, only lanes whose source stays in the row
// are used; C >= 16: shuffle)
template <int C>
__device__ __forceinline__ float lane_down(float v)
{
    if constexpr(C <= 8)
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + C, 0xF, 0xF, false));
    else
        return __shfl_down(v, C, 64);
}

// Active subregions of region r (bits r P .. r P + P - 1 of the bitset): one
// word when P divides 32 (the bits never straddle a word), else count_bits.
__device__ __forceinline__ uint32_t region_active(const uint32_t* __restrict__ act, uint32_t r, uint32_t P)
{
    if(P <= 32u && 32u % P == 0u)
        {
            const uint32_t b = r * P;
            const uint32_t m = P == 32u ? 0xffffffffu : ((1u << P) - 1u) << (b & 31u);
            return (uint32_t)__popc(act[b >> 5] & m);
        }
    return count_bits(act, r * P, P);
}

// Raw score of a region with vc > 0 valid children of tc, cnt of its P
// subregions active (Graph.cu:239-257, Q6: fv^4 = (fv*fv)*(fv*fv))
__device__ __forceinline__ float region_raw_score(const QueryDesc& qd, int32_t vc, int32_t tc, uint32_t cnt)
{
    const float eps = qd.cfg.epsilon;
    // cnt / P, as the product with 1 / P when that is exact (same bits)
    const float cov = qd.invP != 0.0f ? (float)cnt * qd.invP : (float)cnt / (float)qd.P;
    const float vf  = (float)vc;
    const float fv  = (eps + vf) / ((eps + vf) + (float)(tc - vc)) * qd.vol;
    const float f2  = fv * fv;
    const float f4  = f2 * f2;
    const float tt  = (float)tc;
    return f4 / ((1.0f + cov) * (1.0f + tt * tt));
}

// Region scores of one query (block of NT threads, R1 <= kMaxRegions). Q6
// semantics: closed formula, then the pairwise sum a[i] += a[i + st] for
// st = P2/2, ..., 1 over the scores padded with zeros to P2 = pow2(R1). Thread
// t owns regions t + k NT: its inputs are loaded together, and the levels with
// st >= NT add a thread's own registers (i and i + st differ by a multiple of
// NT). The NT partial sums then go through LDS once (a: at least NT floats,
// not read after the call): lane l of the first wave takes sums l + 64 j, so
// the levels 64 <= st < NT add the lane's own registers too, and the last six
// levels pair lanes l and l + st within the wave. Every level adds the same
// two partial sums as the plain tree, so the total has the same bits.
template <int NT>
struct ScoreInputs
{
    static constexpr int KR = kMaxRegions / NT;
    int32_t vc[KR], tc[KR];  // valid and total children of regions t + k NT
    uint32_t cnt[KR];        // their active subregions
};

// the inputs of score_regions_sum for thread t (regions t + k NT), loaded
// together; a caller may issue them before it knows whether it needs them
template <int NT>
__device__ __forceinline__ void score_load(const QueryDesc& qd, const int32_t* __restrict__ total,
                                           const int32_t* __restrict__ validc, const uint32_t* __restrict__ act,
                                           ScoreInputs<NT>& in)
{
    const int R1     = qd.R1;
    const uint32_t P = (uint32_t)qd.P;
    const int t      = (int)threadIdx.x;
#pragma unroll
    for(int k = 0; k < ScoreInputs<NT>::KR; ++k)
        {
            const int r = t + k * NT;
            in.vc[k]    = r < R1 ? validc[r] : 0;
            in.tc[k]    = r < R1 ? total[r] : 0;
            in.cnt[k]   = r < R1 ? region_active(act, (uint32_t)r, P) : 0u;
        }
}

template <int NT>
__device__ void score_regions_sum(const QueryDesc& qd, const ScoreInputs<NT>& in, float* __restrict__ score,
                                  float* a);

template <int NT>
__device__ void score_regions(const QueryDesc& qd, const int32_t* __restrict__ total,
                              const int32_t* __restrict__ validc, const uint32_t* __restrict__ act,
                              float* __restrict__ score, float* a)
{
    ScoreInputs<NT> in;
    score_load<NT>(qd, total, validc, act, in);
    score_regions_sum<NT>(qd, in, score, a);
}

template <int NT>
__device__ void score_regions_sum(const QueryDesc& qd, const ScoreInputs<NT>& in, float* __restrict__ score,
                                  float* a)
{
    static_assert(NT % 64 == 0 && NT <= 1024, "score block");
    constexpr int KR = ScoreInputs<NT>::KR;
    const int R1     = qd.R1;
    const float eps  = qd.cfg.epsilon;
    int P2           = 1;
    while(P2 < R1) P2 <<= 1;
    const int t = (int)threadIdx.x;
    const int32_t* vc   = in.vc;
    const int32_t* tc   = in.tc;
    const uint32_t* cnt = in.cnt;
    float s0[KR], sv[KR];
#pragma unroll
    for(int k = 0; k < KR; ++k)
        {
            const float s = vc[k] > 0 ? region_raw_score(qd, vc[k], tc[k], cnt[k]) : 0.0f;
            s0[k] = sv[k] = s;
        }
    int st = P2 / 2;
    for(; st >= NT; st >>= 1)
        {
            const int m = st / NT;
#pragma unroll
            for(int k = 0; k < KR; ++k)
                if(k < m && k + m < KR) sv[k] = sv[k] + sv[k + m];
        }
    // a[i] for i < NT: the thread's own partial sum (zero past P2)
    a[t] = t < P2 ? sv[0] : 0.0f;
    __syncthreads();
    if(t < 64)
        {
            constexpr int NJ = NT / 64;
            float x[NJ];
#pragma unroll
            for(int j = 0; j < NJ; ++j) x[j] = a[t + 64 * j];
            // levels st = 64 m < min(P2, NT): slot j += slot j + m
#pragma unroll
            for(int m = NJ / 2; m >= 1; m >>= 1)
                if(64 * m < P2)
                    {
#pragma unroll
                        for(int j = 0; j < m; ++j) x[j] = x[j] + x[j + m];
                    }
            float v = x[0];
            if(32 < P2) v = v + lane_down<32>(v);
            if(16 < P2) v = v + lane_down<16>(v);
            if(8 < P2) v = v + lane_down<8>(v);
            if(4 < P2) v = v + lane_down<4>(v);
            if(2 < P2) v = v + lane_down<2>(v);
            if(1 < P2) v = v + lane_down<1>(v);
            if(t == 0) a[0] = v;  // after this wave's reads of a
        }
    __syncthreads();
    const float sum = a[0];
#pragma unroll
    for(int k = 0; k < KR; ++k)
        {
            const int r = t + k * NT;
            if(r < R1) score[r] = vc[k] == 0 ? 1.0f : eps + s0[k] / sum;
        }
}

// fAccept (KGMT.cu:387-388): the host computes (itr * EPSILON) * pow(a, 5)
// with a and the first factor in float and pow's double result, and stores
// the product as float; here the fifth power is (a^2)^2 * a in double (a^2
// is exact) and the product is rounded to float once
__device__ __forceinline__ float faccept(const QueryDesc& qd, uint32_t itr, uint32_t tree, uint32_t next)
{
    const float a   = 1.0f - (float)(tree + next) / (float)qd.cap;
    const double ad = (double)a;
    const double a2 = ad * ad;
    const double a5 = (a2 * a2) * ad;
    return (float)((double)((float)itr * qd.cfg.epsilon) * a5);
}

// children per expansion entry and child count (KGMT.cu:183-201, Q9)
__device__ __forceinline__ void launch_regime(const QueryDesc& qd, uint32_t R, uint32_t tree, uint32_t& per,
                                              uint32_t& nChildren)
{
    const uint32_t B     = qd.cfg.children_per_entry;
    const uint32_t avail = qd.cap - tree;
    if((uint64_t)R * B <= avail)
        per = B;
    else
        {
            const uint32_t qv = avail / R;
            per               = qv > B ? B : qv;
            if(per == 0) per = 1;
        }
    const uint64_t nc = (uint64_t)R * per;
    nChildren         = nc > qd.cap ? qd.cap : (uint32_t)nc;
}

// Work lists of the next iteration into buffer `parity` (block of NT threads,
// run by the last workgroup to arrive).
template <int NT>
__device__ void build_worklists(const QueryState* sts, const WorkLists& W, uint32_t parity, uint32_t* sbuf)
{
    const uint32_t nq = W.nq;
    uint32_t* tileStart = W.tileStart + parity * (nq + 1);
    uint32_t* jobStart  = W.jobStart + parity * (nq + 1);
    uint32_t carryT = 0, carryJ = 0, carryA = 0;
    for(uint32_t base = 0; base < nq; base += NT)
        {
            const uint32_t q = base + threadIdx.x;
            uint32_t tiles = 0, jobs = 0, alive = 0;
            if(q < nq)
                {
                    const uint32_t done = __hip_atomic_load(&sts[q].doneN, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if(!done)
                        {
                            const uint32_t nc = __hip_atomic_load(&sts[q].nChildrenN, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT);
                            const uint32_t tr = __hip_atomic_load(&sts[q].treeSizeN, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT);
                            tiles             = (nc + W.tile - 1) / W.tile;
                            jobs              = (nc + kChunk - 1) / kChunk + (tr + kChunk - 1) / kChunk;
                            alive             = 1;
                        }
                }
            uint32_t tT, tJ, tA;
            const uint32_t eT = block_excl_scan<NT>(tiles, sbuf, tT);
            const uint32_t eJ = block_excl_scan<NT>(jobs, sbuf, tJ);
            block_excl_scan<NT>(alive, sbuf, tA);
            if(q < nq)
                {
                    tileStart[q] = carryT + eT;
                    jobStart[q]  = carryJ + eJ;
                }
            carryT += tT, carryJ += tJ, carryA += tA;
        }
    if(threadIdx.x == 0)
        {
            tileStart[nq]     = carryT;
            jobStart[nq]      = carryJ;
            W.active[parity]  = carryA;
            *W.counter        = 0;
            __hip_atomic_store(W.allDone, carryA == 0 ? 1u : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
}

// Returns true in every thread of the last workgroup to arrive (all writes of
// the earlier ones are visible to it afterwards).
__device__ __forceinline__ bool last_arrival(uint32_t* counter, uint32_t expected, uint32_t* sflag)
{
    __threadfence();
    __syncthreads();
    if(threadIdx.x == 0)
        {
            const uint32_t prev = atomicAdd(counter, 1u);
            const bool last     = prev == expected - 1;
            // every workgroup has arrived: the counter starts at zero for the
            // next launch
            if(last) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            *sflag = last ? 1u : 0u;
        }
    __syncthreads();
    const bool last = *sflag != 0;
    if(last) __threadfence();
    return last;
}

// ---------------------------------------------------------------------------
// init (KGMT.cu:50-81): per-query part of k_init, defined after the scene build
// ---------------------------------------------------------------------------
// the cleared region counters, scores, subregion bits, chunk counts and
// job words of one query: thread `rank` of `ranks` clears every ranks-th word
__device__ void clear_query(const QueryDesc& qd, const Pools& P, uint32_t rank, uint32_t ranks)
{
    for(uint32_t r = rank; r < (uint32_t)qd.R1; r += ranks)
        {
            P.total[qd.regOff + r]  = 0;
            P.validc[qd.regOff + r] = 0;
            P.score[qd.regOff + r]  = 0.0f;
        }
    for(uint32_t w = rank; w < (uint32_t)qd.R2words; w += ranks)
        {
            P.actPrev[qd.subOff + w] = 0u;
            P.actNext[qd.subOff + w] = 0u;
        }
    for(uint32_t c = rank; c <= qd.nChunks; c += ranks) P.partA[qd.chunkOff + c] = 0u;
    for(uint32_t i = rank; i < 4 * (qd.nChunks + 1); i += ranks) P.look[4 * (size_t)qd.chunkOff + i] = 0ull;
}

// root and the first iteration's state of one query (one workgroup)
__device__ void root_query(const QueryDesc& qd, QueryState& st, const Pools& P)
{
    const uint32_t rootRep = qd.cfg.root_repeat < qd.cap ? qd.cfg.root_repeat : qd.cap;
    for(uint32_t k = threadIdx.x; k < rootRep; k += kBlock) P.repIdx[qd.nodeOff + k] = 0u;
    for(int d = threadIdx.x; d < qd.D; d += kBlock) P.treeS[qd.sampleOff + (size_t)d * qd.cap] = qd.start[d];
    if(threadIdx.x == 0)
        {
            P.parent[qd.nodeOff]   = -1;
            P.grand[qd.nodeOff]    = -1;
            P.anc[qd.nodeOff]      = make_int4(-1, -1, -1, -1);
            P.cost[qd.nodeOff]     = 0.0f;
            P.frontier[qd.nodeOff] = 1;
            int ix, iy, iz;
            P.r1node[qd.nodeOff] = region_of(qd, qd.start[0], qd.start[1], qd.start[2], ix, iy, iz);
            st.itr               = 1;
            st.treeSize          = 1;
            st.R                 = qd.cfg.root_repeat;
            uint32_t per, nc;
            launch_regime(qd, st.R, 1, per, nc);
            st.per       = per;
            st.nChildren = nc;
            st.nAcc = 0, st.next = 0;
            st.fAccept = 0.0f;
            st.done = 0, st.success = 0;
            st.goalIdx  = kNoGoal;
            st.pathRows = 0;
            st.pathCost = 0.0f, st.pathLen = 0.0f;
            st.expanded  = 0;
            st.flagged   = 0;
            st.finishItr = 0;
            st.itrN = st.itr, st.treeSizeN = st.treeSize, st.nChildrenN = st.nChildren, st.perN = st.per;
            st.doneN = 0, st.successN = 0, st.finishItrN = 0, st.padN = 0;
            st.expandedN = 0, st.flaggedN = 0;
        }
}

// ---------------------------------------------------------------------------
// propagate: persistent over 256-child tiles
// ---------------------------------------------------------------------------
// LDS of one propagate tile: the draws of its NT / K children and the order
// in which the lanes take them. The per-wave candidate lists of the two-phase
// check exist only with one lane per child (the only shape that runs it), so
// they take no LDS, and no occupancy, from the other shapes.
template <int NT, int K>
struct TileLds
{
    static constexpr int kCandWaves = K == 1 ? NT / 64 : 1;
    static constexpr int kCandBoxes = K == 1 ? kCand : 1;
    uint32_t hist[NT];      // children per step bucket, then bucket offsets
    uint16_t slot[NT];      // lane -> child of the tile
    float ctrl[KGMT_MAX_CONTROL][NT];
    int32_t steps[NT];
    float ufront[NT];
    uint32_t sbuf[NT / 64 + 1];
    float4 cand[kCandWaves][2 * kCandBoxes];  // per-wave candidate boxes (two-phase check)
    // per-tile sums of the region counters and ORs of the subregion bitset
    // words, flushed with one global atomic per distinct key (open addressing,
    // key -1 = empty; at most NT distinct keys per tile, so a probe always ends)
    int32_t rkey[NT];
    uint32_t rtot[NT], rval[NT];
    int32_t wkey[NT];
    uint32_t wmask[NT];
    // flagged children of the tile, one bit per child, by 32-child segment;
    // two buffers alternating by tile, so a tile empties the other one
    uint32_t fbits[2][NT / 32];
};

// empties thread t's slot of both tables and of flag buffer `fb`
template <int NT, int K>
__device__ __forceinline__ void clear_tables(TileLds<NT, K>& L, uint32_t fb)
{
    const uint32_t t = threadIdx.x;
    L.rkey[t] = -1, L.rtot[t] = 0, L.rval[t] = 0;
    L.wkey[t] = -1, L.wmask[t] = 0;
    if(t < NT / 32) L.fbits[fb][t] = 0;
}

// slot of `key` in an NT-entry open-addressing table of LDS keys (claimed by
// compare-and-swap); the table holds fewer than NT other keys
template <int NT>
__device__ __forceinline__ uint32_t tab_slot(int32_t* keys, int32_t key)
{
    uint32_t slot = ((uint32_t)key * 2654435761u) & (uint32_t)(NT - 1);
    for(;;)
        {
            const int32_t old = atomicCAS(&keys[slot], -1, key);
            if(old == -1 || old == key) return slot;
            slot = (slot + 1) & (uint32_t)(NT - 1);
        }
}

// One tile of TC = NT / K children u = base .. base + TC - 1 of a query (TC
// divides kChunk, so a tile never straddles two chunks), K lanes per child.
// Every child writes its flag and region; a flagged child (the only ones the
// tree append reads) also writes its row and parent, at the next free slot of
// its 32-child segment (segment start + flagged children before it in the
// segment), so the rows of a segment are dense. fb: this tile's flag buffer.
// The children are drawn in tile order, then handed to the lane groups
// longest-first (counting sort on the step count) so that each wave's groups
// integrate about the same number of steps. The K lanes of a group integrate
// the same edge (same operations, same bits) and split its box tests; the
// group's first lane writes the child. Each child still writes only its own
// slot and the counters are integer atomics, so the outputs do not depend on
// the order or on K.
template <int MODEL, int NT, int K>
__device__ __forceinline__ void child_tile(const QueryDesc& qd, const QueryState& st, uint32_t base,
                                           const Pools& P, float4* sObs, TileLds<NT, K>& L, uint32_t fb,
                                           bool stage)
{
    constexpr int S = ModelDims<MODEL>::S, C = ModelDims<MODEL>::C, D = S + C + 1;
    constexpr uint32_t TC = NT / K;
    const uint32_t nC  = st.nChildrenN;
    const uint32_t cap = qd.cap;
    const size_t nodeO = qd.nodeOff;
    const uint32_t t   = threadIdx.x;
    const uint32_t per = st.perN;
    const float* T     = P.treeS + qd.sampleOff;
    // K == 1: the parent of child t (tile order) and its state, loaded while
    // the draws and the sort run; after the sort the wave passes them to the
    // lanes that take the children, through its candidate area (written
    // again only by the two-phase check)
    static_assert(K > 1 || (2 + S) * 64 <= 8 * TileLds<NT, K>::kCandBoxes, "parent stage fits a wave's candidates");
    uint32_t pp = 0;
    int32_t pg  = -1;
    float ps[S];
#pragma unroll
    for(int d = 0; d < S; ++d) ps[d] = 0.0f;
    const bool pre = K == 1 && t < TC && base + t < nC;
    if(pre) pp = P.repIdx[nodeO + (base + t) / per];
    // the query's scene into LDS on the workgroup's first tile of it (its
    // loads in flight with the parent index)
    if(stage)
        {
            stage_scene(P.obs, qd.obsOff, qd.nGrp, sObs);
            __syncthreads();
        }
    // draws of the child in tile order
    {
        const uint32_t un = base + t;
        float uc[C];
        int32_t steps = 0;
        float uf      = 0.0f;
        // K == 1: each wave sorts its own 64 children (they come from one or
        // two parents, so the wave's edges stay close together and share few
        // candidate boxes); K > 1: the whole tile is sorted
        constexpr uint32_t BW = K == 1 ? 64u : (uint32_t)NT;  // buckets per sorted group
        const uint32_t g0     = K == 1 ? (t / 64u) * 64u : 0u;
        uint32_t bucket       = g0 + BW - 1;                  // past the end: last
        if(t < TC && un < nC)
            {
                draw_child<C>(qd.cfg, qd.cfg.query_id, st.itrN, un, uc, steps, uf);
                const int32_t k = steps - qd.cfg.steps_min;
                bucket = g0 + (BW - 2) - (uint32_t)(k < (int32_t)BW - 2 ? k : (int32_t)BW - 2);
            }
        L.hist[t] = 0;
        if constexpr(K == 1)
            {
                // The wave's 64 buckets hold exactly its 64 lanes, so its
                // offsets are a wave scan from g0, and every LDS word written
                // here is read only by this wave: no workgroup barrier (the
                // LDS runs one wave's operations in order; the wave barriers
                // keep the compiler from reordering them).
                wave_sync();
                const uint32_t rank = atomicAdd(&L.hist[bucket], 1u);
                wave_sync();
                const uint32_t h = L.hist[t];
                L.hist[t]        = g0 + wave_incl_scan(h) - h;
                wave_sync();
                L.slot[L.hist[bucket] + rank] = (uint16_t)t;
#pragma unroll
                for(int k = 0; k < C; ++k) L.ctrl[k][t] = uc[k];
                L.steps[t]  = steps;
                L.ufront[t] = uf;
                if(pre)
                    {
#pragma unroll
                        for(int d = 0; d < S; ++d) ps[d] = T[(size_t)d * cap + pp];
                        pg = P.parent[nodeO + pp];
                    }
                // parent, grandparent and state of child t at word j * 64 + t % 64
                float* stg          = reinterpret_cast<float*>(L.cand[t / 64]);
                const uint32_t l    = t % 64u;
                stg[l]              = __uint_as_float(pp);
                stg[64 + l]         = __int_as_float(pg);
#pragma unroll
                for(int d = 0; d < S; ++d) stg[(2 + d) * 64 + l] = ps[d];
                wave_sync();
            }
        else
            {
                __syncthreads();
                uint32_t rank = 0;
                if(t < TC) rank = atomicAdd(&L.hist[bucket], 1u);
                __syncthreads();
                uint32_t tot;
                const uint32_t off = block_excl_scan<NT>(L.hist[t], L.sbuf, tot);
                L.hist[t]          = off;
                __syncthreads();
                if(t < TC)
                    {
                        L.slot[L.hist[bucket] + rank] = (uint16_t)t;
#pragma unroll
                        for(int k = 0; k < C; ++k) L.ctrl[k][t] = uc[k];
                        L.steps[t]  = steps;
                        L.ufront[t] = uf;
                    }
                __syncthreads();
            }
    }
    const uint32_t i    = L.slot[t / K];
    const uint32_t u    = base + i;
    const bool has      = u < nC;
    const bool lead     = K == 1 || t % K == 0;
    int r1              = 0;
    bool ok = false, fn = false;
    bool setBit   = false;  // valid child in a subregion unseen at iteration start
    uint32_t r2bit = 0;
    uint32_t p         = 0;
    int32_t gp         = -1;  // parent of the parent
    float s[S], uc[C];
    int32_t steps = 0;
    float uf      = 0.0f;
#pragma unroll
    for(int d = 0; d < S; ++d) s[d] = 0.0f;
#pragma unroll
    for(int k = 0; k < C; ++k) uc[k] = 0.0f;
    if(has)
        {
            if constexpr(K == 1)
                {
                    // child i is one of this wave's 64
                    const float* stg = reinterpret_cast<const float*>(L.cand[t / 64]);
                    const uint32_t l = i % 64u;
                    p                = __float_as_uint(stg[l]);
                    gp               = __float_as_int(stg[64 + l]);
#pragma unroll
                    for(int d = 0; d < S; ++d) s[d] = stg[(2 + d) * 64 + l];
                }
            else
                {
                    p = P.repIdx[nodeO + u / per];
#pragma unroll
                    for(int d = 0; d < S; ++d) s[d] = T[(size_t)d * cap + p];
                    gp = P.parent[nodeO + p];
                }
#pragma unroll
            for(int k = 0; k < C; ++k) uc[k] = L.ctrl[k][i];
            steps = L.steps[i];
            uf    = L.ufront[i];
        }
    if constexpr(K == 1) wave_sync();  // the stage is read before the check writes candidates
    const Scene sc{sObs, (int)qd.nGrp, (int)(t % K)};
    if constexpr(K == 1 && MODEL != KGMT_MODEL_QUADCOPTER)
        ok = propagate_two_phase<MODEL>(qd, s, uc, steps, sc, L.cand[t / 64], has);  // every lane
    else if(has)
        ok = propagate<MODEL, K>(qd, s, uc, steps, sc);
    if(has)
        {
            if(lead)
                {
                    // region ids and the loads they need first: a load issued
                    // after the stores would wait for them (one in-order counter)
                    int ix, iy, iz;
                    r1 = region_of(qd, s[0], s[1], s[2], ix, iy, iz);
                    if(ok)
                        {
                            const int r2    = subregion_of(qd, s[0], s[1], s[2], r1, ix, iy, iz);
                            const bool seen = test_bit(P.actPrev + qd.subOff, (uint32_t)r2);
                            fn              = (uf < P.score[qd.regOff + r1]) || !seen;
                            setBit          = !seen;
                            r2bit           = (uint32_t)r2;
                        }
                    if(u % per == 0) P.frontier[nodeO + p] = 0;
                    P.r1child[nodeO + u] = r1;
                    P.fnext[nodeO + u]   = fn ? 1 : 0;
                    if(fn) atomicOr(&L.fbits[fb][i >> 5], 1u << (i & 31u));
                }
        }
    // region counters and new subregion bits: summed per tile in LDS, then
    // one global atomic per distinct region / bitset word
    if(has && lead)
        {
            const uint32_t e = tab_slot<NT>(L.rkey, r1);
            atomicAdd(&L.rtot[e], 1u);
            if(ok) atomicAdd(&L.rval[e], 1u);
        }
    if(setBit)
        {
            const uint32_t e = tab_slot<NT>(L.wkey, (int32_t)(r2bit >> 5));
            atomicOr(&L.wmask[e], 1u << (r2bit & 31u));
        }
    // the tile lies inside one 1024-element chunk
    const uint64_t fm = __ballot(fn);
    if(lane_id() == 0 && fm) atomicAdd(&P.partA[qd.chunkOff + base / kChunk], (uint32_t)__popcll(fm));
    __syncthreads();
    if(fn)
        {
            // row slot: segment start + flagged children before it; the row
            // (parent bits, sample, zero padding) in whole float4 stores
            constexpr int RW = row_words(D);
            const uint32_t v = base + (i & ~31u) + (uint32_t)__popc(L.fbits[fb][i >> 5] & ((1u << (i & 31u)) - 1u));
            float w[RW];
            w[0] = __int_as_float((int32_t)p);
            w[1] = __int_as_float(gp);
#pragma unroll
            for(int d = 0; d < S; ++d) w[2 + d] = s[d];
#pragma unroll
            for(int k = 0; k < C; ++k) w[2 + S + k] = uc[k];
            w[D + 1] = qd.cfg.dt * (float)steps;
#pragma unroll
            for(int e = D + 2; e < RW; ++e) w[e] = 0.0f;
            float4* row = reinterpret_cast<float4*>(P.childR + qd.rowOffC + (size_t)v * RW);
#pragma unroll
            for(int j = 0; j < RW / 4; ++j) row[j] = make_float4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
        }
    {
        const int32_t rk = L.rkey[t];
        if(rk >= 0)
            {
                atomicAdd(&P.total[qd.regOff + rk], (int32_t)L.rtot[t]);
                if(L.rval[t]) atomicAdd(&P.validc[qd.regOff + rk], (int32_t)L.rval[t]);
            }
        const int32_t wk = L.wkey[t];
        if(wk >= 0) atomicOr(&P.actNext[qd.subOff + (uint32_t)wk], L.wmask[t]);
        // empty for the next tile, which uses the other flag buffer (this
        // tile's is emptied by the next one, after the barrier below)
        clear_tables(L, fb ^ 1u);
    }
    __syncthreads();  // L is reused by the next tile
}

template <int MODEL, int NT, int K>
__global__ __launch_bounds__(NT, 4) void k_propagate(const QueryDesc* __restrict__ qds,
                                                  QueryState* __restrict__ sts, Pools P, WorkLists W,
                                                  uint32_t parity)
{
    static_assert(NT % 64 == 0 && kChunk % (NT / K) == 0, "tile size");
    extern __shared__ float4 sObs[];
    __shared__ TileLds<NT, K> L;
    const uint32_t nq        = W.nq;
    // the first workgroup commits the next-state fields (written when the
    // last iteration closed) to the current ones, which the update kernel
    // reads; propagate itself reads only the next-state fields
    if(blockIdx.x == 0)
        for(uint32_t q = threadIdx.x; q < nq; q += NT) commit_state(sts[q]);
    const uint32_t* tileList = W.tiles(parity);
    const uint32_t nTiles    = tileList[nq];
    uint32_t staged          = 0xffffffffu;
    uint32_t fb              = 0;  // flag buffer of the next tile
    // the tables and both flag buffers start empty (each tile empties them
    // again after its flush; the first tile's staging barrier comes before
    // any insert)
    clear_tables(L, 0u);
    clear_tables(L, 1u);
    for(uint32_t t = blockIdx.x; t < nTiles; t += gridDim.x)
        {
            const uint32_t q    = locate(tileList, nq, t);
            if(sts[q].doneN) continue;  // stale list entry of a finished query
            const QueryDesc& qd = qds[q];
            // a new query's scene: the tile stages it (after the barrier
            // that ends the last tile, so no wave still reads the old one)
            const bool stage = q != staged;
            staged           = q;
            const uint32_t base = (t - tileList[q]) * (NT / K);
            if constexpr(MODEL >= 0)
                child_tile<MODEL, NT, K>(qd, sts[q], base, P, sObs, L, fb, stage);
            else
                {
                    switch(qd.cfg.model)
                        {
                            case KGMT_MODEL_DOUBLE_INTEGRATOR:
                                child_tile<KGMT_MODEL_DOUBLE_INTEGRATOR, NT, K>(qd, sts[q], base, P, sObs, L, fb, stage);
                                break;
                            case KGMT_MODEL_DUBINS_AIRPLANE:
                                child_tile<KGMT_MODEL_DUBINS_AIRPLANE, NT, K>(qd, sts[q], base, P, sObs, L, fb, stage);
                                break;
                            case KGMT_MODEL_DUBINS_5D:
                                child_tile<KGMT_MODEL_DUBINS_5D, NT, K>(qd, sts[q], base, P, sObs, L, fb, stage);
                                break;
                            // (a batch never mixes planar and 3D queries, so
                            // a mixed batch holds no unicycle)
                            default: child_tile<KGMT_MODEL_QUADCOPTER, NT, K>(qd, sts[q], base, P, sObs, L, fb, stage); break;
                        }
                }
            fb ^= 1u;
        }
}

// ---------------------------------------------------------------------------
// closing an iteration
// ---------------------------------------------------------------------------
// End of an iteration for one query (one thread): tree size, stop rules,
// next launch regime. The goal path is walked once the plan has ended
// (k_paths).
__device__ void finish_query(const QueryDesc& qd, QueryState& st, uint32_t R)
{
    const uint32_t tree = st.treeSize + st.next;
    uint32_t itr = st.itr, per = st.per, nc = st.nChildren, done = 0, success = 0, fin = st.finishItr;
    if(st.goalIdx != kNoGoal)
        {
            success = 1;
            done    = 1;
            fin     = itr;
        }
    else if(itr >= qd.cfg.max_iter || tree == qd.cap)
        {
            // out of iterations, or a full tree that can take no further node
            done = 1;
            fin  = itr;
        }
    else
        {
            launch_regime(qd, R, tree, per, nc);
            itr = itr + 1;
        }
    st.treeSizeN  = tree;
    st.expandedN  = st.expanded + st.nChildren;
    st.flaggedN   = st.flagged + st.nAcc;
    st.itrN       = itr;
    st.perN       = per;
    st.nChildrenN = nc;
    st.doneN      = done;
    st.successN   = success;
    st.finishItrN = fin;
}

// One more iteration closed in this plan: the count for the host, which
// stays a few launch units ahead of it (one thread per iteration: the closing
// job of a single query, k_finish for a batch)
__device__ __forceinline__ void count_closed(const WorkLists& W)
{
    const uint32_t n = *W.closedDev + 1u;
    *W.closedDev     = n;
    __hip_atomic_store(W.closed, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Work lists of a single-query table for its next iteration (parity
// buffer), from the next-state fields (one thread).
__device__ __forceinline__ void single_worklists(const QueryState& st, const WorkLists& W, uint32_t parity)
{
    uint32_t* tileStart = W.tileStart + parity * 2;
    uint32_t* jobStart  = W.jobStart + parity * 2;
    const bool alive    = st.doneN == 0;
    tileStart[0]        = 0;
    jobStart[0]         = 0;
    tileStart[1]        = alive ? (st.nChildrenN + W.tile - 1) / W.tile : 0u;
    jobStart[1] = alive ? (st.nChildrenN + kChunk - 1) / kChunk + (st.treeSizeN + kChunk - 1) / kChunk : 0u;
    W.active[parity] = alive ? 1u : 0u;
}

__device__ void walk_path(const QueryDesc& qd, QueryState& st, const Pools& P, float* __restrict__ out, int32_t goal);

// goal -> root rows of a solved single query by one wave (every lane): the
// chain is followed six levels per round of loads (parent, grandparent and
// the ancestors at 3..6), lane l takes the row of the l-th node, and the
// rows go out together; cost and length are summed in walk order, so they
// have the bits of walk_path's. Paths of more than 64 rows take walk_path.
__device__ void walk_path_wave(const QueryDesc& qd, QueryState& st, const Pools& P, float* __restrict__ out,
                               int32_t goal)
{
    const uint32_t lane    = lane_id();
    const uint32_t cap     = qd.cap;
    const int D            = qd.D;
    const uint32_t maxRows = qd.cfg.max_iter + 2;
    const uint32_t lim     = maxRows < 64u ? maxRows : 64u;
    const int32_t* par     = P.parent + qd.nodeOff;
    const int32_t* gpar    = P.grand + qd.nodeOff;
    const int4* anc        = P.anc + qd.nodeOff;
    int32_t mine = lane == 0 ? goal : -1;  // this lane's node
    int32_t f    = goal;                    // the last node found
    uint32_t cnt = 1;                       // nodes found
    bool root    = false;                   // the chain ended at the root
    while(!root && cnt < lim)
        {
            const int32_t p = par[f], g = gpar[f];
            const int4 a    = anc[f];
            const int32_t up[6] = {p, g, a.x, a.y, a.z, a.w};
#pragma unroll
            for(int j = 0; j < 6; ++j)
                if(!root && cnt < lim)
                    {
                        if(up[j] < 0 || (uint32_t)up[j] >= cap)  // -1: past the root
                            root = true;
                        else
                            {
                                if(lane == cnt) mine = up[j];
                                f = up[j];
                                ++cnt;
                            }
                    }
        }
    // stopped at the plan's bound of rows: the last row's parent decides
    // whether its duration counts
    if(!root && cnt >= maxRows) root = par[f] == -1;
    if(!root && cnt < maxRows)
        {
            // longer than one wave's rows
            if(lane == 0) walk_path(qd, st, P, out, goal);
            return;
        }
    const float* __restrict__ T = P.treeS + qd.sampleOff;
    float r[KGMT_MAX_SAMPLE] = {};
    float dn = 0.0f;
    if(lane < cnt)
        {
#pragma unroll
            for(int k = 0; k < KGMT_MAX_SAMPLE; ++k)
                if(k < D) r[k] = T[(size_t)k * cap + mine];
            dn = T[(size_t)(D - 1) * cap + mine];
#pragma unroll
            for(int k = 0; k < KGMT_MAX_SAMPLE; ++k)
                if(k < D) out[lane * D + k] = r[k];
        }
    // segment length from the row before (lane l - 1) to this one
    const float qx = __shfl_up(r[0], 1, 64), qy = __shfl_up(r[1], 1, 64), qz = __shfl_up(r[2], 1, 64);
    const float e  = lane > 0 && lane < cnt ? pos_dist(qd, qx, qy, qz, r[0], r[1], r[2]) : 0.0f;
    // a node's edge duration counts when it has a parent: every node but the
    // root
    float pc = 0.0f, pl = 0.0f;
    for(uint32_t l = 0; l < cnt; ++l)
        {
            const float dl = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dn), l));
            if(!(root && l == cnt - 1)) pc = pc + dl;
            if(l > 0) pl = pl + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e), l));
        }
    if(lane == 0)
        {
            st.pathRows = cnt;
            st.pathCost = pc;
            st.pathLen  = pl;
        }
}

// How a single-query plan ended.
struct Outcome
{
    uint32_t success, iterations, treeSize;
    int32_t goal;
    uint64_t expanded, flagged;
};

// End of a single-query plan (one thread): the goal path rows and the result
// into host-mapped memory, then the completion word, so the host can return
// without waiting for the launches it queued past the last iteration (they
// find the query done and exit). Without a host result buffer only the
// completion word is set.
__device__ void publish_single(const QueryDesc& qd, QueryState& st, const Pools& P, const WorkLists& W,
                               const Outcome& r)
{
    if(!W.result)
        {
            __hip_atomic_store(W.allDone, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            return;
        }
    if(r.success && st.pathRows == 0) walk_path(qd, st, P, W.hostPath, r.goal);
    HostResult* h = W.result;
    h->success    = r.success;
    h->iterations = r.iterations;
    h->treeSize   = r.treeSize;
    h->pathRows   = r.success ? st.pathRows : 0u;
    h->goalIdx    = r.success ? r.goal : -1;
    h->pathCost   = r.success ? st.pathCost : 0.0f;
    h->pathLen    = r.success ? st.pathLen : 0.0f;
    h->expanded   = r.expanded;
    h->flagged    = r.flagged;
    __threadfence_system();
    __hip_atomic_store(W.allDone, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---------------------------------------------------------------------------
// update: the rest of an iteration in one launch: region scores, tree
// append, goal test, re-activation and the next expansion list
// (Graph.cu:207-358, KGMT.cu:305-403)
// ---------------------------------------------------------------------------
// A query's jobs, in the order of their entries in the next expansion list:
// its old-node chunks (one entry per re-activated node, ascending), then its
// child chunks (each appended node repeated 15 or 1 times). A job finds the
// entries before it by a decoupled look-back (Merrill and Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016) over one
// 64-bit word per job (two sets, by iteration parity; each launch empties
// the other set for the next): it publishes its own count and smallest goal
// node (aggregate) as soon as it has them, reads the words before it until
// one holds a prefix that includes its job (inclusive), and then publishes
// its own inclusive word. A job waits only for jobs held by workgroups that
// are running, and those publish their aggregates without waiting, so every
// wait ends: with no more jobs than workgroups, workgroup j does job j (the
// dispatcher starts the workgroups of an XCD in index order, so a workgroup
// whose predecessor has not started leaves that predecessor's XCD slots to
// lower-index workgroups or to other kernels, all of which finish); with
// more, workgroups take jobs from a counter in order. The wait is bounded all
// the same (kLookBackPolls); a job that runs out sets the error word, which
// the host reports, and goes on with what it has.
// Every job derives the iteration's totals it needs itself rather than
// waiting for a launch that computes them once: the flagged children of all
// chunks (and of the chunks before its own) from the chunk counts, and, for
// a re-activation job, the region scores (their sum is one fixed pairwise
// tree, so every job gets the same bits). Job 0 stores the scores and the
// subregion snapshot that the next propagate reads.
// update workgroups: kUpd threads, each taking kPer consecutive slots of a
// job's chunk (one slot: one block scan covers the chunk, and a thread's
// flag byte is one load), and all of them the region scores
constexpr uint32_t kUpd = 4 * kBlock;
constexpr uint32_t kPer = kChunk / kUpd;
static_assert(kPer == 1, "one slot per thread");
constexpr uint32_t kLbNone        = (1u << 27) - 1;  // no goal node (node indices stay below 2^26)
constexpr uint64_t kLbAggregate   = 1ull << 62;
constexpr uint64_t kLbInclusive   = 2ull << 62;
constexpr uint32_t kLookBackPolls = 1u << 20;

// LDS of an update workgroup
struct UpdateLds
{
    union
    {
        float raw[kMaxRegions];        // raw region scores (-1: no valid child) of a re-activation job
        uint32_t stage[kMaxRegions];   // then the job's expansion-list entries
    };
    float part[kUpd];                   // the score tree's partial sums
    uint32_t sbuf[3 * (kUpd / 64)];     // block scans
    uint32_t job, before, goalBefore, nAcc;
    float sum;
};

// word: flag (bits 62-63), smallest goal node (32-58), entries (0-31)
__device__ __forceinline__ uint64_t lb_word(uint64_t flag, uint32_t entries, uint32_t goal)
{
    return flag | ((uint64_t)goal << 32) | entries;
}

__device__ __forceinline__ uint32_t wave_min_u(uint32_t v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1)
        {
            const uint32_t t = (uint32_t)__shfl_xor((int)v, o, 64);
            v                = t < v ? t : v;
        }
    return v;
}

// Entries and smallest goal node of the jobs before job li (words: the
// query's job words); every lane of one wave. A round reads the 256 words
// before `base` (lane l, slot k: word base - 1 - l - 64 k), all loads in
// flight together, so the jobs of one launch (every aggregate published at
// about the same time) find their prefix in one or two round trips. False if
// a word did not appear within the poll bound (the sums then cover the words
// seen).
__device__ bool look_back(const uint64_t* words, uint32_t li, uint32_t& entries, uint32_t& goal)
{
    constexpr int KL   = 4;
    const int32_t lane = (int32_t)lane_id();
    uint32_t c = 0, g = kLbNone, polls = 0;
    int32_t base = (int32_t)li;
    while(true)
        {
            uint64_t w[KL];
#pragma unroll
            for(int k = 0; k < KL; ++k)
                {
                    const int32_t idx = base - 1 - lane - 64 * k;
                    // before the first job: an inclusive word of nothing
                    w[k] = idx >= 0 ? __hip_atomic_load(words + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : lb_word(kLbInclusive, 0u, kLbNone);
                }
            // per slot, the lanes up to and including the nearest inclusive
            // word (none past the slot that holds it)
            uint64_t need[KL];
            bool found = false, wait = false;
#pragma unroll
            for(int k = 0; k < KL; ++k)
                {
                    const uint32_t flag = (uint32_t)(w[k] >> 62);
                    const uint64_t incl = __ballot(flag == 2u);
                    const uint64_t none = __ballot(flag == 0u);
                    need[k] = found ? 0ull : incl ? ((incl & (~incl + 1ull)) << 1) - 1ull : ~0ull;
                    wait    = wait || (none & need[k]) != 0ull;
                    found   = found || incl != 0ull;
                }
            if(wait)
                {
                    if(++polls >= kLookBackPolls)
                        {
                            entries = c, goal = g;
                            return false;
                        }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
            uint32_t wc = 0u, wg = kLbNone;
#pragma unroll
            for(int k = 0; k < KL; ++k)
                if((need[k] >> lane) & 1ull)
                    {
                        wc += (uint32_t)w[k];
                        const uint32_t gk = (uint32_t)(w[k] >> 32) & kLbNone;
                        wg                = gk < wg ? gk : wg;
                    }
            c += (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(wc), 63);
            const uint32_t gm = wave_min_u(wg);
            g = gm < g ? gm : g;
            if(found)
                {
                    entries = c, goal = g;
                    return true;
                }
            base -= 64 * KL;
        }
}

// Exclusive scan of v and minimum of m over a block of NT threads, one
// barrier pair; sbuf: 2 * NT/64 words. Call uniformly.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan_min(uint32_t v, uint32_t m, uint32_t* sbuf, uint32_t& total,
                                                        uint32_t& mmin)
{
    constexpr int NW = NT / 64;
    static_assert(NW >= 2 && NW <= 16, "block of 128 to 1024 threads");
    const uint32_t incl = wave_incl_scan(v);
    const uint32_t lane = lane_id();
    const int wave      = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
    if(__ballot(m != kLbNone)) m = wave_min_u(m);
    if(lane == 63) sbuf[wave] = incl, sbuf[NW + wave] = m;
    __syncthreads();
    const uint32_t w  = row_incl_scan(lane < (uint32_t)NW ? sbuf[lane] : 0u);
    const uint32_t pw = wave == 0 ? 0u : (uint32_t)__builtin_amdgcn_readlane((int)w, wave - 1);
    total             = (uint32_t)__builtin_amdgcn_readlane((int)w, NW - 1);
    uint32_t mm       = kLbNone;
#pragma unroll
    for(int i = 0; i < NW; ++i) mm = sbuf[NW + i] < mm ? sbuf[NW + i] : mm;
    mmin = mm;
    __syncthreads();
    return pw + incl - v;
}

// Region scores for a re-activation job (Graph.cu:239-257 per region,
// :297-304 and :349-357 for the sum): raw scores into L.raw (-1 for a region
// without valid children), their sum by the pairwise tree of
// score_regions_sum, here over all kMaxRegions leaves (the levels above the
// query's own padded size add zeros, which leave every partial sum's bits
// unchanged), into L.sum; with the flagged children of the iteration
// (nAcc, acc: this thread's share of the chunk counts) into L.nAcc.
// KB regions per thread and round. Barriers inside: call uniformly.
template <int KB>
__device__ void job_scores(const QueryDesc& qd, const Pools& P, uint32_t acc, UpdateLds& L)
{
    const int R1             = qd.R1;
    const uint32_t Pn        = (uint32_t)qd.P;
    const int t              = (int)threadIdx.x;
    const int32_t* total     = P.total + qd.regOff;
    const int32_t* validc    = P.validc + qd.regOff;
    const uint32_t* act      = P.actNext + qd.subOff;
    // KB regions per thread and round: every load of a round is issued
    // before any is used (indices clamped into the query's regions, so no
    // load sits behind a branch)
    const bool oneWord   = Pn <= 32u && 32u % Pn == 0u;
    for(int k0 = t; k0 < R1; k0 += KB * (int)kUpd)
        {
            int32_t vc[KB], tc[KB];
            uint32_t aw[KB];
#pragma unroll
            for(int k = 0; k < KB; ++k)
                {
                    const int r  = k0 + k * (int)kUpd;
                    const int rr = r < R1 ? r : R1 - 1;
                    vc[k]        = validc[rr];
                    tc[k]        = total[rr];
                    aw[k]        = oneWord ? act[((uint32_t)rr * Pn) >> 5] : 0u;
                }
#pragma unroll
            for(int k = 0; k < KB; ++k)
                {
                    const int r = k0 + k * (int)kUpd;
                    if(r < R1 && vc[k] > 0)
                        {
                            const uint32_t b   = (uint32_t)r * Pn;
                            const uint32_t m   = Pn == 32u ? 0xffffffffu : ((1u << (Pn & 31u)) - 1u) << (b & 31u);
                            const uint32_t cnt = oneWord ? (uint32_t)__popc(aw[k] & m) : count_bits(act, b, Pn);
                            L.raw[r]           = region_raw_score(qd, vc[k], tc[k], cnt);
                        }
                    else if(r < R1)
                        L.raw[r] = -1.0f;
                }
        }
    // the flagged children of the iteration, per wave (used after the tree)
    const uint32_t lane    = lane_id();
    const int wave         = __builtin_amdgcn_readfirstlane(t / 64);
    const uint32_t waveAcc = (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(acc), 63);
    if(lane == 0) L.sbuf[wave] = waveAcc;
    __syncthreads();
    // the thread's leaves t + k kUpd; the levels with stride >= kUpd add
    // the thread's own registers
    constexpr int KR = kMaxRegions / (int)kUpd;
    float sv[KR];
#pragma unroll
    for(int k = 0; k < KR; ++k)
        {
            const int r = t + k * (int)kUpd;
            const float x = r < R1 ? L.raw[r] : 0.0f;
            sv[k]         = x < 0.0f ? 0.0f : x;
        }
#pragma unroll
    for(int m = KR / 2; m >= 1; m >>= 1)
        {
#pragma unroll
            for(int k = 0; k < m; ++k) sv[k] = sv[k] + sv[k + m];
        }
    L.part[t] = sv[0];
    __syncthreads();
    if(t < 64)
        {
            // lane l holds partial sums l + 64 j; then the last six levels
            // pair lanes within the wave (as score_regions_sum)
            constexpr int NJ = (int)kUpd / 64;
            float x[NJ];
#pragma unroll
            for(int j = 0; j < NJ; ++j) x[j] = L.part[t + 64 * j];
#pragma unroll
            for(int m = NJ / 2; m >= 1; m >>= 1)
                {
#pragma unroll
                    for(int j = 0; j < m; ++j) x[j] = x[j] + x[j + m];
                }
            float v = x[0];
            v       = v + lane_down<32>(v);
            v       = v + lane_down<16>(v);
            v       = v + lane_down<8>(v);
            v       = v + lane_down<4>(v);
            v       = v + lane_down<2>(v);
            v       = v + lane_down<1>(v);
            if(t == 0)
                {
                    L.sum       = v;
                    uint32_t na = 0;
#pragma unroll
                    for(int w = 0; w < NJ; ++w) na += L.sbuf[w];
                    L.nAcc = na;
                }
        }
    __syncthreads();
}

// score of a region from its raw score and the sum (-1: no valid child)
__device__ __forceinline__ float region_score(const QueryDesc& qd, float raw, float sum)
{
    return raw < 0.0f ? 1.0f : qd.cfg.epsilon + raw / sum;
}

// Old-node chunk o: the nodes not in the frontier whose draw u is at most
// their region's score plus fA join it again (KGMT.cu:364-376); scores from
// L.raw and sum (own) or from P.score; entries: one per re-activated node
// (node[k] = its index, rep[k] = 1)
__device__ __forceinline__ void react_job(const QueryDesc& qd, const QueryState& st, uint32_t o, const Pools& P,
                                          uint32_t fw, const int32_t rr, const float (&u)[kPer], bool own,
                                          const float* raw, float sum, float fA, uint32_t (&node)[kPer],
                                          uint32_t (&rep)[kPer])
{
    const uint32_t tree = st.treeSize;
    const size_t nodeO  = qd.nodeOff;
    const uint32_t i0   = o * kChunk + kPer * threadIdx.x;
#pragma unroll
    for(uint32_t k = 0; k < kPer; ++k) node[k] = i0 + k, rep[k] = 0u;
    if(i0 >= tree) return;
    const int32_t r[kPer] = {rr};
    float sc[kPer];
#pragma unroll
    for(uint32_t k = 0; k < kPer; ++k)
        {
            const bool cand = i0 + k < tree && ((fw >> (8 * k)) & 0xffu) == 0;
            sc[k] = !cand ? 0.0f : own ? region_score(qd, raw[r[k]], sum) : P.score[qd.regOff + r[k]];
        }
    uint32_t rw = 0;
#pragma unroll
    for(uint32_t k = 0; k < kPer; ++k)
        {
            const bool cand = i0 + k < tree && ((fw >> (8 * k)) & 0xffu) == 0;
            if(cand && u[k] <= sc[k] + fA)
                {
                    // byte store: the new nodes' flags past the tree are
                    // written concurrently by the child chunk jobs
                    P.frontier[nodeO + i0 + k] = 1;
                    rw |= 1u << (8 * k);
                    rep[k] = 1u;
                }
        }
    P.react[nodeO + i0] = (uint8_t)rw;
}

// Child chunk c: the flagged children whose rank (flagged children before
// them) is below `next` (nAcc clamped to the free capacity) are appended at
// tree + rank with their distance to the goal (KGMT.cu:305-345); entries:
// each appended node repeated 15 or 1 times by its region's valid count
// (KGMT.cu:336-339); goal: this thread's smallest appended index within
// goal_thresh (kLbNone if none). pAll / pBefore: this thread's share of the
// chunk counts of all chunks / of the chunks before c.
__device__ __forceinline__ void append_job(const QueryDesc& qd, const QueryState& st, uint32_t c, const Pools& P,
                                           uint32_t pAll, uint32_t pBefore, uint32_t* sbuf, uint32_t (&node)[kPer],
                                           uint32_t (&rep)[kPer], uint32_t& goal, uint32_t& nAcc, int32_t& apar)
{
    const uint32_t cap = qd.cap, nC = st.nChildren, tree = st.treeSize;
    const size_t nodeO    = qd.nodeOff;
    float* __restrict__ T = P.treeS + qd.sampleOff;
    const int D           = qd.D;
    const uint32_t u0     = c * kChunk + kPer * threadIdx.x;
    uint32_t fl           = 0;
    int32_t r1[kPer]      = {0};
    int32_t vcount[kPer]  = {0};
    constexpr int kRowMax = row_words(KGMT_MAX_SAMPLE);
    const int RW          = row_words(D);
    float v[kPer][kRowMax];  // rows as stored: parent and grandparent bits, then the sample
    if(u0 < nC)
        {
            // flags and regions together
            const uint32_t w = P.fnext[nodeO + u0];
            r1[0]            = P.r1child[nodeO + u0];
#pragma unroll
            for(uint32_t k = 0; k < kPer; ++k)
                if(u0 + k < nC && ((w >> (8 * k)) & 0xffu)) fl |= 1u << k;
        }
    // Before the rank scan: the regions' valid counts for the repeat count
    // and the rows of the flagged children (appended unless the tree fills
    // up). A flagged child's row sits at its 32-child segment's start plus
    // the flagged children before it in the segment (the 32 threads of half
    // a wave).
    const uint32_t nf = (uint32_t)__popc(fl);
    uint32_t slot     = (u0 & ~31u) + half_incl_scan(nf) - nf;
#pragma unroll
    for(uint32_t k = 0; k < kPer; ++k)
        if((fl >> k) & 1u)
            {
                vcount[k]         = P.validc[qd.regOff + r1[k]];
                const float4* row = reinterpret_cast<const float4*>(P.childR + qd.rowOffC + (size_t)slot * RW);
#pragma unroll
                for(int j = 0; j < kRowMax / 4; ++j)
                    if(4 * j < RW)
                        {
                            const float4 x = row[j];
                            v[k][4 * j] = x.x, v[k][4 * j + 1] = x.y, v[k][4 * j + 2] = x.z, v[k][4 * j + 3] = x.w;
                        }
                ++slot;
            }
    // one scan: ranks inside the chunk, flagged children of all chunks and
    // of the chunks before this one
    const uint32_t in3[3] = {nf, pAll, pBefore};
    uint32_t ex3[3], tot3[3];
    block_excl_scan_n<kUpd, 3>(in3, ex3, tot3, sbuf);
    nAcc                = tot3[1];
    const uint32_t next = nAcc < cap - tree ? nAcc : cap - tree;
    uint32_t rank       = tot3[2] + ex3[0];
    goal                = kLbNone;
    apar                = -1;
    uint32_t t          = tree + rank;
#pragma unroll
    for(uint32_t k = 0; k < kPer; ++k)
        {
            rep[k]  = 0u;
            node[k] = t;
            if((fl >> k) & 1u)
                {
                    if(rank < next)
                        {
#pragma unroll
                            for(int d = 0; d < KGMT_MAX_SAMPLE; ++d)
                                if(d < D) T[(size_t)d * cap + t] = v[k][2 + d];
                            P.parent[nodeO + t]   = __float_as_int(v[k][0]);
                            P.grand[nodeO + t]    = __float_as_int(v[k][1]);
                            apar                  = __float_as_int(v[k][0]);
                            const float dg        = goal_dist(qd, v[k][2], v[k][3], v[k][4]);
                            P.cost[nodeO + t]     = dg;
                            P.frontier[nodeO + t] = 1;
                            P.r1node[nodeO + t]   = r1[k];
                            rep[k] = vcount[k] < qd.cfg.sparse_valid_below ? qd.cfg.repeat_sparse : qd.cfg.repeat_dense;
                            if(dg < qd.cfg.goal_thresh && goal == kLbNone) goal = t;
                        }
                    ++rank;
                    ++t;
                }
        }
}

// One job of query q (li: its index among the query's jobs). own: the
// scores come from the region counts (the planner); otherwise from P.score
// (the stage entry, which is given them).
__device__ void update_job(const QueryDesc* __restrict__ qds, QueryState* __restrict__ sts, const Pools& P,
                           const WorkLists& W, uint32_t parity, uint32_t q, uint32_t li, bool own, UpdateLds& L)
{
    QueryState& st      = sts[q];
    const QueryDesc& qd = qds[q];
    const uint32_t nq   = W.nq;
    const uint32_t tree = st.treeSize, nC = st.nChildren, cap = qd.cap;
    const uint32_t nOC  = (tree + kChunk - 1) / kChunk;
    const uint32_t nCC  = (nC + kChunk - 1) / kChunk;
    const uint32_t nJ   = nOC + nCC;
    const uint32_t* pa  = P.partA + qd.chunkOff;  // flagged children per child chunk
    const size_t wpq    = 2 * ((size_t)qd.nChunks + 1);  // words of one parity
    uint64_t* words     = P.look + 4 * (size_t)qd.chunkOff + parity * wpq;
    uint64_t* wordsNext = P.look + 4 * (size_t)qd.chunkOff + (parity ^ 1u) * wpq;
    const uint32_t c    = li - nOC;  // child chunk of a child job
    // this thread's share of the chunk counts: the first without a branch
    // (the index clamped into the query's chunks), the rest past kUpd chunks
    const uint32_t i0c = threadIdx.x < qd.nChunks ? threadIdx.x : qd.nChunks;
    const uint32_t x0  = pa[i0c];
    uint32_t pAll      = threadIdx.x < nCC ? x0 : 0u;
    uint32_t pBefore   = threadIdx.x < nCC && threadIdx.x < c ? x0 : 0u;
    for(uint32_t i = threadIdx.x + kUpd; i < nCC; i += kUpd)
        {
            const uint32_t x = pa[i];
            pAll += x;
            pBefore += i < c ? x : 0u;
        }
    uint32_t node[kPer], rep[kPer], goal = kLbNone, nAcc;
    int32_t apar = -1;  // parent of the node this thread appends
    if(li < nOC)
        {
            // the job's nodes first, then the totals (their loads overlap)
            const uint32_t i0 = li * kChunk + kPer * threadIdx.x;
            uint32_t fw       = 0;
            int32_t rr        = 0;
            if(i0 < tree)
                {
                    fw = P.frontier[qd.nodeOff + i0];
                    rr = P.r1node[qd.nodeOff + i0];
                }
            // the draws need neither loads nor scores: computed while they come
            float u[kPer];
#pragma unroll
            for(uint32_t k = 0; k < kPer; ++k) u[k] = draw_react(qd.cfg, qd.cfg.query_id, st.itr, i0 + k);
            // one round of loads: 4 regions per thread up to 4096 regions, 8 above
            if(own && qd.R1 <= 4 * (int)kUpd)
                job_scores<4>(qd, P, pAll, L);
            else if(own)
                job_scores<8>(qd, P, pAll, L);
            else
                {
                    uint32_t tot;
                    block_excl_scan<kUpd>(pAll, L.sbuf, tot);
                    if(threadIdx.x == 0) L.nAcc = tot, L.sum = 0.0f;
                    __syncthreads();
                }
            nAcc                = L.nAcc;
            const uint32_t next = nAcc < cap - tree ? nAcc : cap - tree;
            react_job(qd, st, li, P, fw, rr, u, own, L.raw, L.sum, faccept(qd, st.itr, tree, next), node, rep);
        }
    else
        append_job(qd, st, c, P, pAll, pBefore, L.sbuf, node, rep, goal, nAcc, apar);
    uint32_t total, jobGoal;
    const uint32_t ex = block_excl_scan_min<kUpd>(rep[0], goal, L.sbuf, total, jobGoal);
    if(threadIdx.x == 0)
        {
            __hip_atomic_store(words + li, lb_word(li == 0 ? kLbInclusive : kLbAggregate, total, jobGoal),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            L.before = 0u, L.goalBefore = kLbNone;
        }
    // the aggregate goes out before the look-back's loads
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    // the appended node's ancestors at distances 3..6 are its parent's at
    // 2..5, for the path walk (no job of this launch reads them): written
    // while the look-back runs, by wave 0 after it
    auto link = [&]() {
        if(apar >= 0 && (uint32_t)apar < cap)
            {
                const int32_t g = P.grand[qd.nodeOff + apar];
                const int4 a    = P.anc[qd.nodeOff + apar];
                P.anc[qd.nodeOff + node[0]] = make_int4(g, a.x, a.y, a.z);
            }
    };
    if(threadIdx.x >= 64) link();
    // the job's entries into LDS while the look-back runs (written out below
    // as whole rows; a job with more entries than the stage holds, and job 0,
    // whose region scores stay in L.raw, write them directly)
    const bool staged = total <= (uint32_t)kMaxRegions && !(li == 0 && own);
    if(staged && threadIdx.x >= 64)
        {
            uint32_t p = ex;
#pragma unroll
            for(uint32_t k = 0; k < kPer; ++k)
                {
                    for(uint32_t m = 0; m < rep[k]; ++m) L.stage[p + m] = node[k];
                    p += rep[k];
                }
        }
    if(li > 0 && threadIdx.x < 64)
        {
            uint32_t pc, pg;
            const bool ok = look_back(words, li, pc, pg);
            if(threadIdx.x == 0)
                {
                    if(!ok) __hip_atomic_store(W.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    __hip_atomic_store(words + li, lb_word(kLbInclusive, pc + total, jobGoal < pg ? jobGoal : pg),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    L.before = pc, L.goalBefore = pg;
                }
        }
    if(threadIdx.x < 64) link();
    if(staged && threadIdx.x < 64)
        {
            uint32_t p = ex;
#pragma unroll
            for(uint32_t k = 0; k < kPer; ++k)
                {
                    for(uint32_t m = 0; m < rep[k]; ++m) L.stage[p + m] = node[k];
                    p += rep[k];
                }
        }
    // (work that no job of this launch reads) the other parity's words start
    // empty for the next launch; job 0 stores the scores and the subregion
    // snapshot for the next propagate
    for(size_t i = (size_t)li * kUpd + threadIdx.x; i < wpq; i += (size_t)nJ * kUpd) wordsNext[i] = 0ull;
    if(li == 0 && own)
        {
            for(int r = threadIdx.x; r < qd.R1; r += kUpd) P.score[qd.regOff + r] = region_score(qd, L.raw[r], L.sum);
            for(int w = threadIdx.x; w < qd.R2words; w += kUpd) P.actPrev[qd.subOff + w] = P.actNext[qd.subOff + w];
        }
    __syncthreads();
    const uint32_t before = L.before, goalBefore = L.goalBefore;
    // the job's entries of the expansion list at before.., its first
    // capacity entries kept
    uint32_t* out = P.repIdx + qd.nodeOff;
    if(staged)
        {
            for(uint32_t e = threadIdx.x; e < total; e += kUpd)
                if(before + e < cap) out[before + e] = L.stage[e];
        }
    else
        {
            uint32_t pos = before + ex;
#pragma unroll
            for(uint32_t k = 0; k < kPer; ++k)
                {
                    for(uint32_t m = 0; m < rep[k]; ++m)
                        if(pos + m < cap) out[pos + m] = node[k];
                    pos += rep[k];
                }
        }
    const uint32_t next = nAcc < cap - tree ? nAcc : cap - tree;
    const bool last     = li == nJ - 1;
    // the first goal of a single query's iteration: wave 0 walks its path
    // into host memory before its first lane publishes the result
    if(nq == 1 && jobGoal != kLbNone && goalBefore == kLbNone && W.result != nullptr && threadIdx.x < 64 &&
       st.pathRows == 0u)
        walk_path_wave(qd, st, P, W.hostPath, (int32_t)jobGoal);
    // The last job's look-back has seen a word of every other job, each
    // published after that job read the chunk counts: they start at zero
    // again for the next propagate.
    if(last)
        for(uint32_t i = threadIdx.x; i < nCC; i += kUpd) P.partA[qd.chunkOff + i] = 0u;
    if(threadIdx.x == 0)
        {
            const uint32_t goalAll = jobGoal < goalBefore ? jobGoal : goalBefore;
            // a single query's job holding the smallest goal node publishes
            // the solved plan: the goal row was written by this workgroup,
            // the rest of its path by earlier launches
            if(nq == 1 && jobGoal != kLbNone && goalBefore == kLbNone)
                {
                    const Outcome r{1u, st.itr, tree + next, (int32_t)jobGoal, st.expanded + nC, st.flagged + nAcc};
                    publish_single(qd, st, P, W, r);
                }
            // the last job knows the list size and the goal: it closes the
            // iteration of a single query (the next-state fields and work
            // lists, read from the next kernel on), k_finish those of a batch
            if(last)
                {
                    const uint32_t R = before + total;
                    st.R             = R;
                    st.goalIdx       = goalAll == kLbNone ? kNoGoal : (int32_t)goalAll;
                    st.nAcc          = nAcc;
                    st.next          = next;
                    st.fAccept       = faccept(qd, st.itr, tree, next);
                    if(nq == 1)
                        {
                            finish_query(qd, st, R);
                            single_worklists(st, W, parity ^ 1u);
                            count_closed(W);
                            if(st.doneN && !st.successN)
                                {
                                    const Outcome r{0u, st.finishItrN, st.treeSizeN, -1, st.expandedN, st.flaggedN};
                                    publish_single(qd, st, P, W, r);
                                }
                        }
                }
        }
}

__global__ __launch_bounds__(kUpd) void k_update(const QueryDesc* __restrict__ qds, QueryState* __restrict__ sts,
                                                   Pools P, WorkLists W, uint32_t parity, uint32_t own)
{
    __shared__ UpdateLds L;
    const uint32_t nq       = W.nq;
    const uint32_t* jobList = W.jobs(parity);
    // one query: its jobs 0 .. nJ-1 straight from its state (the list holds
    // the same), without waiting for the list first
    const QueryState& st0 = sts[0];
    const uint32_t nJobs  = nq > 1 ? jobList[nq]
                            : st0.done ? 0u
                                       : (st0.treeSize + kChunk - 1) / kChunk + (st0.nChildren + kChunk - 1) / kChunk;
    // with no more jobs than workgroups, workgroup j does job j; with more,
    // every job comes from the counter (the other parity's counter starts at
    // zero for the next launch)
    if(blockIdx.x >= nJobs) return;
    if(blockIdx.x == 0 && threadIdx.x == 0) W.claim[parity ^ 1u] = 0u;
    const bool more = nJobs > gridDim.x;
    while(true)
        {
            uint32_t j = blockIdx.x;
            if(more)
                {
                    if(threadIdx.x == 0)
                        L.job = __hip_atomic_fetch_add(W.claim + parity, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __syncthreads();
                    j = L.job;
                    if(j >= nJobs) return;
                }
            const uint32_t q = nq > 1 ? locate(jobList, nq, j) : 0u;
            // (a stale list entry of a finished query has nothing to do)
            if(!sts[q].done) update_job(qds, sts, P, W, parity, q, nq > 1 ? j - jobList[q] : j, own != 0, L);
            if(!more) return;
            __syncthreads();  // L is rewritten by the next job
        }
}

// ---------------------------------------------------------------------------
// finish: one workgroup closes the iteration of every query and builds the
// next iteration's work lists
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_finish(const QueryDesc* __restrict__ qds, QueryState* __restrict__ sts,
                                                   WorkLists W, uint32_t parity)
{
    __shared__ uint32_t sbuf[kBlock / 64 + 1];
    for(uint32_t q = threadIdx.x; q < W.nq; q += kBlock)
        if(!sts[q].done) finish_query(qds[q], sts[q], sts[q].R);
    __syncthreads();
    build_worklists<kBlock>(sts, W, parity ^ 1u, sbuf);
    if(threadIdx.x == 0) count_closed(W);
}

// ---------------------------------------------------------------------------
// paths: goal -> root rows of every solved query, exact depth (Q13); path
// cost = sum of edge durations, length = sum of segment lengths between the
// node positions
// ---------------------------------------------------------------------------
// goal -> root rows of a solved query (pathRows etc. of st); two levels per
// round of loads: a and its parent b are known, their rows, b's parent c and
// grandparent d are loaded together
__device__ void walk_path(const QueryDesc& qd, QueryState& st, const Pools& P, float* __restrict__ out, int32_t goal)
{
    const uint32_t cap            = qd.cap;
    const int D                   = qd.D;
    const float* __restrict__ T   = P.treeS + qd.sampleOff;
    const float* __restrict__ dur = T + (size_t)(D - 1) * cap;  // edge durations
    const int32_t* par            = P.parent + qd.nodeOff;
    const int32_t* gpar           = P.grand + qd.nodeOff;
    uint32_t rows                 = 0;
    float pc = 0.0f, pl = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
    const uint32_t maxRows = qd.cfg.max_iter + 2;
    // one node's row, in walk order (its parent pn decides whether its edge
    // duration dn counts); indices into r are compile-time constants
    auto emit = [&](const float* r, float dn, int32_t pn) {
#pragma unroll
        for(int k = 0; k < KGMT_MAX_SAMPLE; ++k)
            if(k < D) out[rows * D + k] = r[k];
        if(pn != -1) pc = pc + dn;
        if(rows > 0) pl = pl + pos_dist(qd, px, py, pz, r[0], r[1], r[2]);
        px = r[0], py = r[1], pz = r[2];
        ++rows;
    };
    int32_t a = goal;
    int32_t b = par[a];
    while(a != -1 && rows < maxRows)
        {
            float ra[KGMT_MAX_SAMPLE], rb[KGMT_MAX_SAMPLE];
            float da = dur[a], db = 0.0f;
            int32_t c = -1, d = -1;
#pragma unroll
            for(int k = 0; k < KGMT_MAX_SAMPLE; ++k)
                if(k < D) ra[k] = T[(size_t)k * cap + a];
            if(b != -1)
                {
#pragma unroll
                    for(int k = 0; k < KGMT_MAX_SAMPLE; ++k)
                        if(k < D) rb[k] = T[(size_t)k * cap + b];
                    db = dur[b];
                    c  = par[b];
                    d  = gpar[b];
                }
            emit(ra, da, b);
            if(b == -1 || rows >= maxRows) break;
            emit(rb, db, c);
            a = c;
            b = d;
        }
    st.pathRows = rows;
    st.pathCost = pc;
    st.pathLen  = pl;
}

__global__ __launch_bounds__(64) void k_paths(const QueryDesc* __restrict__ qds, QueryState* __restrict__ sts,
                                              Pools P, uint32_t nq)
{
    const uint32_t q = blockIdx.x * 64 + threadIdx.x;
    if(q >= nq) return;
    QueryState& st = sts[q];
    commit_state(st);  // the closed last iteration
    if(!st.success || st.pathRows) return;
    walk_path(qds[q], st, P, P.path + qds[q].pathOff, st.goalIdx);
}

// ---------------------------------------------------------------------------
// scene build: one workgroup per scene. Boxes sorted along a Morton curve of
// their centres (10 bits per axis over the bounds of the finite centres; ties
// by input order), grouped by kGroup; each group's bounds are the min/max of
// its boxes per axis, or NaN on an axis where a member has a NaN face (such a
// group is never skipped); groups past the boxes hold lo = +inf, hi = -inf and
// are never near. Output: nGrp group (lo, hi) pairs, then nGrp * kGroup boxes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t morton10(uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t code = 0;
    for(int b = 9; b >= 0; --b) code = (code << 3) | (((x >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((z >> b) & 1u);
    return code;
}

struct SceneLds
{
    float box[kMaxObs][6];
    uint32_t key[kMaxObs];
    float red[6][kBlock / 64];
};

// Face k (0..2 min x y z, 3..5 max x y z) of box i of rows of 2 * W floats; a
// planar box (W = 2) spans z [-1, 1], which holds the planar positions' z = 0.
template <int W>
__device__ __forceinline__ float row_face(const float* __restrict__ rows, uint32_t i, int k)
{
    if constexpr(W == 3)
        return rows[6 * (size_t)i + k];
    else
        return k == 2 ? -1.0f : k == 5 ? 1.0f : rows[4 * (size_t)i + (k < 3 ? k : k - 1)];
}

template <int W>
__device__ void build_scene_wg(const float* __restrict__ rows, uint32_t n, uint32_t nGrp, float4* __restrict__ out,
                               SceneLds& L)
{
    // bounds of the finite centres
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(uint32_t i = threadIdx.x; i < n; i += kBlock)
        for(int d = 0; d < 3; ++d)
            {
                const float c = 0.5f * (row_face<W>(rows, i, d) + row_face<W>(rows, i, 3 + d));
                if(__builtin_isfinite(c)) lo[d] = c < lo[d] ? c : lo[d], hi[d] = c > hi[d] ? c : hi[d];
            }
    for(int d = 0; d < 3; ++d)
        {
            for(int o = 32; o > 0; o >>= 1)
                {
                    const float a = __shfl_xor(lo[d], o, 64), b = __shfl_xor(hi[d], o, 64);
                    lo[d] = a < lo[d] ? a : lo[d], hi[d] = b > hi[d] ? b : hi[d];
                }
            if(lane_id() == 0) L.red[d][threadIdx.x / 64] = lo[d], L.red[3 + d][threadIdx.x / 64] = hi[d];
        }
    __syncthreads();
    for(int d = 0; d < 3; ++d)
        {
            lo[d] = INFINITY, hi[d] = -INFINITY;
            for(int w = 0; w < kBlock / 64; ++w)
                {
                    const float a = L.red[d][w], b = L.red[3 + d][w];
                    lo[d] = a < lo[d] ? a : lo[d], hi[d] = b > hi[d] ? b : hi[d];
                }
        }
    // Morton keys
    for(uint32_t i = threadIdx.x; i < n; i += kBlock)
        {
            uint32_t q[3];
            for(int d = 0; d < 3; ++d)
                {
                    const float c = 0.5f * (row_face<W>(rows, i, d) + row_face<W>(rows, i, 3 + d));
                    float t       = hi[d] > lo[d] ? (c - lo[d]) / (hi[d] - lo[d]) : 0.0f;
                    if(!(t >= 0.0f)) t = 0.0f;
                    if(t > 1.0f) t = 1.0f;
                    q[d] = (uint32_t)(t * 1023.0f);
                }
            L.key[i] = morton10(q[0], q[1], q[2]);
        }
    __syncthreads();
    // rank = position in (key, index) order; box i goes to slot rank
    for(uint32_t i = threadIdx.x; i < n; i += kBlock)
        {
            const uint32_t ki = L.key[i];
            uint32_t rank     = 0;
            for(uint32_t j = 0; j < n; ++j)
                {
                    const uint32_t kj = L.key[j];
                    rank += (kj < ki || (kj == ki && j < i)) ? 1u : 0u;
                }
            for(int d = 0; d < 6; ++d) L.box[rank][d] = row_face<W>(rows, i, d);
        }
    __syncthreads();
    float4* box = out + 2 * nGrp;
    for(uint32_t k = threadIdx.x; k < nGrp * kGroup; k += kBlock)
        {
            if(k < n)
                {
                    box[2 * k]     = make_float4(L.box[k][0], L.box[k][1], L.box[k][2], 0.0f);
                    box[2 * k + 1] = make_float4(L.box[k][3], L.box[k][4], L.box[k][5], 0.0f);
                }
            else
                {
                    box[2 * k]     = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
                    box[2 * k + 1] = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
                }
        }
    for(uint32_t g = threadIdx.x; g < nGrp; g += kBlock)
        {
            float gl[3] = {INFINITY, INFINITY, INFINITY}, gh[3] = {-INFINITY, -INFINITY, -INFINITY};
            bool nanAxis[3] = {false, false, false};
            const uint32_t b1 = n < (g + 1) * kGroup ? n : (g + 1) * kGroup;
            for(uint32_t k = g * kGroup; k < b1; ++k)
                for(int d = 0; d < 3; ++d)
                    {
                        const float l = L.box[k][d], h = L.box[k][3 + d];
                        if(l != l || h != h) nanAxis[d] = true;
                        gl[d] = l < gl[d] ? l : gl[d];  // a NaN member leaves the bound as is
                        gh[d] = gh[d] < h ? h : gh[d];
                    }
            for(int d = 0; d < 3; ++d)
                if(nanAxis[d]) gl[d] = gh[d] = NAN;
            out[2 * g]     = make_float4(gl[0], gl[1], gl[2], 0.0f);
            out[2 * g + 1] = make_float4(gh[0], gh[1], gh[2], 0.0f);
        }
}

// ---------------------------------------------------------------------------
// init: one workgroup per query (KGMT.cu:50-81): the query's scene from the
// caller's obstacle rows (rows null: no obstacles), then the root, cleared
// counters and the first iteration's state; the last workgroup to arrive
// builds the first work lists.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_init(const QueryDesc* __restrict__ qds, QueryState* __restrict__ sts,
                                                 Pools P, WorkLists W, uint32_t nq, const float* __restrict__ rows)
{
    __shared__ SceneLds L;
    __shared__ uint32_t sbuf[kBlock / 64 + 1];
    __shared__ uint32_t sflag;
    const uint32_t q    = blockIdx.x;
    const QueryDesc& qd = qds[q];
    if(rows && qd.nObs)
        {
            if(qd.wdim == 2)
                build_scene_wg<2>(rows + 4 * (size_t)qd.rowOff, qd.nObs, qd.nGrp, P.obs + 2 * (size_t)qd.obsOff, L);
            else
                build_scene_wg<3>(rows + 6 * (size_t)qd.rowOff, qd.nObs, qd.nGrp, P.obs + 2 * (size_t)qd.obsOff, L);
        }
    clear_query(qd, P, threadIdx.x, kBlock);
    root_query(qd, sts[q], P);
    if(q == 0 && threadIdx.x < 2) W.claim[threadIdx.x] = 0u;
    if(q == 0 && threadIdx.x == 0) *W.closedDev = 0u;
    if(last_arrival(W.counter, nq, &sflag)) build_worklists<kBlock>(sts, W, 0, sbuf);
}

// k_init of a single query, its descriptor a kernel argument that the first
// thread stores for the kernels after it (no upload before the launch):
// workgroup 0 builds the scene, the root, the state and the first work
// lists (no hand-off between workgroups), workgroups 1.. clear the counters
__global__ __launch_bounds__(kBlock) void k_init_one(const QueryDesc qd, QueryDesc* __restrict__ dq,
                                                     QueryState* __restrict__ sts, Pools P, WorkLists W,
                                                     const float* __restrict__ rows)
{
    __shared__ SceneLds L;
    if(blockIdx.x > 0)
        {
            clear_query(qd, P, (blockIdx.x - 1) * kBlock + threadIdx.x, (gridDim.x - 1) * kBlock);
            return;
        }
    if(threadIdx.x == 0) *dq = qd;
    if(rows && qd.nObs)
        {
            if(qd.wdim == 2)
                build_scene_wg<2>(rows + 4 * (size_t)qd.rowOff, qd.nObs, qd.nGrp, P.obs + 2 * (size_t)qd.obsOff, L);
            else
                build_scene_wg<3>(rows + 6 * (size_t)qd.rowOff, qd.nObs, qd.nGrp, P.obs + 2 * (size_t)qd.obsOff, L);
        }
    root_query(qd, sts[0], P);
    if(threadIdx.x < 2) W.claim[threadIdx.x] = 0u;
    // the first thread wrote the state: it builds the first work lists
    if(threadIdx.x == 0)
        {
            *W.closedDev = 0u;
            single_worklists(sts[0], W, 0u);
        }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_build_scene(const float* __restrict__ rows, uint32_t n, uint32_t nGrp,
                                                        float4* __restrict__ out)
{
    __shared__ SceneLds L;
    build_scene_wg<W>(rows, n, nGrp, out, L);
}

// ---------------------------------------------------------------------------
// a query's tree in the reference's layout: rows of D floats
// (d_treeSamples_), parents and costs; entries past the tree are 0, -1, 0 as
// the reference's fills leave them (KGMT.cu:62-64)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_tree_rows(const float* __restrict__ treeS,
                                                      const int32_t* __restrict__ parent,
                                                      const float* __restrict__ cost, uint32_t cap, int D,
                                                      uint32_t treeSize, float* __restrict__ rows,
                                                      int32_t* __restrict__ parents, float* __restrict__ costs)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if(i >= cap) return;
    const bool in = i < treeSize;
    for(int d = 0; d < D; ++d) rows[(size_t)i * D + d] = in ? treeS[(size_t)d * cap + i] : 0.0f;
    parents[i] = in ? parent[i] : -1;
    costs[i]   = in ? cost[i] : 0.0f;
}

// ---------------------------------------------------------------------------
// stage-level kernels (single query, caller buffers)
// ---------------------------------------------------------------------------
template <int C>
__global__ __launch_bounds__(kBlock) void k_stage_draw(const QueryDesc* __restrict__ qd, uint32_t query,
                                                       uint32_t itr, uint32_t u0, uint32_t n, float* controls,
                                                       int32_t* steps, float* ufront)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if(i >= n) return;
    float uc[C];
    int32_t st;
    float uf;
    draw_child<C>(qd->cfg, query, itr, u0 + i, uc, st, uf);
#pragma unroll
    for(int k = 0; k < C; ++k) controls[(size_t)k * n + i] = uc[k];
    steps[i]  = st;
    ufront[i] = uf;
}

template <int MODEL>
__global__ __launch_bounds__(kBlock) void k_stage_propagate(const QueryDesc* __restrict__ qdp,
                                                            const int32_t* __restrict__ parentIdx,
                                                            const float* __restrict__ parentStates,
                                                            uint32_t nParents, const float* __restrict__ controls,
                                                            const int32_t* __restrict__ steps, uint32_t n,
                                                            const float4* __restrict__ obs, uint32_t nGrp,
                                                            float* child, uint8_t* valid, int32_t* r1o, int32_t* r2o)
{
    extern __shared__ float4 sObs[];
    stage_scene(obs, 0, nGrp, sObs);
    __syncthreads();
    stage_propagate_child<MODEL>(*qdp, parentIdx, parentStates, nParents, controls, steps, n, sObs, nGrp, child,
                                 valid, r1o, r2o);
}

// The planner's per-child path on caller buffers: children in waves of 64
// consecutive indices, handed to the wave's lanes longest-first by the same
// counting sort child_tile runs with one lane per child, then
// propagate_two_phase called by every lane of the wave as child_tile calls it
// (lanes without a child take part with an empty box), or the one-pass check
// for the quadcopter. check / box (optional) receive the check that decided
// each child and its edge box.
template <int MODEL>
__global__ __launch_bounds__(kBlock) void k_stage_propagate_planner(
  const QueryDesc* __restrict__ qdp, const int32_t* __restrict__ parentIdx, const float* __restrict__ parentStates,
  uint32_t nParents, const float* __restrict__ controls, const int32_t* __restrict__ steps, uint32_t n,
  const float4* __restrict__ obs, uint32_t nGrp, float* child, uint8_t* valid, int32_t* r1o, int32_t* r2o,
  uint8_t* checkOut, float* boxOut)
{
    constexpr int S = ModelDims<MODEL>::S, C = ModelDims<MODEL>::C, D = S + C + 1;
    extern __shared__ float4 sObs[];
    __shared__ float4 cand[kBlock / 64][2 * kCand];
    __shared__ uint32_t hist[kBlock];
    __shared__ uint16_t slot[kBlock];
    const QueryDesc& qd = *qdp;
    stage_scene(obs, 0, nGrp, sObs);
    __syncthreads();
    const uint32_t t    = threadIdx.x;
    const uint32_t base = blockIdx.x * kBlock;
    // counting sort of the wave's 64 children by step count, longest first
    // (bucket 63 of the wave: lanes past the end)
    {
        const uint32_t g0 = (t / 64u) * 64u;
        uint32_t bucket   = g0 + 63u;
        if(base + t < n)
            {
                int32_t k = steps[base + t] - qd.cfg.steps_min;
                k         = k < 0 ? 0 : (k > 62 ? 62 : k);
                bucket    = g0 + 62u - (uint32_t)k;
            }
        hist[t] = 0;
        wave_sync();
        const uint32_t rank = atomicAdd(&hist[bucket], 1u);
        wave_sync();
        const uint32_t h = hist[t];
        hist[t]          = g0 + wave_incl_scan(h) - h;
        wave_sync();
        slot[hist[bucket] + rank] = (uint16_t)t;
        wave_sync();
    }
    const uint32_t i = base + slot[t];
    const int32_t p  = i < n ? parentIdx[i] : -1;
    const bool has   = i < n && p >= 0 && (uint32_t)p < nParents;
    float s[S], uc[C];
    int32_t st = 0;
#pragma unroll
    for(int d = 0; d < S; ++d) s[d] = has ? parentStates[(size_t)d * nParents + p] : 0.0f;
#pragma unroll
    for(int k = 0; k < C; ++k) uc[k] = has ? controls[(size_t)k * n + i] : 0.0f;
    if(has) st = steps[i];
    const Scene sc{sObs, (int)nGrp, 0};
    bool ok = false;
    TwoPhaseTrace tr{KGMT_CHECK_ONE_PASS, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    if constexpr(MODEL != KGMT_MODEL_QUADCOPTER)
        ok = propagate_two_phase<MODEL>(qd, s, uc, st, sc, cand[t / 64], has, &tr);  // every lane
    else if(has)
        ok = propagate<MODEL>(qd, s, uc, st, sc);
    if(i >= n) return;
    if(!has)
        {
            // caller index out of range: no state is read; flagged by r1 = r2 = -1
            for(int d = 0; d < D; ++d) child[(size_t)d * n + i] = 0.0f;
            valid[i] = 0;
            r1o[i]   = -1;
            r2o[i]   = -1;
            return;
        }
#pragma unroll
    for(int d = 0; d < S; ++d) child[(size_t)d * n + i] = s[d];
#pragma unroll
    for(int k = 0; k < C; ++k) child[(size_t)(S + k) * n + i] = uc[k];
    child[(size_t)(D - 1) * n + i] = qd.cfg.dt * (float)st;
    int ix, iy, iz;
    const int r1 = region_of(qd, s[0], s[1], s[2], ix, iy, iz);
    valid[i]     = ok ? 1 : 0;
    r1o[i]       = r1;
    r2o[i]       = subregion_of(qd, s[0], s[1], s[2], r1, ix, iy, iz);
    if(checkOut) checkOut[i] = (uint8_t)tr.path;
    if(boxOut && MODEL != KGMT_MODEL_QUADCOPTER)
        for(int d = 0; d < 3; ++d) boxOut[(size_t)d * n + i] = tr.lo[d], boxOut[(size_t)(3 + d) * n + i] = tr.hi[d];
}

// Frontier update stage: the caller's tree and children staged into a
// one-query pool the way the iteration's propagate and score kernels leave
// it, so that the planner's own k_append and k_expand run on it. Per tree
// node its region (the reference's getRegion of the node, KGMT.cu:369); per
// child its region, its flag and, when flagged, its row (parent, parent's
// parent, sample) at the next free slot of its 32-child segment.
__global__ __launch_bounds__(kBlock) void k_stage_fu_nodes(const QueryDesc* __restrict__ qdp, Pools P, uint32_t tree,
                                                           const float* __restrict__ child,
                                                           const int32_t* __restrict__ childParent,
                                                           const uint8_t* __restrict__ fnext, uint32_t nC)
{
    const QueryDesc& qd = *qdp;
    const uint32_t i    = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t cap  = qd.cap;
    int ix, iy, iz;
    if(i < tree)
        P.r1node[i] = region_of(qd, P.treeS[i], P.treeS[(size_t)cap + i], P.treeS[2 * (size_t)cap + i], ix, iy, iz);
    if(i >= nC) return;
    const bool f = fnext[i] != 0;
    P.fnext[i]   = f ? 1 : 0;
    P.r1child[i] = region_of(qd, child[i], child[(size_t)nC + i], child[2 * (size_t)nC + i], ix, iy, iz);
    if(!f) return;
    const uint32_t s0 = i & ~31u;
    uint32_t before   = 0;
    for(uint32_t k = s0; k < i; ++k) before += fnext[k] ? 1u : 0u;
    const int32_t p  = childParent[i];
    const int32_t gp = p >= 0 && (uint32_t)p < tree ? P.parent[p] : -1;
    const int D      = qd.D, RW = row_words(D);
    float* row       = P.childR + (size_t)(s0 + before) * RW;
    row[0]           = __int_as_float(p);
    row[1]           = __int_as_float(gp);
    for(int d = 0; d < D; ++d) row[2 + d] = child[(size_t)d * nC + i];
    for(int e = D + 2; e < RW; ++e) row[e] = 0.0f;
}

// flagged children per chunk, the state an iteration's propagate leaves
// (accepted count and fAccept for the info output) and the work lists of the
// update jobs (one workgroup)
__global__ __launch_bounds__(kBlock) void k_stage_fu_state(const QueryDesc* __restrict__ qdp, QueryState* st,
                                                           Pools P, WorkLists W, uint32_t itr, uint32_t tree,
                                                           uint32_t nC)
{
    __shared__ uint32_t cnt[1024];
    const QueryDesc& qd = *qdp;
    const uint32_t nCC  = (nC + kChunk - 1) / kChunk;
    for(uint32_t c = threadIdx.x; c < nCC; c += kBlock)
        {
            uint32_t n = 0;
            for(uint32_t u = c * kChunk; u < nC && u < (c + 1) * kChunk; ++u) n += P.fnext[u];
            cnt[c] = n;
        }
    __syncthreads();
    if(threadIdx.x != 0) return;
    uint32_t acc = 0;
    for(uint32_t c = 0; c < nCC; ++c)
        {
            P.partA[c] = cnt[c];
            acc += cnt[c];
        }
    const uint32_t avail = qd.cap - tree;
    const uint32_t next  = acc < avail ? acc : avail;
    QueryState s{};
    s.itr = s.itrN = itr;
    s.treeSize = s.treeSizeN = tree;
    s.nChildren = s.nChildrenN = nC;
    s.nAcc      = acc;
    s.next      = next;
    s.fAccept   = faccept(qd, itr, tree, next);
    s.goalIdx   = kNoGoal;
    *st         = s;
    const uint32_t jobs = nCC + (tree + kChunk - 1) / kChunk;
    for(uint32_t i = 0; i < jobs; ++i) P.look[i] = 0ull;  // parity 0's words
    W.claim[0] = W.claim[1] = 0u;
    *W.error   = 0u;
    for(uint32_t par = 0; par < 2; ++par)
        {
            W.jobStart[2 * par] = 0, W.jobStart[2 * par + 1] = jobs;
            W.tileStart[2 * par] = 0, W.tileStart[2 * par + 1] = 0;
            W.active[par] = 1;
        }
}

// info[4] = {appended, expansion multiset size, goal node or -1, re-activated}
__global__ void k_stage_fu_info(const QueryDesc* __restrict__ qdp, const QueryState* st, Pools P, uint32_t tree,
                                int32_t* info)
{
    if(threadIdx.x != 0) return;
    // the inclusive entries of the last old-node job: the re-activated count
    const uint32_t nOC   = (tree + kChunk - 1) / kChunk;
    const uint32_t react = (uint32_t)P.look[nOC - 1];
    info[0] = (int32_t)st->next;
    info[1] = (int32_t)st->R;
    info[2] = st->goalIdx == kNoGoal ? -1 : st->goalIdx;
    info[3] = (int32_t)react;
}

__global__ __launch_bounds__(kBlock) void k_stage_region_update(const QueryDesc* __restrict__ qdp,
                                                                const int32_t* __restrict__ r1,
                                                                const int32_t* __restrict__ r2,
                                                                const uint8_t* __restrict__ valid,
                                                                const float* __restrict__ ufront, uint32_t n,
                                                                const float* __restrict__ score,
                                                                const uint32_t* __restrict__ actPrev,
                                                                uint32_t* actNext, int32_t* total, int32_t* validc,
                                                                uint8_t* fnext)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const QueryDesc& qd = *qdp;
    int key             = i < n ? r1[i] : 0;
    const int sub       = i < n ? r2[i] : 0;
    // children with a region id out of range are ignored
    const bool has = i < n && key >= 0 && key < qd.R1 && sub >= 0 && sub < qd.R2;
    if(!has) key = 0;
    const bool ok = has && valid[i] != 0;
    bool fn       = false;
    bool setBit   = false;
    if(ok)
        {
            const bool seen = test_bit(actPrev, (uint32_t)sub);
            fn              = (ufront[i] < score[key]) || !seen;
            setBit          = !seen;
        }
    if(i < n) fnext[i] = fn ? 1 : 0;
    wave_set_bits(actNext, has ? (uint32_t)sub : 0u, setBit);
    wave_count(total, validc, key, has, ok);
}

__global__ __launch_bounds__(kScoreThreads) void k_stage_score(const QueryDesc* __restrict__ qd,
                                                               const int32_t* __restrict__ total,
                                                               const int32_t* __restrict__ validc,
                                                               const uint32_t* __restrict__ act, float* score)
{
    __shared__ float a[kScoreThreads];
    score_regions<kScoreThreads>(*qd, total, validc, act, score, a);
}

// compaction: chunk counts -> one-workgroup scan -> scatter
__global__ __launch_bounds__(kBlock) void k_stage_count(const uint8_t* __restrict__ flags, uint32_t n, uint32_t* tmp)
{
    __shared__ uint32_t sbuf[kBlock / 64 + 1];
    const uint32_t c = blockIdx.x;
    uint32_t cnt     = 0;
    for(uint32_t k = 0; k < kChunk / kBlock; ++k)
        {
            const uint32_t i = c * kChunk + k * kBlock + threadIdx.x;
            cnt += (i < n && flags[i]) ? 1u : 0u;
        }
    uint32_t tot;
    block_excl_scan<kBlock>(cnt, sbuf, tot);
    if(threadIdx.x == 0) tmp[c] = tot;
}

__global__ __launch_bounds__(kScoreThreads) void k_stage_scan(uint32_t* tmp, uint32_t m, uint32_t* total)
{
    __shared__ uint32_t sbuf[kScoreThreads / 64 + 1];
    const uint32_t t = block_scan_array<kScoreThreads>(tmp, m, sbuf);
    if(threadIdx.x == 0) *total = t;
}

__global__ __launch_bounds__(kBlock) void k_stage_scatter(const uint8_t* __restrict__ flags, uint32_t n,
                                                          const uint32_t* __restrict__ tmp, uint32_t* out)
{
    __shared__ uint32_t sbuf[kBlock / 64 + 1];
    const uint32_t c = blockIdx.x;
    uint32_t run     = tmp[c];
    for(uint32_t k = 0; k < kChunk / kBlock; ++k)
        {
            const uint32_t i = c * kChunk + k * kBlock + threadIdx.x;
            const bool f     = i < n && flags[i] != 0;
            uint32_t tot;
            const uint32_t ex = block_excl_scan<kBlock>(f ? 1u : 0u, sbuf, tot);
            if(f) out[run + ex] = i;
            run += tot;
        }
}

// segmented repeat: chunk sums -> scan -> write
__global__ __launch_bounds__(kBlock) void k_stage_sum(const uint32_t* __restrict__ counts, uint32_t n, uint32_t* tmp)
{
    __shared__ uint32_t sbuf[kBlock / 64 + 1];
    const uint32_t c = blockIdx.x;
    uint32_t sum     = 0;
    for(uint32_t k = 0; k < kChunk / kBlock; ++k)
        {
            const uint32_t i = c * kChunk + k * kBlock + threadIdx.x;
            sum += i < n ? counts[i] : 0u;
        }
    uint32_t tot;
    block_excl_scan<kBlock>(sum, sbuf, tot);
    if(threadIdx.x == 0) tmp[c] = tot;
}

__global__ __launch_bounds__(kBlock) void k_stage_repeat(const uint32_t* __restrict__ ids,
                                                         const uint32_t* __restrict__ counts, uint32_t n,
                                                         const uint32_t* __restrict__ tmp, uint32_t* out,
                                                         uint32_t outCap)
{
    __shared__ uint32_t sbuf[kBlock / 64 + 1];
    const uint32_t c = blockIdx.x;
    uint32_t run     = tmp[c];
    for(uint32_t k = 0; k < kChunk / kBlock; ++k)
        {
            const uint32_t i   = c * kChunk + k * kBlock + threadIdx.x;
            const uint32_t cnt = i < n ? counts[i] : 0u;
            uint32_t tot;
            const uint32_t ex = block_excl_scan<kBlock>(cnt, sbuf, tot);
            for(uint32_t m = 0; m < cnt; ++m)
                if(run + ex + m < outCap) out[run + ex + m] = ids[i];
            run += tot;
        }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_init(const QueryDesc* qds, QueryState* sts, const Pools& P, const WorkLists& W, const float* rows,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_init, dim3(W.nq), dim3(kBlock), 0, s, qds, sts, P, W, W.nq, rows);
    return hipGetLastError();
}

hipError_t launch_init_one(const QueryDesc& qd, QueryDesc* dq, QueryState* sts, const Pools& P, const WorkLists& W,
                           const float* rows, hipStream_t s)
{
    // about 4 K words to clear per clearing workgroup
    const uint32_t words = 3u * (uint32_t)qd.R1 + 2u * (uint32_t)qd.R2words + 9u * (qd.nChunks + 1u);
    const uint32_t clear = std::min(32u, std::max(1u, (words + 4095u) / 4096u));
    hipLaunchKernelGGL(k_init_one, dim3(1 + clear), dim3(kBlock), 0, s, qd, dq, sts, P, W, rows);
    return hipGetLastError();
}

template <int NT, int K>
static hipError_t launch_propagate_nt(int model, const QueryDesc* qds, QueryState* sts, const Pools& P,
                                      const WorkLists& W, uint32_t parity, uint32_t maxPairs, const Launch& L,
                                      hipStream_t s)
{
    const size_t lds = (size_t)maxPairs * 2 * sizeof(float4);
    const dim3 g(L.propGrid), b(NT);
    switch(model)
        {
            case KGMT_MODEL_DOUBLE_INTEGRATOR:
                hipLaunchKernelGGL((k_propagate<KGMT_MODEL_DOUBLE_INTEGRATOR, NT, K>), g, b, lds, s, qds, sts, P, W,
                                   parity);
                break;
            case KGMT_MODEL_DUBINS_AIRPLANE:
                hipLaunchKernelGGL((k_propagate<KGMT_MODEL_DUBINS_AIRPLANE, NT, K>), g, b, lds, s, qds, sts, P, W,
                                   parity);
                break;
            case KGMT_MODEL_DUBINS_5D:
                hipLaunchKernelGGL((k_propagate<KGMT_MODEL_DUBINS_5D, NT, K>), g, b, lds, s, qds, sts, P, W, parity);
                break;
            case KGMT_MODEL_QUADCOPTER:
                hipLaunchKernelGGL((k_propagate<KGMT_MODEL_QUADCOPTER, NT, K>), g, b, lds, s, qds, sts, P, W, parity);
                break;
            case KGMT_MODEL_UNICYCLE:
                hipLaunchKernelGGL((k_propagate<KGMT_MODEL_UNICYCLE, NT, K>), g, b, lds, s, qds, sts, P, W, parity);
                break;
            default: hipLaunchKernelGGL((k_propagate<-1, NT, K>), g, b, lds, s, qds, sts, P, W, parity); break;
        }
    return hipGetLastError();
}

bool launch_supported(uint32_t block, uint32_t lanes)
{
    return (lanes == 1 && (block == 128 || block == 256 || block == 512)) ||
           (block == 256 && (lanes == 2 || lanes == 4));
}

hipError_t launch_propagate(int model, const QueryDesc* qds, QueryState* sts, const Pools& P, const WorkLists& W,
                            uint32_t parity, uint32_t maxPairs, const Launch& L, hipStream_t s)
{
    if(!launch_supported(L.propBlock, L.lanes) || W.tile != L.propBlock / L.lanes) return hipErrorInvalidValue;
    if(L.lanes == 2) return launch_propagate_nt<256, 2>(model, qds, sts, P, W, parity, maxPairs, L, s);
    if(L.lanes == 4) return launch_propagate_nt<256, 4>(model, qds, sts, P, W, parity, maxPairs, L, s);
    switch(L.propBlock)
        {
            case 128: return launch_propagate_nt<128, 1>(model, qds, sts, P, W, parity, maxPairs, L, s);
            case 256: return launch_propagate_nt<256, 1>(model, qds, sts, P, W, parity, maxPairs, L, s);
            default: return launch_propagate_nt<512, 1>(model, qds, sts, P, W, parity, maxPairs, L, s);
        }
}

hipError_t launch_update(const QueryDesc* qds, QueryState* sts, const Pools& P, const WorkLists& W,
                         uint32_t parity, const Launch& L, bool own, hipStream_t s)
{
    hipLaunchKernelGGL(k_update, dim3(L.chunkGrid), dim3(kUpd), 0, s, qds, sts, P, W, parity, own ? 1u : 0u);
    return hipGetLastError();
}

hipError_t launch_build_scene(const float* rows, uint32_t n, uint32_t nGrp, float4* out, int wdim, hipStream_t s)
{
    if(n == 0) return hipSuccess;
    if(wdim == 2)
        hipLaunchKernelGGL(k_build_scene<2>, dim3(1), dim3(kBlock), 0, s, rows, n, nGrp, out);
    else
        hipLaunchKernelGGL(k_build_scene<3>, dim3(1), dim3(kBlock), 0, s, rows, n, nGrp, out);
    return hipGetLastError();
}

hipError_t launch_finish(const QueryDesc* qds, QueryState* sts, const WorkLists& W, uint32_t parity, hipStream_t s)
{
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, s, qds, sts, W, parity);
    return hipGetLastError();
}

hipError_t launch_paths(const QueryDesc* qds, QueryState* sts, const Pools& P, uint32_t nq, hipStream_t s)
{
    if(nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_paths, dim3((nq + 63) / 64), dim3(64), 0, s, qds, sts, P, nq);
    return hipGetLastError();
}

hipError_t launch_tree_rows(const float* treeS, const int32_t* parent, const float* cost, uint32_t cap, int D,
                            uint32_t treeSize, float* rows, int32_t* parents, float* costs, hipStream_t s)
{
    if(cap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tree_rows, dim3((cap + kBlock - 1) / kBlock), dim3(kBlock), 0, s, treeS, parent, cost, cap,
                       D, treeSize, rows, parents, costs);
    return hipGetLastError();
}

hipError_t launch_stage_draw(const QueryDesc* qd, int C, uint32_t query, uint32_t itr, uint32_t u0, uint32_t n,
                             float* controls, int32_t* steps, float* ufront, hipStream_t s)
{
    const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
    if(n == 0) return hipSuccess;
    switch(C)
        {
            case 2: hipLaunchKernelGGL(k_stage_draw<2>, g, b, 0, s, qd, query, itr, u0, n, controls, steps, ufront); break;
            case 3: hipLaunchKernelGGL(k_stage_draw<3>, g, b, 0, s, qd, query, itr, u0, n, controls, steps, ufront); break;
            case 4: hipLaunchKernelGGL(k_stage_draw<4>, g, b, 0, s, qd, query, itr, u0, n, controls, steps, ufront); break;
            default: return hipErrorInvalidValue;
        }
    return hipGetLastError();
}

hipError_t launch_stage_propagate(int model, const QueryDesc* qd, const int32_t* parentIdx, const float* parentStates,
                                  uint32_t nParents, const float* controls, const int32_t* steps, uint32_t n,
                                  const float4* obs, uint32_t nGrp, uint32_t nObs, float* child, uint8_t* valid,
                                  int32_t* r1, int32_t* r2, hipStream_t s)
{
    if(n == 0) return hipSuccess;
    const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
    const size_t lds = (size_t)(nGrp ? nGrp * (1 + kGroup) : 1) * 2 * sizeof(float4);
    (void)nObs;
#define KGMT_STAGE_PROP(M)                                                                                     \
    hipLaunchKernelGGL(k_stage_propagate<M>, g, b, lds, s, qd, parentIdx, parentStates, nParents, controls, steps, \
                       n, obs, nGrp, child, valid, r1, r2)
    switch(model)
        {
            case KGMT_MODEL_DOUBLE_INTEGRATOR: KGMT_STAGE_PROP(KGMT_MODEL_DOUBLE_INTEGRATOR); break;
            case KGMT_MODEL_DUBINS_AIRPLANE: KGMT_STAGE_PROP(KGMT_MODEL_DUBINS_AIRPLANE); break;
            case KGMT_MODEL_DUBINS_5D: KGMT_STAGE_PROP(KGMT_MODEL_DUBINS_5D); break;
            case KGMT_MODEL_QUADCOPTER: KGMT_STAGE_PROP(KGMT_MODEL_QUADCOPTER); break;
            case KGMT_MODEL_UNICYCLE: KGMT_STAGE_PROP(KGMT_MODEL_UNICYCLE); break;
            default: return hipErrorInvalidValue;
        }
#undef KGMT_STAGE_PROP
    return hipGetLastError();
}

hipError_t launch_stage_propagate_planner(int model, const QueryDesc* qd, const int32_t* parentIdx,
                                          const float* parentStates, uint32_t nParents, const float* controls,
                                          const int32_t* steps, uint32_t n, const float4* obs, uint32_t nGrp,
                                          float* child, uint8_t* valid, int32_t* r1, int32_t* r2, uint8_t* check,
                                          float* box, hipStream_t s)
{
    if(n == 0) return hipSuccess;
    const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
    const size_t lds = (size_t)(nGrp ? nGrp * (1 + kGroup) : 1) * 2 * sizeof(float4);
#define KGMT_STAGE_PROP(M)                                                                                        \
    hipLaunchKernelGGL(k_stage_propagate_planner<M>, g, b, lds, s, qd, parentIdx, parentStates, nParents, controls, \
                       steps, n, obs, nGrp, child, valid, r1, r2, check, box)
    switch(model)
        {
            case KGMT_MODEL_DOUBLE_INTEGRATOR: KGMT_STAGE_PROP(KGMT_MODEL_DOUBLE_INTEGRATOR); break;
            case KGMT_MODEL_DUBINS_AIRPLANE: KGMT_STAGE_PROP(KGMT_MODEL_DUBINS_AIRPLANE); break;
            case KGMT_MODEL_DUBINS_5D: KGMT_STAGE_PROP(KGMT_MODEL_DUBINS_5D); break;
            case KGMT_MODEL_QUADCOPTER: KGMT_STAGE_PROP(KGMT_MODEL_QUADCOPTER); break;
            case KGMT_MODEL_UNICYCLE: KGMT_STAGE_PROP(KGMT_MODEL_UNICYCLE); break;
            default: return hipErrorInvalidValue;
        }
#undef KGMT_STAGE_PROP
    return hipGetLastError();
}

hipError_t launch_stage_frontier_update(const QueryDesc* qd, QueryState* st, const Pools& P, const WorkLists& W,
                                        uint32_t itr, uint32_t tree, const float* child, const int32_t* childParent,
                                        const uint8_t* fnext, uint32_t nC, int32_t* info, hipStream_t s)
{
    const uint32_t n = tree > nC ? tree : nC;
    if(n) hipLaunchKernelGGL(k_stage_fu_nodes, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, qd, P, tree, child,
                             childParent, fnext, nC);
    hipLaunchKernelGGL(k_stage_fu_state, dim3(1), dim3(kBlock), 0, s, qd, st, P, W, itr, tree, nC);
    Launch L;
    L.chunkGrid = 256;
    const hipError_t e = launch_update(qd, st, P, W, 0, L, false, s);
    if(e != hipSuccess) return e;
    hipLaunchKernelGGL(k_stage_fu_info, dim3(1), dim3(64), 0, s, qd, st, P, tree, info);
    return hipGetLastError();
}

hipError_t launch_stage_region_update(const QueryDesc* qd, const int32_t* r1, const int32_t* r2, const uint8_t* valid,
                                      const float* ufront, uint32_t n, const float* score, const uint32_t* actPrev,
                                      uint32_t* actNext, int32_t* total, int32_t* validc, uint8_t* fnext,
                                      hipStream_t s)
{
    if(n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_region_update, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, qd, r1, r2, valid,
                       ufront, n, score, actPrev, actNext, total, validc, fnext);
    return hipGetLastError();
}

hipError_t launch_stage_score(const QueryDesc* qd, const int32_t* total, const int32_t* validc, const uint32_t* act,
                              float* score, hipStream_t s)
{
    hipLaunchKernelGGL(k_stage_score, dim3(1), dim3(kScoreThreads), 0, s, qd, total, validc, act, score);
    return hipGetLastError();
}

hipError_t launch_stage_compact(const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t* count, uint32_t* tmp,
                                hipStream_t s)
{
    const uint32_t m = (n + kChunk - 1) / kChunk;
    if(m) hipLaunchKernelGGL(k_stage_count, dim3(m), dim3(kBlock), 0, s, flags, n, tmp);
    hipLaunchKernelGGL(k_stage_scan, dim3(1), dim3(kScoreThreads), 0, s, tmp, m, count);
    if(m) hipLaunchKernelGGL(k_stage_scatter, dim3(m), dim3(kBlock), 0, s, flags, n, tmp, out);
    return hipGetLastError();
}

hipError_t launch_stage_repeat(const uint32_t* ids, const uint32_t* counts, uint32_t n, uint32_t* out,
                               uint32_t outCap, uint32_t* total, uint32_t* tmp, hipStream_t s)
{
    const uint32_t m = (n + kChunk - 1) / kChunk;
    if(m) hipLaunchKernelGGL(k_stage_sum, dim3(m), dim3(kBlock), 0, s, counts, n, tmp);
    hipLaunchKernelGGL(k_stage_scan, dim3(1), dim3(kScoreThreads), 0, s, tmp, m, total);
    if(m) hipLaunchKernelGGL(k_stage_repeat, dim3(m), dim3(kBlock), 0, s, ids, counts, n, tmp, out, outCap);
    return hipGetLastError();
}

}  // namespace kgmt
