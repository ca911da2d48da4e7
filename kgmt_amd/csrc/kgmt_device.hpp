// kgmt_device.hpp — device math of the KGMT expansion path for gfx950.
//
// Everything here is evaluated per lane in registers. The file is compiled with
// -ffp-contract=off: each float operator is one IEEE operation, in source
// order, so results match a CPU build of the same expressions bit for bit
// (division and sqrt are correctly rounded in HIP by default).
//
// Reference (paths relative to the KGMT repository):
//   dynamics      src/statePropagator/statePropagator.cu:76-142, 147-241, 246-370
//   collision     src/collisionCheck/collisionCheck.cu:3-10, 49-62
//   regions       src/graphs/Graph.cu:53-97 (min corners), 99-153, 155-205
//   goal distance include/helper/helper.cuh:138-152
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/kgmt.h"

namespace kgmt
{

constexpr uint32_t kChunk   = 1024;  // elements per bookkeeping chunk (one workgroup job)
constexpr uint32_t kNodeAlign = 16;  // per-query node ranges start 16-aligned (vector loads of 4 flags)
constexpr uint32_t kTile    = 256;   // default children per propagate tile (128, 256 or 512)
constexpr uint32_t kBlock   = 256;   // threads of chunk / tile workgroups
constexpr uint32_t kMaxObs  = 1500;  // obstacles a query may carry
constexpr uint32_t kGroup   = 8;     // boxes per bounding group of a scene
constexpr uint32_t kGroupBatch = 4;  // groups tested together (scenes pad to a multiple)
// LDS of one scene: G groups + 8 G boxes, G = ceil(n / 32) * 4, as (lo, hi)
// float4 pairs: at most 1692 pairs = 54.1 KB (+ 58.5 KB of tile state and
// candidate lists for a 512-thread propagate tile: 113 KB of gfx950's 160 KB)
constexpr int32_t kNoGoal   = 0x7fffffff;

// -------------------------------------------------------------------------
// Per-query descriptor (read-only during a plan) and mutable state.
// -------------------------------------------------------------------------
struct QueryDesc
{
    kgmt_config cfg;
    int32_t S, C, D;                 // state, control, sample dims
    int32_t R1, P, R2, R2words;      // regions, subregions per region, subregions, bitset words
    uint32_t cap, nChunks;           // capacity, ceil(cap / kChunk)
    uint64_t sampleOff;              // float offset of this query's [D][cap] sample blocks
    uint64_t rowOffC;                // float offset of this query's flagged-child rows ([cap][row_words(D)])
    uint64_t nodeOff;                // element offset of per-node arrays (cap entries)
    uint32_t chunkOff;               // offset of per-chunk arrays (nChunks + 1 entries)
    uint32_t regOff;                 // offset of per-region arrays (R1 entries)
    uint32_t subOff;                 // offset of subregion bitset words (R2words entries)
    uint32_t obsOff, nObs, nGrp;     // first pair, boxes, groups (multiple of 4) of the scene in the pool
    uint32_t pathOff;                // float offset of the path rows ((max_iter+2)*D)
    float r1size, vol;               // W_R1_SIZE, W_R1_VOL
    float invSpan, invR1, invP;      // exact 1 / (ws_max - ws_min), 1 / r1size, 1 / P when powers of two, else 0
    float goal[3];
    float start[KGMT_MAX_SAMPLE];
    float qminv, qax, qbx, qay, qby, qaz, qbz;  // quad: 1/m, (Iy-Iz)/Ix, 1/Ix, ...
    uint32_t rowOff;                 // first obstacle row (2 * wdim floats) of the query in the scene's source rows
    int32_t wdim;                    // workspace dimension W_DIM: 2 (planar: x y, z = 0) or 3
};

// words of a flagged-child row: parent and grandparent indices + D sample
// words, padded to whole float4s
__host__ __device__ constexpr int row_words(int D)
{
    return (D + 5) & ~3;
}

struct QueryState
{
    uint32_t itr;        // current iteration (1-based)
    uint32_t treeSize;   // nodes in the tree
    uint32_t nChildren;  // children to propagate this iteration
    uint32_t per;        // children per expansion entry
    uint32_t R;          // size of the expansion multiset
    uint32_t nAcc;       // children flagged for the next frontier
    uint32_t next;       // children appended (nAcc clamped to capacity)
    float fAccept;       // re-activation bonus
    uint32_t done;       // 1 once solved or out of iterations
    uint32_t success;
    int32_t goalIdx;     // smallest tree index within goal_thresh (kNoGoal if none)
    uint32_t pathRows;
    float pathCost, pathLen;
    uint64_t expanded;
    uint64_t flagged;    // children flagged for the tree over all iterations (rows written)
    uint32_t finishItr;
    uint32_t pad[3];
    // The next iteration's values, written when an iteration closes (by the
    // last update job of a single-query plan, by k_finish for a batch);
    // propagate reads them and its first workgroup copies them to the fields
    // above (k_paths after the last iteration), which the update kernel reads.
    uint32_t itrN, treeSizeN, nChildrenN, perN, doneN, successN, finishItrN, padN;
    uint64_t expandedN, flaggedN;
};

__device__ __forceinline__ void commit_state(QueryState& st)
{
    st.itr       = st.itrN;
    st.treeSize  = st.treeSizeN;
    st.nChildren = st.nChildrenN;
    st.per       = st.perN;
    st.done      = st.doneN;
    st.success   = st.successN;
    st.finishItr = st.finishItrN;
    st.expanded  = st.expandedN;
    st.flagged   = st.flaggedN;
}

// Pointers of the device pools (shared by all queries of a batch).
struct Pools
{
    float* treeS;        // tree samples        [D][cap] per query
    float* childR;       // rows of this iteration's flagged children, [cap][row_words(D)] per query: the
                         // parent's and grandparent's index bits, then the D sample words (zero padded);
                         // dense per 32-child segment
    int32_t* parent;     // tree parents
    int32_t* grand;      // tree grandparents (-1 for the root and its children)
    int4* anc;           // tree ancestors at distances 3, 4, 5, 6 (-1 past the root)
    float* cost;         // distance of node position to goal
    uint8_t* frontier;   // frontier flag per tree node
    int32_t* r1node;     // region of each tree node
    int32_t* r1child;    // region of each child
    uint8_t* fnext;      // child flagged for the next frontier
    uint8_t* react;      // old node re-activated this iteration
    uint32_t* repIdx;    // expansion list: node id of multiset entry r
    uint32_t* partA;     // per child chunk: flagged children, then exclusive offsets
    uint64_t* look;      // look-back word per update job: 4 x (nChunks + 1) per query, two sets by parity
    int32_t* total;      // per region: children counted
    int32_t* validc;     // per region: valid children
    float* score;        // per region score
    uint32_t* actPrev;   // subregion bitset as of iteration start
    uint32_t* actNext;   // subregion bitset being built
    float4* obs;         // scenes: per query nGrp group bounds then nObs boxes, (lo, hi) float4 pairs
    float* path;         // goal->root rows
};

// Result of a single-query plan, written by the device into host-mapped
// memory before the completion word (the host then returns at once, leaving
// the launches it queued past the last iteration to exit on the GPU).
struct HostResult
{
    uint32_t success, iterations, treeSize, pathRows;
    int32_t goalIdx;
    float pathCost, pathLen;
    uint32_t pad;
    uint64_t expanded, flagged;
};

// Work lists over queries, double-buffered by iteration parity: the kernels
// of iteration i read buffer (i-1)&1 while the last finisher of that
// iteration writes the other one. Prefix sums of nq + 1 entries each.
struct WorkLists
{
    uint32_t* tileStart;   // 2 x (nq + 1): propagate tiles
    uint32_t* jobStart;    // 2 x (nq + 1): update jobs
    uint32_t* active;      // 2: queries not done in each list
    uint32_t* counter;     // arrivals of k_init's workgroups (the last builds the first work lists)
    uint32_t* claim;       // 2 (by parity): update jobs taken in a launch (the launch before zeroes it)
    uint32_t* allDone;     // host-mapped word: 1 once every query is done
    uint32_t* error;       // 1 once an update job's look-back wait ran out (host-mapped for plans)
    uint32_t* closed;      // host-mapped word: iterations closed in this plan (the host paces on it)
    uint32_t* closedDev;   // the same count in device memory (its one writer reads it back)
    HostResult* result;    // host-mapped result of a single-query plan (null: batch, or no early result)
    float* hostPath;       // host-mapped goal -> root rows of that plan ((max_iter + 2) x D floats)
    uint32_t nq;
    uint32_t tile;         // children per propagate tile (= its workgroup size)

    __device__ __forceinline__ const uint32_t* tiles(uint32_t parity) const { return tileStart + parity * (nq + 1); }
    __device__ __forceinline__ const uint32_t* jobs(uint32_t parity) const { return jobStart + parity * (nq + 1); }
};

// -------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11), published multipliers and Weyl keys.
// -------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for(int round = 0; round < 10; ++round)
        {
            if(round > 0)
                {
                    k0 += 0x9E3779B9u;
                    k1 += 0xBB67AE85u;
                }
            const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
            const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
            const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
            c0 = n0, c1 = lo1, c2 = n2, c3 = lo0;
        }
    out[0] = c0, out[1] = c1, out[2] = c2, out[3] = c3;
}

// uniform in (0, 1]: top 24 bits plus one, times 2^-24
__device__ __forceinline__ float u01(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f;
}

__device__ __forceinline__ void draw_block(const kgmt_config& c, uint32_t q, uint32_t itr, uint32_t idx,
                                           uint32_t stream, uint32_t out[4])
{
    philox4x32_10(idx, itr, stream, q, (uint32_t)c.seed, (uint32_t)(c.seed >> 32), out);
}

// Controls (sample order) from words d[0..C), steps from d[C], frontier
// uniform from d[7]; blocks 0 and 1 of counter (u, itr, *, q).
template <int C>
__device__ __forceinline__ void draw_child(const kgmt_config& c, uint32_t q, uint32_t itr, uint32_t u, float* ctrl,
                                           int32_t& steps, float& ufront)
{
    uint32_t d[8];
    draw_block(c, q, itr, u, 0, d);
    draw_block(c, q, itr, u, 1, d + 4);
#pragma unroll
    for(int i = 0; i < C; ++i)
        {
            const float range = c.ctrl_max[i] - c.ctrl_min[i];
            ctrl[i]           = c.ctrl_min[i] + u01(d[i]) * range;
        }
    const int32_t n = c.steps_max - c.steps_min + 1;
    int32_t k       = (int32_t)(u01(d[C]) * (float)n);
    if(k > n - 1) k = n - 1;
    steps  = c.steps_min + k;
    ufront = u01(d[7]);
}

__device__ __forceinline__ float draw_react(const kgmt_config& c, uint32_t q, uint32_t itr, uint32_t node)
{
    uint32_t d[4];
    draw_block(c, q, itr, node, 2, d);
    return u01(d[0]);
}

// -------------------------------------------------------------------------
// sin/cos: Cody–Waite reduction by pi/2 (split constants with 8 and 11
// significant bits, so k*P1, k*P2 are exact for |k| < 2^13) and Taylor
// polynomials (sin to x^9, cos to x^12) on [-pi/4, pi/4].
// -------------------------------------------------------------------------
__device__ __forceinline__ void sincos_k(float x, float& s, float& c)
{
    const float fk = __builtin_rintf(x * 0x1.45f306p-1f);
    float r        = x - fk * 0x1.92p+0f;
    r              = r - fk * 0x1.fb4p-12f;
    r              = r - fk * 0x1.4442d2p-24f;
    const int q    = (int)fk;
    const float z  = r * r;
    float ps       = 0x1.71de3ap-19f;
    ps             = ps * z + -0x1.a01a02p-13f;
    ps             = ps * z + 0x1.111112p-7f;
    ps             = ps * z + -0x1.555556p-3f;
    const float sr = r + (r * z) * ps;
    float pc       = 0x1.1eed8ep-29f;
    pc             = pc * z + -0x1.27e4fcp-22f;
    pc             = pc * z + 0x1.a01a02p-16f;
    pc             = pc * z + -0x1.6c16c2p-10f;
    pc             = pc * z + 0x1.555556p-5f;
    const float cr = (1.0f - 0.5f * z) + (z * z) * pc;
    const int m    = q & 3;
    const float a  = (m & 1) ? cr : sr;
    const float b  = (m & 1) ? sr : cr;
    s              = (m & 2) ? -a : a;
    c              = ((m + 1) & 2) ? -b : b;
}

// -------------------------------------------------------------------------
// Collision: segment AABB vs boxes staged in LDS as (lo, hi) float4 pairs.
// A box is hit unless some axis separates (bbMax <= lo or hi <= bbMin),
// exactly the reference's test (a NaN face never separates, like there).
// The scene holds the boxes in groups of kGroup with a bounding box per group
// (componentwise min of lo / max of hi, NaN if a member has NaN); a group
// that no lane of the wave can touch is skipped as a whole. Groups are padded
// to whole batches with dummy boxes lo = +inf, hi = -inf, which separate from
// every segment that has one non-NaN axis (a segment with NaN on all axes
// already hits every real box of that group). Skipping and padding are exact:
// a box hit implies a hit of its group's bounds, and the verdict is an OR
// over boxes, independent of their order.
// -------------------------------------------------------------------------
// A scene staged in LDS: nGrp group (lo, hi) pairs, then the lo corners of
// its kGroup * nGrp boxes, then their hi corners (lanes that read
// consecutive boxes read consecutive 16-byte words: no bank conflicts).
struct Scene
{
    const float4* grp;  // the staged scene
    int nGrp;
    int sub = 0;        // lane within the child's group of K lanes
    __device__ __forceinline__ const float4* box_lo() const { return grp + 2 * nGrp; }
    __device__ __forceinline__ const float4* box_hi() const { return grp + 2 * nGrp + nGrp * (int)kGroup; }
};

// copy a scene of nGrp groups, stored as (lo, hi) float4 pairs at pool + 2
// off, into LDS in the layout of Scene
__device__ __forceinline__ void stage_scene(const float4* __restrict__ pool, uint32_t off, uint32_t nGrp,
                                            float4* sObs)
{
    const uint32_t nBox = nGrp * kGroup, n = 2 * (nGrp + nBox);
    for(uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        {
            const uint32_t k = i - 2 * nGrp;  // box word: box k / 2, corner k % 2
            sObs[i < 2 * nGrp ? i : 2 * nGrp + (k & 1u) * nBox + (k >> 1)] = pool[2 * (size_t)off + i];
        }
}

// OR of a flag over each group of K consecutive lanes (K = 1, 2, 4), by DPP
// quad permutations; every lane of a group must be active.
template <int K>
__device__ __forceinline__ bool group_or(bool v)
{
    static_assert(K == 1 || K == 2 || K == 4, "lanes per child");
    if constexpr(K == 1)
        return v;
    else
        {
            int x = v ? 1 : 0;
            x |= __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]: lane ^ 1
            if constexpr(K == 4) x |= __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);  // [2,3,0,1]: lane ^ 2
            return x != 0;
        }
}

__device__ __forceinline__ bool separated(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, float4 lo,
                                          float4 hi)
{
    return (mxx <= lo.x) | (hi.x <= mnx) | (mxy <= lo.y) | (hi.y <= mny) | (mxz <= lo.z) | (hi.z <= mnz);
}

// Segment vs every box of the scene (collisionCheck.cu:49-62 per step). With
// K lanes per child, lane `sc.sub` of the group tests groups sub, sub + K, ...
// and the verdicts are ORed over the group, so every lane returns the same.
template <int K = 1>
__device__ __forceinline__ bool segment_free(float ax, float ay, float az, float bx, float by, float bz,
                                             const Scene& sc)
{
    const float mnx = ax > bx ? bx : ax, mxx = ax > bx ? ax : bx;
    const float mny = ay > by ? by : ay, mxy = ay > by ? ay : by;
    const float mnz = az > bz ? bz : az, mxz = az > bz ? az : bz;
    const float4* blo = sc.box_lo();
    const float4* bhi = sc.box_hi();
    bool hit          = false;
    for(int g0 = 0; g0 < sc.nGrp; g0 += (int)kGroupBatch * K)
        {
            bool near[kGroupBatch];
            int gi[kGroupBatch];
#pragma unroll
            for(int k = 0; k < (int)kGroupBatch; ++k)
                {
                    if constexpr(K == 1)
                        {
                            gi[k]   = g0 + k;
                            near[k] = !separated(mnx, mny, mnz, mxx, mxy, mxz, sc.grp[2 * gi[k]], sc.grp[2 * gi[k] + 1]);
                        }
                    else
                        {
                            // past the last group: re-test the last one (an OR
                            // does not change when a box is tested twice)
                            const int g = g0 + k * K + sc.sub;
                            gi[k]       = g < sc.nGrp ? g : sc.nGrp - 1;
                            near[k]     = !separated(mnx, mny, mnz, mxx, mxy, mxz, sc.grp[2 * gi[k]], sc.grp[2 * gi[k] + 1]);
                        }
                }
#pragma unroll
            for(int k = 0; k < (int)kGroupBatch; ++k)
                {
                    if(__any(near[k] && !hit))
                        {
                            const float4* lo = blo + (int)kGroup * gi[k];
                            const float4* hi = bhi + (int)kGroup * gi[k];
#pragma unroll 4
                            for(int m = 0; m < (int)kGroup; ++m)
                                hit |= !separated(mnx, mny, mnz, mxx, mxy, mxz, lo[m], hi[m]);
                        }
                }
        }
    return !group_or<K>(hit);
}

// Segment-check policies of the dynamics below: chk(px, py, pz, x, y, z)
// returns whether the step's segment is free.
template <int K>
struct FullCheck  // every box of the scene (K lanes split the groups)
{
    Scene sc;
    __device__ __forceinline__ bool operator()(float px, float py, float pz, float x, float y, float z) const
    {
        return segment_free<K>(px, py, pz, x, y, z, sc);
    }
};

struct DryRun  // no boxes; records the box spanned by every checked segment
{
    float lo[3], hi[3];
    bool finite = true;
    __device__ __forceinline__ void start(float x, float y, float z)
    {
        lo[0] = hi[0] = x, lo[1] = hi[1] = y, lo[2] = hi[2] = z;
        finite = (x == x) && (y == y) && (z == z);
    }
    __device__ __forceinline__ bool operator()(float, float, float, float x, float y, float z)
    {
        finite = finite && (x == x) && (y == y) && (z == z);
        lo[0] = x < lo[0] ? x : lo[0], hi[0] = x > hi[0] ? x : hi[0];
        lo[1] = y < lo[1] ? y : lo[1], hi[1] = y > hi[1] ? y : hi[1];
        lo[2] = z < lo[2] ? z : lo[2], hi[2] = z > hi[2] ? z : hi[2];
        return true;
    }
};

// separated() for a segment box with finite faces and a box whose faces are
// finite, -inf low or +inf high, or NaN: on each axis the two face gaps
// lo - max and min - hi are >= 0 exactly when the comparisons of separated()
// hold (denormals are kept, so a difference is 0 only for equal operands),
// a NaN gap counts as not separating, and fmaxf keeps the non-NaN operand.
__device__ __forceinline__ bool separated_gap(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                              float4 lo, float4 hi)
{
    const float gx = fmaxf(lo.x - mxx, mnx - hi.x);
    const float gy = fmaxf(lo.y - mxy, mny - hi.y);
    const float gz = fmaxf(lo.z - mxz, mnz - hi.z);
    return fmaxf(fmaxf(gx, gy), gz) >= 0.0f;
}

// A lane's own candidate boxes: the entries of the wave's candidate list
// ((lo, hi) float4 pairs in LDS) that overlap the lane's edge box. Every
// segment the lane tests lies inside its edge box, so a box outside it can
// meet none of them, and testing only these gives the verdicts of testing the
// whole list. The first R are held in registers for the whole edge (the loop
// runs to the most any lane of the wave holds; a lane tests only its own),
// the rest are read from LDS. Positions tested here are finite (the dry run
// or the Dubins bound was finite), and a candidate overlaps a finite box, so
// its low faces are < +inf and its high faces > -inf: separated_gap gives
// separated()'s verdict.
__device__ __forceinline__ float wave_maxf(float v);

struct LaneCandCheck
{
    static constexpr int R = 4;
    const float4* c;
    uint64_t restLo, restHi;  // own candidates past the first R (bit = list index, minus 64 for Hi)
    int mine;       // own candidates in registers
    int top;        // most own candidates in registers over the wave
    bool anyRest;   // some lane of the wave has candidates past R
    float4 r[2 * R];
    __device__ __forceinline__ LaneCandCheck(const float4* list, uint64_t ownLo, uint64_t ownHi) : c(list)
    {
        mine = 0;
#pragma unroll
        for(int k = 0; k < R; ++k)
            {
                const bool lo    = ownLo != 0;
                const uint64_t w = lo ? ownLo : ownHi;
                const int idx    = w ? __ffsll((unsigned long long)w) - 1 + (lo ? 0 : 64) : 0;
                r[2 * k]         = w ? c[2 * idx] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                r[2 * k + 1]     = w ? c[2 * idx + 1] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                mine += w ? 1 : 0;
                if(lo)
                    ownLo &= ownLo - 1;
                else
                    ownHi &= ownHi - 1;
            }
        restLo  = ownLo;
        restHi  = ownHi;
        top     = (int)wave_maxf((float)mine);
        anyRest = __any((restLo | restHi) != 0);
    }
    __device__ __forceinline__ bool operator()(float ax, float ay, float az, float bx, float by, float bz) const
    {
        // the positions are finite here, so fminf / fmaxf give the
        // reference's box (up to the sign of a zero, which no face gap's
        // comparison with 0 sees)
        const float mnx = fminf(ax, bx), mxx = fmaxf(ax, bx);
        const float mny = fminf(ay, by), mxy = fmaxf(ay, by);
        const float mnz = fminf(az, bz), mxz = fmaxf(az, bz);
        bool hit = false;
#pragma unroll
        for(int k = 0; k < R; ++k)
            {
                if(k >= top) break;  // wave-uniform
                hit |= (k < mine) & !separated_gap(mnx, mny, mnz, mxx, mxy, mxz, r[2 * k], r[2 * k + 1]);
            }
        if(anyRest)
            {
                for(uint64_t m = restLo; m; m &= m - 1)
                    {
                        const int idx = __ffsll((unsigned long long)m) - 1;
                        hit |= !separated_gap(mnx, mny, mnz, mxx, mxy, mxz, c[2 * idx], c[2 * idx + 1]);
                    }
                for(uint64_t m = restHi; m; m &= m - 1)
                    {
                        const int idx = __ffsll((unsigned long long)m) + 63;
                        hit |= !separated_gap(mnx, mny, mnz, mxx, mxy, mxz, c[2 * idx], c[2 * idx + 1]);
                    }
            }
        return !hit;
    }
};

// bitwise ORs: every comparison is evaluated (no short-circuit branches)
__device__ __forceinline__ bool ws_out(const kgmt_config& c, float x, float y, float z)
{
    return (x < c.ws_min) | (x > c.ws_max) | (y < c.ws_min) | (y > c.ws_max) | (z < c.ws_min) | (z > c.ws_max);
}

// -------------------------------------------------------------------------
// Dynamics. s[] holds the state on entry and the final (break-point) state on
// exit; returns validity of the whole edge.
// -------------------------------------------------------------------------
template <int MODEL>
struct ModelDims;
template <>
struct ModelDims<KGMT_MODEL_DOUBLE_INTEGRATOR>
{
    static constexpr int S = 6, C = 3;
};
template <>
struct ModelDims<KGMT_MODEL_DUBINS_AIRPLANE>
{
    static constexpr int S = 6, C = 3;
};
template <>
struct ModelDims<KGMT_MODEL_QUADCOPTER>
{
    static constexpr int S = 12, C = 4;
};
template <>
struct ModelDims<KGMT_MODEL_DUBINS_5D>
{
    static constexpr int S = 5, C = 3;
};
template <>
struct ModelDims<KGMT_MODEL_UNICYCLE>
{
    static constexpr int S = 4, C = 2;
};

// z of a state's position: the planar unicycle moves in z = 0
template <int MODEL>
__device__ __forceinline__ float pos_z(const float* s)
{
    if constexpr(MODEL == KGMT_MODEL_UNICYCLE)
        return 0.0f;
    else
        return s[2];
}

// statePropagator.cu:13-71: Euler steps of the planar unicycle, workspace
// check on x y, then the step's segment (z = 0; planar boxes span z [-1, 1],
// so the 3D test gives the planar verdict). tan(steering) = sin / cos of
// sincos_k.
template <class CHK>
__device__ __forceinline__ bool prop_unicycle(const QueryDesc& qd, float* s, const float* u, int steps, CHK& chk)
{
    const kgmt_config& c = qd.cfg;
    const float h        = c.dt;
    const float a = u[0], steer = u[1];
    float ss, cs;
    sincos_k(steer, ss, cs);
    const float tn = ss / cs;
    float x = s[0], y = s[1], theta = s[2], v = s[3];
    bool valid = true;
    for(int i = 0; i < steps; ++i)
        {
            const float px = x, py = y;
            float st, ct;
            sincos_k(theta, st, ct);
            x     = x + (v * ct) * h;
            y     = y + (v * st) * h;
            theta = theta + ((v / c.uni_length) * tn) * h;
            v     = v + a * h;
            // workspace then segment, both ending the edge: one branch
            const bool free = chk(px, py, 0.0f, x, y, 0.0f);
            if((x < c.ws_min) | (x > c.ws_max) | (y < c.ws_min) | (y > c.ws_max) | !free)
                {
                    valid = false;
                    break;
                }
        }
    s[0] = x, s[1] = y, s[2] = theta, s[3] = v;
    return valid;
}

// statePropagator.cu:76-142
template <class CHK>
__device__ __forceinline__ bool prop_di(const QueryDesc& qd, float* s, const float* u, int steps, CHK& chk)
{
    const kgmt_config& c = qd.cfg;
    const float h        = c.dt;
    const float ax = u[0], ay = u[1], az = u[2];
    float x = s[0], y = s[1], z = s[2], vx = s[3], vy = s[4], vz = s[5];
    const bool exact = c.integrator == KGMT_DI_EXACT;
    bool valid       = true;
    for(int i = 0; i < steps; ++i)
        {
            const float px = x, py = y, pz = z;
            if(exact)
                {
                    x  = x + (vx * h + ((ax * h) * h) / 2.0f);
                    y  = y + (vy * h + ((ay * h) * h) / 2.0f);
                    z  = z + (vz * h + ((az * h) * h) / 2.0f);
                    vx = vx + ax * h;
                    vy = vy + ay * h;
                    vz = vz + az * h;
                }
            else
                {
                    x  = x + (vx + (vx + 2.0f * (vx + ax * h / 2.0f) + (vx + ax * h))) * h / 6.0f;
                    y  = y + (vy + (vy + 2.0f * (vy + ay * h / 2.0f) + (vy + ay * h))) * h / 6.0f;
                    z  = z + (vz + (vz + 2.0f * (vz + az * h / 2.0f) + (vz + az * h))) * h / 6.0f;
                    vx = vx + (ax + 2.0f * ax + 2.0f * ax + ax) * h / 6.0f;
                    vy = vy + (ay + 2.0f * ay + 2.0f * ay + ay) * h / 6.0f;
                    vz = vz + (az + 2.0f * az + 2.0f * az + az) * h / 6.0f;
                }
            // velocity, workspace, segment: each ends the edge, so one branch
            const bool free = chk(px, py, pz, x, y, z);
            if((vx < c.v_min) | (vx > c.v_max) | (vy < c.v_min) | (vy > c.v_max) | (vz < c.v_min) | (vz > c.v_max) |
               ws_out(c, x, y, z) | !free)
                {
                    valid = false;
                    break;
                }
        }
    s[0] = x, s[1] = y, s[2] = z, s[3] = vx, s[4] = vy, s[5] = vz;
    return valid;
}

// statePropagator.cu:147-241; FIVE_D: speed is control u[2], no acceleration.
// Within an edge yaw and pitch advance by the constant h*yr and h*pr per step,
// so the sin/cos of a step's mid (+hh*rate) and end (+h*rate) angles are its
// start values rotated by those increments (angle addition); the end values
// start the next step. sincos_k is evaluated once per angle and increment per
// edge instead of four times per step (the reference calls sin/cos of every
// stage angle).
struct Rot  // sin/cos of an angle increment
{
    float s, c;
};
// angle addition with one fused multiply-add per component (the oracle
// evaluates the same fused operations)
__device__ __forceinline__ void rotate(float s0, float c0, Rot r, float& s, float& c)
{
    s = __builtin_fmaf(s0, r.c, c0 * r.s);
    c = __builtin_fmaf(c0, r.c, -(s0 * r.s));
}

template <bool FIVE_D, class CHK>
__device__ __forceinline__ bool prop_dubins(const QueryDesc& qd, float* s, const float* u, int steps, CHK& chk)
{
    const kgmt_config& c = qd.cfg;
    const float h        = c.dt;
    const float hh       = 0.5f * h;
    const float yr = u[0], pr = u[1];
    const float a = FIVE_D ? 0.0f : u[2];
    float x = s[0], y = s[1], z = s[2], yaw = s[3], pitch = s[4];
    float v    = FIVE_D ? u[2] : s[5];
    bool valid = true;
    float sp0, cp0, sy0, cy0;
    sincos_k(pitch, sp0, cp0);
    sincos_k(yaw, sy0, cy0);
    Rot pH, pE, yH, yE;
    sincos_k(hh * pr, pH.s, pH.c);
    sincos_k(h * pr, pE.s, pE.c);
    sincos_k(hh * yr, yH.s, yH.c);
    sincos_k(h * yr, yE.s, yE.c);
    // FIVE_D: the acceleration is 0, so fma(hh, 0, v), fma(h, 0, v) and the
    // speed after every step are v + 0 (dt is positive and finite: the
    // products are +0), the same bits computed once; only the first step's
    // v * cos * cos terms see v itself
    const float vn = v + 0.0f;
    for(int i = 0; i < steps; ++i)
        {
            const float px = x, py = y, pz = z;
            // RK4 with fused multiply-adds at fixed places (the oracle's
            // prop_dubins evaluates the same operations)
            const float vm = FIVE_D ? vn : __builtin_fmaf(hh, a, v);
            const float ve = FIVE_D ? vn : __builtin_fmaf(h, a, v);
            float spm, cpm, sym, cym, spe, cpe, sye, cye;
            rotate(sp0, cp0, pH, spm, cpm);
            rotate(sy0, cy0, yH, sym, cym);
            rotate(sp0, cp0, pE, spe, cpe);
            rotate(sy0, cy0, yE, sye, cye);
            const float mx = vm * cpm * cym;
            const float my = vm * cpm * sym;
            const float mz = vm * spm;
            const float h6 = h / 6.0f;
            x = __builtin_fmaf(h6, __builtin_fmaf(ve * cpe, cye, __builtin_fmaf(2.0f, mx + mx, v * cp0 * cy0)), x);
            y = __builtin_fmaf(h6, __builtin_fmaf(ve * cpe, sye, __builtin_fmaf(2.0f, my + my, v * cp0 * sy0)), y);
            z = __builtin_fmaf(h6, __builtin_fmaf(ve, spe, __builtin_fmaf(2.0f, mz + mz, v * sp0)), z);
            yaw   = __builtin_fmaf(h, yr, yaw);
            pitch = __builtin_fmaf(h, pr, pitch);
            v     = FIVE_D ? vn : __builtin_fmaf(h6, __builtin_fmaf(2.0f, a + a, a) + a, v);
            sp0 = spe, cp0 = cpe, sy0 = sye, cy0 = cye;
            // the reference's checks in its order (speed, pitch, workspace,
            // segment) all end the edge as invalid at this step, so they are
            // combined; the segment test runs for every lane (one branch)
            const bool free = chk(px, py, pz, x, y, z);
            if((!FIVE_D & ((v < c.v_min) | (v > c.v_max))) | (pitch < c.pitch_min) | (pitch > c.pitch_max) |
               ws_out(c, x, y, z) | !free)
                {
                    valid = false;
                    break;
                }
        }
    s[0] = x, s[1] = y, s[2] = z, s[3] = yaw, s[4] = pitch;
    if(!FIVE_D) s[5] = v;
    return valid;
}

// ode, statePropagator.cu:325-370
__device__ __forceinline__ void quad_ode(const QueryDesc& qd, const float* st, const float* uc, float* xd)
{
    const kgmt_config& c = qd.cfg;
    const float phi = st[3], theta = st[4], psi = st[5];
    const float u = st[6], v = st[7], w = st[8], p = st[9], q = st[10], r = st[11];
    float sph, cph, sth, cth, sps, cps;
    sincos_k(phi, sph, cph);
    sincos_k(theta, sth, cth);
    sincos_k(psi, sps, cps);
    const float tth = sth / cth;
    const float g   = c.gravity;
    xd[0]           = cth * cps * u + (sph * sth * cps - cph * sps) * v + (cph * sth * cps + sph * sps) * w;
    xd[1]           = cth * sps * u + (sph * sth * sps + cph * cps) * v + (cph * sth * sps - sph * cps) * w;
    xd[2]           = -sth * u + sph * cth * v + cph * cth * w;
    xd[3]           = p + (q * sph + r * cph) * tth;
    xd[4]           = q * cph - r * sph;
    xd[5]           = (q * sph + r * cph) / cth;
    const float XYZ = -c.quad_nu * __builtin_sqrtf(u * u + v * v + w * w);
    xd[6]           = (r * v - q * w) - g * sth + qd.qminv * XYZ * u;
    xd[7]           = (p * w - r * u) + g * cth * sph + qd.qminv * XYZ * v;
    xd[8]           = (q * u - p * v) + g * cth * cph + qd.qminv * XYZ * w + qd.qminv * uc[0];
    const float LMN = -c.quad_mu * __builtin_sqrtf(p * p + q * q + r * r);
    xd[9]           = qd.qax * q * r + qd.qbx * LMN * p + qd.qbx * uc[1];
    xd[10]          = qd.qay * p * r + qd.qby * LMN * q + qd.qby * uc[2];
    xd[11]          = qd.qaz * p * q + qd.qbz * LMN * r + qd.qbz * uc[3];
}

// statePropagator.cu:246-323
template <class CHK>
__device__ __forceinline__ bool prop_quad(const QueryDesc& qd, float* s, const float* uc, int steps, CHK& chk)
{
    const kgmt_config& c = qd.cfg;
    const float h = c.dt, hh = 0.5f * h, h6 = h / 6.0f;
    bool valid = true;
    for(int i = 0; i < steps; ++i)
        {
            const float px = s[0], py = s[1], pz = s[2];
            float k1[12], k2[12], k3[12], k4[12], st[12];
            quad_ode(qd, s, uc, k1);
#pragma unroll
            for(int j = 0; j < 12; ++j) st[j] = s[j] + hh * k1[j];
            quad_ode(qd, st, uc, k2);
#pragma unroll
            for(int j = 0; j < 12; ++j) st[j] = s[j] + hh * k2[j];
            quad_ode(qd, st, uc, k3);
#pragma unroll
            for(int j = 0; j < 12; ++j) st[j] = s[j] + h * k3[j];
            quad_ode(qd, st, uc, k4);
#pragma unroll
            for(int j = 0; j < 12; ++j) s[j] = s[j] + h6 * (k1[j] + 2.0f * k2[j] + 2.0f * k3[j] + k4[j]);
            // body velocity, workspace, segment: each ends the edge, so one branch
            const bool free = chk(px, py, pz, s[0], s[1], s[2]);
            if((s[6] < c.v_min) | (s[6] > c.v_max) | (s[7] < c.v_min) | (s[7] > c.v_max) | (s[8] < c.v_min) |
               (s[8] > c.v_max) | ws_out(c, s[0], s[1], s[2]) | !free)
                {
                    valid = false;
                    break;
                }
        }
    return valid;
}

template <int MODEL, class CHK>
__device__ __forceinline__ bool propagate_with(const QueryDesc& qd, float* s, const float* u, int steps, CHK& chk)
{
    if constexpr(MODEL == KGMT_MODEL_DOUBLE_INTEGRATOR)
        return prop_di(qd, s, u, steps, chk);
    else if constexpr(MODEL == KGMT_MODEL_DUBINS_AIRPLANE)
        return prop_dubins<false>(qd, s, u, steps, chk);
    else if constexpr(MODEL == KGMT_MODEL_DUBINS_5D)
        return prop_dubins<true>(qd, s, u, steps, chk);
    else if constexpr(MODEL == KGMT_MODEL_UNICYCLE)
        return prop_unicycle(qd, s, u, steps, chk);
    else
        return prop_quad(qd, s, u, steps, chk);
}

// Every box of the scene tested at every step (statePropagator.cu order:
// integrate, bound checks, then the AABB test of the step's segment).
template <int MODEL, int K = 1>
__device__ __forceinline__ bool propagate(const QueryDesc& qd, float* s, const float* u, int steps,
                                          const Scene& sc)
{
    FullCheck<K> chk{sc};
    return propagate_with<MODEL>(qd, s, u, steps, chk);
}

constexpr int kCand = 128;  // candidate boxes per wave of the two-phase check (two 64-bit lane masks)

// Two-phase edge check with the same verdict and final state as propagate():
// 1. integrate without boxes, recording the box T spanned by every segment the
//    real run would test (the run stops at the same bound check);
// 2. the wave's candidates: boxes overlapping some lane's T (a box that meets
//    a segment overlaps the segment's box, which lies in T; group bounds
//    contain their boxes), gathered into this wave's LDS list `cand`;
// 3. integrate again (same operations, same bits) testing only candidates;
//    with no candidate at all the dry run's result stands.
// When a lane's positions are not finite, or the candidates exceed kCand, the
// wave tests every box instead. Wave-uniform decisions only; call with the
// lanes that propagate a child active.
// Interval helpers for the Dubins edge bound.
__device__ __forceinline__ void imul(float alo, float ahi, float blo, float bhi, float& lo, float& hi)
{
    const float p0 = alo * blo, p1 = alo * bhi, p2 = ahi * blo, p3 = ahi * bhi;
    lo = fminf(fminf(p0, p1), fminf(p2, p3));
    hi = fmaxf(fmaxf(p0, p1), fmaxf(p2, p3));
}

// Ranges of the components of (s0, c0) rotated by every angle d in [da, db]
// (da <= 0 <= db): (s0 cos d + c0 sin d, c0 cos d - s0 sin d), with
// cos d in [1 - D^2 / 2, 1] (D = max(-da, db)) and sin d in [da, db]
// (|sin d| <= |d|, with the sign of d while |d| < pi). Past D = 1 the ranges
// are the whole [-1, 1]. Every bound is widened by 1e-6 (a few ulps of
// rounding) and the rotated vector's length is 1 to within that, so the
// ranges hold the sin/cos pairs that start at (s0, c0) and turn by d. A
// non-finite s0, c0, da or db makes the ranges NaN (the clamps keep NaN).
__device__ __forceinline__ void rot_range(float s0, float c0, float da, float db, float& slo, float& shi, float& clo,
                                          float& chi)
{
    constexpr float e = 1e-6f, one = 1.0f + 1e-6f;
    const float z     = 0.0f * (s0 + c0) + 0.0f * (da + db);  // 0, or NaN
    const float D     = fmaxf(-da, db);
    if(D > 1.0f)
        {
            slo = clo = z - one, shi = chi = z + one;
            return;
        }
    const float k  = 1.0f - 0.5f * D * D;
    const float sk = s0 * k, ck = c0 * k;
    const float b0 = c0 * da, b1 = c0 * db, d0 = -(s0 * da), d1 = -(s0 * db);
    auto lower = [&](float x) { return x < -one ? -one : x; };
    auto upper = [&](float x) { return x > one ? one : x; };
    slo = lower(z + fminf(sk, s0) + fminf(b0, b1) - e), shi = upper(z + fmaxf(sk, s0) + fmaxf(b0, b1) + e);
    clo = lower(z + fminf(ck, c0) + fminf(d0, d1) - e), chi = upper(z + fmaxf(ck, c0) + fmaxf(d0, d1) + e);
}

// Box containing every position of a Dubins edge (statePropagator.cu:147-241)
// without integrating it. Within the edge the angles move linearly (yaw +
// h yr and pitch + h pr per step) and the speed by h a per step, so every RK4
// stage samples the velocity v (cos p cos y, cos p sin y, sin p) with angles
// and speed inside the edge's ranges; a step moves the position by h times a
// weighted mean (weights 1/6, 4/6, 1/6) of three samples, so after k steps
// the offset lies in k h [lo, hi] of the velocity's interval bound. The box
// is widened by a margin that covers float rounding of the steps.
template <bool FIVE_D>
__device__ __forceinline__ void dubins_edge_box(const QueryDesc& qd, const float* s, const float* u, int steps,
                                                float* lo, float* hi, bool& ok)
{
    const float h = qd.cfg.dt, T = h * (float)steps;
    const float yr = u[0], pr = u[1];
    const float a  = FIVE_D ? 0.0f : u[2];
    const float v0 = FIVE_D ? u[2] : s[5];
    const float dp = pr * T, dy = yr * T, dv = a * T;
    const float angPad = 1e-5f * (1.0f + fabsf(s[3]) + fabsf(s[4]) + fabsf(dp) + fabsf(dy));
    const float vlo = v0 + fminf(0.0f, dv), vhi = v0 + fmaxf(0.0f, dv);
    // the start angles' sin/cos as the integration evaluates them, turned by
    // the edge's angle ranges
    float sp0, cp0, sy0, cy0;
    sincos_k(s[4], sp0, cp0);
    sincos_k(s[3], sy0, cy0);
    float spl, sph, cpl, cph, syl, syh, cyl, cyh;
    rot_range(sp0, cp0, fminf(0.0f, dp) - angPad, fmaxf(0.0f, dp) + angPad, spl, sph, cpl, cph);
    rot_range(sy0, cy0, fminf(0.0f, dy) - angPad, fmaxf(0.0f, dy) + angPad, syl, syh, cyl, cyh);
    float dl[3], dh[3], tl, th;
    imul(cpl, cph, cyl, cyh, tl, th);
    imul(vlo, vhi, tl, th, dl[0], dh[0]);
    imul(cpl, cph, syl, syh, tl, th);
    imul(vlo, vhi, tl, th, dl[1], dh[1]);
    imul(vlo, vhi, spl, sph, dl[2], dh[2]);
    const float vmax = fmaxf(fabsf(vlo), fabsf(vhi));
    // a NaN velocity bound (non-finite angles or speed) would be dropped by
    // fminf / fmaxf below: such an edge has no box
    ok = (dl[0] == dl[0]) & (dh[0] == dh[0]) & (dl[1] == dl[1]) & (dh[1] == dh[1]) & (dl[2] == dl[2]) &
         (dh[2] == dh[2]);
#pragma unroll
    for(int d = 0; d < 3; ++d)
        {
            // rounding of the steps, and of the rotated sin/cos (a few ulps
            // per step)
            const float pad = 1e-4f * (1.0f + fabsf(s[d]) + T * vmax) + 1e-6f * (float)steps * T * vmax;
            lo[d]           = s[d] + fminf(0.0f, T * dl[d]) - pad;
            hi[d]           = s[d] + fmaxf(0.0f, T * dh[d]) + pad;
        }
}

template <int CTRL>
__device__ __forceinline__ float dpp_minf(float v)
{
    return fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false)));
}
template <int CTRL>
__device__ __forceinline__ float dpp_maxf(float v)
{
    return fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false)));
}

// float bits of lane `l` (readlane moves 32-bit integers)
__device__ __forceinline__ float readlane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// min / max of v over the 64 lanes (all active), wave-uniform: DPP within each
// row of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror), then the rows' lane 0
__device__ __forceinline__ float wave_minf(float v)
{
    v = dpp_minf<0x140>(dpp_minf<0x141>(dpp_minf<0x4E>(dpp_minf<0xB1>(v))));
    return fminf(fminf(readlane_f(v, 0), readlane_f(v, 16)), fminf(readlane_f(v, 32), readlane_f(v, 48)));
}
__device__ __forceinline__ float wave_maxf(float v)
{
    v = dpp_maxf<0x140>(dpp_maxf<0x141>(dpp_maxf<0x4E>(dpp_maxf<0xB1>(v))));
    return fmaxf(fmaxf(readlane_f(v, 0), readlane_f(v, 16)), fmaxf(readlane_f(v, 32), readlane_f(v, 48)));
}

// What propagate_two_phase did for one lane: the check that decided the
// verdict (KGMT_CHECK_* of kgmt.h) and the lane's edge box (the box holding
// every segment the edge tests). Only the stage-level entry asks for it; the
// planner passes no trace and the recording folds away.
struct TwoPhaseTrace
{
    int path;
    float lo[3], hi[3];
};

// Called by every lane of the wave; lanes without a child (has == false)
// pass steps = 0 and take part in the wave-wide steps with an empty box.
template <int MODEL>
__device__ __forceinline__ bool propagate_two_phase(const QueryDesc& qd, float* s, const float* u, int steps,
                                                    const Scene& sc, float4* cand, bool has,
                                                    TwoPhaseTrace* trace = nullptr)
{
    constexpr int S      = ModelDims<MODEL>::S;
    constexpr bool BOUND = MODEL == KGMT_MODEL_DUBINS_5D || MODEL == KGMT_MODEL_DUBINS_AIRPLANE;
    DryRun dry;
    float s1[S];
    bool dryOk = false;
    if constexpr(BOUND)
        {
            // the edge's box from the bound, without a dry run
            bool ok;
            dubins_edge_box<MODEL == KGMT_MODEL_DUBINS_5D>(qd, s, u, steps, dry.lo, dry.hi, ok);
            dry.finite = ok && __builtin_isfinite(dry.lo[0]) && __builtin_isfinite(dry.lo[1]) &&
                         __builtin_isfinite(dry.lo[2]) && __builtin_isfinite(dry.hi[0]) &&
                         __builtin_isfinite(dry.hi[1]) && __builtin_isfinite(dry.hi[2]);
        }
    else
        {
            dry.start(s[0], s[1], pos_z<MODEL>(s));
#pragma unroll
            for(int d = 0; d < S; ++d) s1[d] = s[d];
            dryOk = propagate_with<MODEL>(qd, s1, u, steps, dry);
        }
    if(trace)
        {
#pragma unroll
            for(int d = 0; d < 3; ++d) trace->lo[d] = dry.lo[d], trace->hi[d] = dry.hi[d];
        }
    if(!has)
        {
            // an empty box: no candidate, no effect on the wave's union
            dry.lo[0] = dry.lo[1] = dry.lo[2] = INFINITY;
            dry.hi[0] = dry.hi[1] = dry.hi[2] = -INFINITY;
            dry.finite = true;
        }
    int n        = 0;
    uint64_t ownLo = 0, ownHi = 0;  // list entries 0..63 / 64..127 overlapping this lane's box
    if(!__any(!dry.finite))
        {
            // candidates = boxes that some lane's box overlaps, in box order.
            // Lane j first tests box b0 + j against the union of the lanes'
            // boxes (a box meeting a lane's box meets the union); the boxes
            // that pass are then tested against every lane's box.
            float ulo[3], uhi[3];
#pragma unroll
            for(int d = 0; d < 3; ++d) ulo[d] = wave_minf(dry.lo[d]), uhi[d] = wave_maxf(dry.hi[d]);
            const float4* blo = sc.box_lo();
            const float4* bhi = sc.box_hi();
            const int nBox    = sc.nGrp * (int)kGroup;
            for(int b0 = 0; b0 < nBox && n <= kCand; b0 += 64)
                {
                    const int b = b0 + (int)__lane_id();
                    bool pass   = false;
                    if(b < nBox)
                        pass = !separated(ulo[0], ulo[1], ulo[2], uhi[0], uhi[1], uhi[2], blo[b], bhi[b]);
                    uint64_t m = __ballot(pass);
                    // a box that some lane's box meets joins the list (same
                    // value in every lane; n past kCand ends the gather)
                    auto take = [&](const float4& lo, const float4& hi, bool c) {
                        if(n <= kCand && __any(c))
                            {
                                if(n < kCand)
                                    {
                                        cand[2 * n] = lo, cand[2 * n + 1] = hi;
                                        if(c)
                                            {
                                                if(n < 64)
                                                    ownLo |= 1ull << n;
                                                else
                                                    ownHi |= 1ull << (n - 64);
                                            }
                                    }
                                ++n;
                            }
                    };
                    // many passing boxes (children of many parents): four per
                    // round, their reads and tests independent of each other,
                    // taken in box order
                    while(__popcll(m) >= 4 && n <= kCand)
                        {
                            float4 lo[4], hi[4];
#pragma unroll
                            for(int q = 0; q < 4; ++q)
                                {
                                    const int k = __ffsll((unsigned long long)m) - 1;
                                    m &= m - 1;
                                    lo[q] = blo[b0 + k], hi[q] = bhi[b0 + k];
                                }
                            // the face gaps first (VALU only, one vote for all
                            // four): a gap says separated only where
                            // separated() does, so four boxes that no lane's
                            // box meets by the gaps join nothing
                            bool near = false;
#pragma unroll
                            for(int q = 0; q < 4; ++q)
                                near |= !separated_gap(dry.lo[0], dry.lo[1], dry.lo[2], dry.hi[0], dry.hi[1], dry.hi[2],
                                                       lo[q], hi[q]);
                            if(!__any(near)) continue;
                            bool c[4];
#pragma unroll
                            for(int q = 0; q < 4; ++q)
                                c[q] = !separated(dry.lo[0], dry.lo[1], dry.lo[2], dry.hi[0], dry.hi[1], dry.hi[2], lo[q],
                                                  hi[q]);
#pragma unroll
                            for(int q = 0; q < 4; ++q) take(lo[q], hi[q], c[q]);
                        }
                    while(m && n <= kCand)
                        {
                            const int k = __ffsll((unsigned long long)m) - 1;
                            m &= m - 1;
                            const float4 lo = blo[b0 + k], hi = bhi[b0 + k];
                            take(lo, hi,
                                 !separated(dry.lo[0], dry.lo[1], dry.lo[2], dry.hi[0], dry.hi[1], dry.hi[2], lo, hi));
                        }
                }
        }
    else
        n = kCand + 1;
    if(n > kCand)
        {
            if(trace) trace->path = KGMT_CHECK_SCENE;
            FullCheck<1> chk{sc};
            return propagate_with<MODEL>(qd, s, u, steps, chk);
        }
    if(!BOUND && n == 0)
        {
            // no box near any lane: the dry run is the run
            if(trace) trace->path = KGMT_CHECK_DRY_RUN;
#pragma unroll
            for(int d = 0; d < S; ++d) s[d] = s1[d];
            return dryOk;
        }
    if(trace) trace->path = KGMT_CHECK_CANDIDATES;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    LaneCandCheck chk(cand, ownLo, ownHi);
    return propagate_with<MODEL>(qd, s, u, steps, chk);
}

// -------------------------------------------------------------------------
// Regions (workspace-only grid, last axis fastest). Graph.cu:99-153, 155-205; the
// region minimum corner is ws_min + index * W_R1_SIZE (Graph.cu:67-70).
// -------------------------------------------------------------------------
__device__ __forceinline__ int grid_index(float f, int L)
{
    if(!(f >= 0.0f)) return 0;
    if(f >= (float)L) return L - 1;
    const int i = (int)f;
    return i >= L ? L - 1 : i;
}

// Planar queries (wdim 2) ignore z: regions ix * L + iy.
__device__ __forceinline__ int region_of(const QueryDesc& qd, float x, float y, float z, int& ix, int& iy, int& iz)
{
    const kgmt_config& c = qd.cfg;
    const int L          = c.r1_len;
    const float span     = c.ws_max - c.ws_min;
    const float fL       = (float)L;
    // a / span, or the same correctly rounded value as a product with the
    // exact reciprocal when the span is a power of two (a uniform branch)
    const float inv = qd.invSpan;
    const bool mul  = inv != 0.0f;
    const float ay = fL * (y - c.ws_min), ax = fL * (x - c.ws_min);
    iy = grid_index(mul ? ay * inv : ay / span, L);
    ix = grid_index(mul ? ax * inv : ax / span, L);
    if(qd.wdim == 2)
        {
            iz = 0;
            return ix * L + iy;
        }
    const float az = fL * (z - c.ws_min);
    iz             = grid_index(mul ? az * inv : az / span, L);
    return (ix * L + iy) * L + iz;
}

__device__ __forceinline__ int subregion_of(const QueryDesc& qd, float x, float y, float z, int r1, int ix, int iy,
                                            int iz)
{
    const kgmt_config& c = qd.cfg;
    const int L2         = c.r2_len;
    const float fL2      = (float)L2;
    const float size     = qd.r1size;
    const float inv  = qd.invR1;  // as in region_of
    const bool mul   = inv != 0.0f;
    const float mx = c.ws_min + (float)ix * size, my = c.ws_min + (float)iy * size;
    const float ay = fL2 * (y - my), ax = fL2 * (x - mx);
    const int ly   = grid_index(mul ? ay * inv : ay / size, L2);
    const int lx   = grid_index(mul ? ax * inv : ax / size, L2);
    if(qd.wdim == 2) return r1 * qd.P + lx * L2 + ly;
    const float mz = c.ws_min + (float)iz * size;
    const float az = fL2 * (z - mz);
    const int lz   = grid_index(mul ? az * inv : az / size, L2);
    return r1 * qd.P + (lx * L2 + ly) * L2 + lz;
}

// helper.cuh:138-152: Euclidean distance over the W_DIM position axes
__device__ __forceinline__ float pos_dist(const QueryDesc& qd, float ax, float ay, float az, float bx, float by,
                                          float bz)
{
    const float dx = ax - bx, dy = ay - by;
    if(qd.wdim == 2) return __builtin_sqrtf(dx * dx + dy * dy);
    const float dz = az - bz;
    return __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ float goal_dist(const QueryDesc& qd, float x, float y, float z)
{
    return pos_dist(qd, x, y, z, qd.goal[0], qd.goal[1], qd.goal[2]);
}

// -------------------------------------------------------------------------
// Wave / block helpers (wave64).
// -------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id()
{
    return __lane_id();
}

// Orders one wave's LDS accesses across its lanes without a workgroup
// barrier: the LDS performs a wave's operations in issue order, and the
// fences and wave barrier keep the compiler from moving accesses across.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v of the lane `CTRL` selects (DPP), 0 where the source lies outside the row
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

// inclusive scan of v over each row of 16 lanes (row_shr 1, 2, 4, 8)
__device__ __forceinline__ uint32_t row_incl_scan(uint32_t v)
{
    v += dpp_u<0x111>(v);
    v += dpp_u<0x112>(v);
    v += dpp_u<0x114>(v);
    v += dpp_u<0x118>(v);
    return v;
}

// inclusive scan of v over each half wave (lanes 0-31, 32-63): row scans,
// then the first row's total added to the second row of each half
__device__ __forceinline__ uint32_t half_incl_scan(uint32_t v)
{
    v                   = row_incl_scan(v);
    const uint32_t r15  = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
    const uint32_t r47  = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
    const uint32_t row  = lane_id() >> 4;
    return v + (row == 1u ? r15 : row == 3u ? r47 : 0u);
}

// inclusive scan of v over the 64 lanes of a wave: row scans, then the row
// totals (lanes 15, 31, 47) added to the rows after them
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    v                   = row_incl_scan(v);
    const uint32_t r0   = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
    const uint32_t r1   = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
    const uint32_t r2   = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
    const uint32_t row  = lane_id() >> 4;
    const uint32_t r01  = r0 + r1;
    return v + (row == 0 ? 0u : row == 1 ? r0 : row == 2 ? r01 : r01 + r2);
}

// Exclusive scan of one value per thread over a block of NT threads
// (NT/64 <= 16 waves). Returns the exclusive prefix; *total receives the
// block sum. Every wave scans the wave totals itself (one row scan), so there
// is one barrier before and one after (sbuf reuse). sbuf: at least NT/64 + 1
// words of LDS. Contains barriers: call uniformly.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sbuf, uint32_t& total)
{
    constexpr int NW = NT / 64;
    static_assert(NW >= 1 && NW <= 16, "block of 64 to 1024 threads");
    const uint32_t incl = wave_incl_scan(v);
    if constexpr(NW == 1)
        {
            total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            __syncthreads();
            return incl - v;
        }
    else
        {
            const uint32_t lane = lane_id();
            const int wave      = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
            if(lane == 63) sbuf[wave] = incl;
            __syncthreads();
            const uint32_t w  = row_incl_scan(lane < (uint32_t)NW ? sbuf[lane] : 0u);
            const uint32_t pw = wave == 0 ? 0u : (uint32_t)__builtin_amdgcn_readlane((int)w, wave - 1);
            total             = (uint32_t)__builtin_amdgcn_readlane((int)w, NW - 1);
            __syncthreads();
            return pw + incl - v;
        }
}

// Exclusive scans of NV values per thread over a block of NT threads (2 to
// 16 waves) with one barrier pair: ex[i] = prefix of v[i], total[i] = block
// sum of v[i]. sbuf: at least NV * NT/64 words of LDS. Call uniformly.
template <int NT, int NV>
__device__ __forceinline__ void block_excl_scan_n(const uint32_t (&v)[NV], uint32_t (&ex)[NV], uint32_t (&total)[NV],
                                                  uint32_t* sbuf)
{
    constexpr int NW = NT / 64;
    static_assert(NW >= 2 && NW <= 16, "block of 128 to 1024 threads");
    const uint32_t lane = lane_id();
    const int wave      = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
    uint32_t incl[NV];
#pragma unroll
    for(int i = 0; i < NV; ++i) incl[i] = wave_incl_scan(v[i]);
    if(lane == 63)
        {
#pragma unroll
            for(int i = 0; i < NV; ++i) sbuf[i * NW + wave] = incl[i];
        }
    __syncthreads();
#pragma unroll
    for(int i = 0; i < NV; ++i)
        {
            const uint32_t w  = row_incl_scan(lane < (uint32_t)NW ? sbuf[i * NW + lane] : 0u);
            const uint32_t pw = wave == 0 ? 0u : (uint32_t)__builtin_amdgcn_readlane((int)w, wave - 1);
            total[i]          = (uint32_t)__builtin_amdgcn_readlane((int)w, NW - 1);
            ex[i]             = pw + incl[i] - v[i];
        }
    __syncthreads();
}

// Exclusive scan (in place) of n words with one block of NT threads,
// returns the sum. Barriers inside: call uniformly.
template <int NT>
__device__ __forceinline__ uint32_t block_scan_array(uint32_t* a, uint32_t n, uint32_t* sbuf)
{
    uint32_t carry = 0;
    for(uint32_t base = 0; base < n; base += NT)
        {
            const uint32_t i = base + threadIdx.x;
            const uint32_t v = i < n ? a[i] : 0u;
            uint32_t tot;
            const uint32_t ex = block_excl_scan<NT>(v, sbuf, tot);
            if(i < n) a[i] = carry + ex;
            carry += tot;
        }
    return carry;
}

__device__ __forceinline__ bool test_bit(const uint32_t* words, uint32_t bit)
{
    return (words[bit >> 5] >> (bit & 31u)) & 1u;
}

// number of set bits in [start, start + len) of a bitset
__device__ __forceinline__ uint32_t count_bits(const uint32_t* words, uint32_t start, uint32_t len)
{
    uint32_t cnt = 0, b = start;
    const uint32_t end = start + len;
    while(b < end)
        {
            const uint32_t w   = b >> 5, off = b & 31u;
            const uint32_t n   = (32u - off) < (end - b) ? (32u - off) : (end - b);
            const uint32_t msk = n == 32u ? 0xffffffffu : (((1u << n) - 1u) << off);
            cnt += __popc(words[w] & msk);
            b += n;
        }
    return cnt;
}

// index of the query owning global job t (prefix array of nq + 1 entries)
__device__ __forceinline__ uint32_t locate(const uint32_t* prefix, uint32_t nq, uint32_t t)
{
    uint32_t lo = 0, hi = nq;  // prefix[lo] <= t < prefix[hi]
    while(hi - lo > 1)
        {
            const uint32_t mid = (lo + hi) >> 1;
            if(prefix[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
    return lo;
}

// one child per thread, the one-pass check against every box of the scene
// (staged in LDS by the caller kernel)
template <int MODEL>
__device__ __forceinline__ void stage_propagate_child(const QueryDesc& qd, const int32_t* __restrict__ parentIdx,
                                                      const float* __restrict__ parentStates, uint32_t nParents,
                                                      const float* __restrict__ controls,
                                                      const int32_t* __restrict__ steps, uint32_t n,
                                                      const float4* sObs, uint32_t nGrp, float* child,
                                                      uint8_t* valid, int32_t* r1o, int32_t* r2o)
{
    constexpr int S = ModelDims<MODEL>::S, C = ModelDims<MODEL>::C, D = S + C + 1;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if(i >= n) return;
    const int32_t p = parentIdx[i];
    if(p < 0 || (uint32_t)p >= nParents)
        {
            // caller index out of range: no state is read; flagged by r1 = r2 = -1
            for(int d = 0; d < D; ++d) child[(size_t)d * n + i] = 0.0f;
            valid[i] = 0;
            r1o[i]   = -1;
            r2o[i]   = -1;
            return;
        }
    float s[S], uc[C];
#pragma unroll
    for(int d = 0; d < S; ++d) s[d] = parentStates[(size_t)d * nParents + p];
#pragma unroll
    for(int k = 0; k < C; ++k) uc[k] = controls[(size_t)k * n + i];
    const int32_t st = steps[i];
    const Scene sc{sObs, (int)nGrp};
    const bool ok    = propagate<MODEL>(qd, s, uc, st, sc);
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
}

}  // namespace kgmt
