// kgmt_engine.cpp — host side of the KGMT expansion path: device pools, the
// per-iteration launch sequence (plain or hipGraph), and the C-ABI of
// include/kgmt.h.
//
// Reference host code replaced (paths relative to the KGMT repository):
//   KGMT::KGMT / Planner::Planner / Graph::Graph  src/planners/KGMT.cu:4-39,
//       src/planners/Planner.cu:4-35, src/graphs/Graph.cu:5-42  -> Engine::Engine
//   KGMT::plan                                    src/planners/KGMT.cu:41-105 -> Engine::run
// The reference reads five sizes back to the host every iteration
// (KGMT.cu:171-172, 180-181, 384, 399); here all sizes stay on the device and
// the host polls one completion word a few iterations behind the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kgmt.h"
#include "kgmt_launch.hpp"

using namespace kgmt;

namespace
{

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do                                                                                           \
        {                                                                                        \
            hipError_t e_ = (expr);                                                              \
            if(e_ != hipSuccess) return fail(KGMT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                        \
    while(0)

int model_dims(int model, int& S, int& C)
{
    switch(model)
        {
            case KGMT_MODEL_DOUBLE_INTEGRATOR:
            case KGMT_MODEL_DUBINS_AIRPLANE: S = 6, C = 3; return 0;
            case KGMT_MODEL_QUADCOPTER: S = 12, C = 4; return 0;
            case KGMT_MODEL_DUBINS_5D: S = 5, C = 3; return 0;
            case KGMT_MODEL_UNICYCLE: S = 4, C = 2; return 0;
            default: return -1;
        }
}

// workspace dimension W_DIM: the unicycle is planar
int model_wdim(int model)
{
    return model == KGMT_MODEL_UNICYCLE ? 2 : 3;
}

int64_t ipow(int64_t b, int e)
{
    int64_t r = 1;
    for(int i = 0; i < e; ++i) r *= b;
    return r;
}

int check_config(const kgmt_config* c)
{
    if(!c) return fail(KGMT_E_ARG, "config is null");
    int S, C;
    if(model_dims(c->model, S, C)) return fail(KGMT_E_ARG, "unknown model " + std::to_string(c->model));
    if(c->capacity < 2 || c->capacity > (1u << 26)) return fail(KGMT_E_ARG, "capacity must be in [2, 2^26]");
    if(c->max_iter < 1 || c->max_iter > 1000000) return fail(KGMT_E_ARG, "max_iter must be in [1, 1e6]");
    if(!(c->dt > 0.0f) || !std::isfinite(c->dt)) return fail(KGMT_E_ARG, "dt must be positive");
    if(c->steps_min < 1 || c->steps_max < c->steps_min || c->steps_max > 100000)
        return fail(KGMT_E_ARG, "steps range must satisfy 1 <= steps_min <= steps_max <= 1e5");
    if(!(c->ws_max > c->ws_min)) return fail(KGMT_E_ARG, "ws_max must exceed ws_min");
    if(c->r1_len < 1 || c->r2_len < 1) return fail(KGMT_E_ARG, "grid lengths must be >= 1");
    if(c->r1_len > 8192 || c->r2_len > 8192) return fail(KGMT_E_ARG, "grid lengths must be <= 8192");
    const int W      = model_wdim(c->model);
    const int64_t R1 = ipow(c->r1_len, W);
    const int64_t P  = ipow(c->r2_len, W);
    if(R1 > 8192) return fail(KGMT_E_ARG, "at most 8192 regions (r1_len <= 20 in 3D, <= 90 planar)");
    if(R1 * P > (1ll << 28)) return fail(KGMT_E_ARG, "too many subregions");
    if(c->children_per_entry < 1 || c->children_per_entry > 1024)
        return fail(KGMT_E_ARG, "children_per_entry must be in [1, 1024]");
    if(c->root_repeat < 1) return fail(KGMT_E_ARG, "root_repeat must be >= 1");
    if(c->repeat_sparse < 1 || c->repeat_sparse > 63 || c->repeat_dense < 1 || c->repeat_dense > 63)
        return fail(KGMT_E_ARG, "repeat counts must be 1..63");
    if(c->model == KGMT_MODEL_QUADCOPTER &&
       (!(c->quad_mass > 0.0f) || !(c->quad_ix > 0.0f) || !(c->quad_iy > 0.0f) || !(c->quad_iz > 0.0f)))
        return fail(KGMT_E_ARG, "quad mass and inertia must be positive");
    if(c->model == KGMT_MODEL_UNICYCLE && !(c->uni_length > 0.0f && std::isfinite(c->uni_length)))
        return fail(KGMT_E_ARG, "uni_length must be positive");
    if(c->reserved != 0) return fail(KGMT_E_ARG, "reserved must be 0");
    for(int i = 0; i < C; ++i)
        if(!(c->ctrl_max[i] >= c->ctrl_min[i])) return fail(KGMT_E_ARG, "control range reversed");
    return KGMT_OK;
}

// 1 / d when d is a power of two whose reciprocal is a normal float (then
// a * (1 / d) and a / d are the same correctly rounded value for every a),
// else 0
float exact_recip(float d)
{
    if(!(d > 0.0f) || !std::isfinite(d)) return 0.0f;
    int e;
    const float m = std::frexp(d, &e);  // d = m 2^e, m in [0.5, 1)
    if(m != 0.5f || e - 1 < -126 || 1 - e > 127) return 0.0f;
    return std::ldexp(1.0f, 1 - e);
}

// descriptor fields that depend only on the configuration
void fill_desc(QueryDesc& qd, const kgmt_config& c)
{
    std::memset(&qd, 0, sizeof(qd));
    qd.cfg = c;
    int S = 0, C = 0;
    model_dims(c.model, S, C);
    qd.S = S, qd.C = C, qd.D = S + C + 1;
    qd.wdim    = model_wdim(c.model);
    qd.R1      = (int32_t)ipow(c.r1_len, qd.wdim);
    qd.P       = (int32_t)ipow(c.r2_len, qd.wdim);
    qd.R2      = qd.R1 * qd.P;
    qd.R2words = (qd.R2 + 31) / 32;
    qd.cap     = c.capacity;
    qd.nChunks = (c.capacity + kChunk - 1) / kChunk;
    // W_R1_SIZE and W_R1_VOL as the float macros of config.h:247-251 evaluate
    // (W_R1_VOL = W_R1_SIZE^W_DIM: an area for a planar workspace)
    qd.r1size = (c.ws_max - c.ws_min) / (float)c.r1_len;
    qd.vol    = qd.wdim == 2 ? qd.r1size * qd.r1size : qd.r1size * qd.r1size * qd.r1size;
    qd.invSpan = exact_recip(c.ws_max - c.ws_min);
    qd.invR1   = exact_recip(qd.r1size);
    qd.invP    = exact_recip((float)qd.P);
    qd.qminv  = 1.0f / c.quad_mass;
    qd.qax = (c.quad_iy - c.quad_iz) / c.quad_ix, qd.qbx = 1.0f / c.quad_ix;
    qd.qay = (c.quad_iz - c.quad_ix) / c.quad_iy, qd.qby = 1.0f / c.quad_iy;
    qd.qaz = (c.quad_ix - c.quad_iy) / c.quad_iz, qd.qbz = 1.0f / c.quad_iz;
}

template <typename T>
int dev_alloc(T*& p, size_t count, std::vector<void*>& owned)
{
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if(hipMalloc(&q, bytes) != hipSuccess)
        return fail(KGMT_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    owned.push_back(q);
    p = static_cast<T*>(q);
    return KGMT_OK;
}

constexpr size_t kHostPathFloats = size_t(1) << 20;  // host-mapped path rows of a single query (4 MiB)
// launch units the host stays ahead of the GPU: one for a plan's first
// kLongAfter iterations (every launch is queued while the one before runs,
// and a plan that finishes leaves one unit behind), two after them (a plan
// that runs long is most often one that runs to the iteration limit: the
// deeper queue absorbs the host's own delays); two while events time the
// launches (a deeper queue keeps the event pairs back to back, so they time
// kernels rather than host gaps)
constexpr uint32_t kLag = 1, kLagLong = 2, kLongAfter = 16, kLagProfiled = 2;

}  // namespace

// groups of a scene of n boxes: whole batches of kGroupBatch groups
uint32_t kgmt::scene_groups(uint32_t n)
{
    const uint32_t per = kGroup * kGroupBatch;
    return (n + per - 1) / per * kGroupBatch;
}

uint32_t kgmt::scene_pairs(uint32_t n)
{
    return scene_groups(n) * (1 + kGroup);
}

// Host array in pinned memory (asynchronous copies to and from the device).
template <typename T>
struct PinnedArray
{
    T* p     = nullptr;
    size_t n = 0;
    PinnedArray() = default;
    PinnedArray(const PinnedArray&) = delete;
    PinnedArray& operator=(const PinnedArray&) = delete;
    ~PinnedArray()
    {
        if(p) (void)hipHostFree(p);
    }
    bool resize(size_t count)
    {
        if(p) (void)hipHostFree(p);
        p = nullptr, n = 0;
        if(hipHostMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess)
            return false;
        std::memset((void*)p, 0, std::max<size_t>(count, 1) * sizeof(T));
        n = count;
        return true;
    }
    T* data() { return p; }
    const T* data() const { return p; }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

// ---------------------------------------------------------------------------
// Engine: device pools and the iteration loop for a table of queries
// ---------------------------------------------------------------------------
struct Engine
{
    uint32_t nq = 0;
    PinnedArray<QueryDesc> hq;
    PinnedArray<QueryState> hs;
    std::vector<void*> owned;
    QueryDesc* dq  = nullptr;
    QueryState* ds = nullptr;
    Pools P{};
    WorkLists W{};
    uint32_t obsCap = 0;     // (lo, hi) pairs the obstacle pool holds
    float4* obsPool = nullptr;
    uint32_t rowsCap = 0;    // obstacle rows the staging buffer holds (host-provided obstacles)
    float* rowsBuf   = nullptr;
    const float* sceneRows = nullptr;  // device rows the scenes of the next run are built from
    bool sceneBuild  = false;
    bool counterDirty = false;  // W.counter may be non-zero (a plan did not complete)
    HostResult* hostRes = nullptr;  // host-mapped result of a single-query plan
    float* hostPath     = nullptr;  // host-mapped path rows of that plan
    uint32_t maxPairs = 0;   // largest per-query scene (pairs) of the current plan
    int model       = -1;    // common model or -1
    uint32_t maxIter = 0;
    hipStream_t stream = nullptr;
    bool ownStream     = false;
    uint32_t* pinned   = nullptr;  // host-mapped words written by the kernels: completion, look-back error,
                                   // iterations closed
    Launch L;
    hipGraph_t graph         = nullptr;
    hipGraphExec_t graphExec = nullptr;
    uint32_t graphObs        = 0xffffffffu;
    Launch graphLaunch;
    std::vector<double> iterMs;
    uint32_t itersRun = 0;
    int device        = 0;
    // profiling: 1 = events around propagate only, 2 = at the boundaries of all
    // 5 kernels of an iteration
    int profiling = 0;
    std::vector<hipEvent_t> profEv;
    kgmt_profile prof{};
    // device copy of query 0's tree in the reference's layout (rows of D
    // floats, parents, costs), written on request after a plan
    float* mirRows    = nullptr;
    int32_t* mirPar   = nullptr;
    float* mirCost    = nullptr;
    bool mirValid     = false;

    ~Engine()
    {
        // a plan may have returned while its trailing launches still drain
        if(stream) (void)hipStreamSynchronize(stream);
        for(auto& e : profEv)
            if(e) (void)hipEventDestroy(e);
        if(graphExec) (void)hipGraphExecDestroy(graphExec);
        if(graph) (void)hipGraphDestroy(graph);
        if(pinned) (void)hipHostFree(pinned);
        if(hostRes) (void)hipHostFree(hostRes);
        if(hostPath) (void)hipHostFree(hostPath);
        for(void* p : owned) (void)hipFree(p);
        if(ownStream && stream) (void)hipStreamDestroy(stream);
    }

    int init(const kgmt_config* cfgs, uint32_t n)
    {
        if(n == 0) return fail(KGMT_E_ARG, "no queries");
        int ndev = 0;
        if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(KGMT_E_NODEVICE, "no HIP device visible");
        HIP_TRY(hipGetDevice(&device));
        nq = n;
        if(!hq.resize(n) || !hs.resize(n)) return fail(KGMT_E_NOMEM, "pinned host allocation failed");
        uint64_t samples = 0, rows = 0, nodes = 0, paths = 0;
        uint32_t chunks = 0, regs = 0, subs = 0;
        model = cfgs[0].model;
        for(uint32_t q = 0; q < n; ++q)
            {
                int rc = check_config(&cfgs[q]);
                if(rc) return rc;
                QueryDesc& qd = hq[q];
                fill_desc(qd, cfgs[q]);
                qd.sampleOff = samples;
                qd.rowOffC   = rows;
                qd.nodeOff   = nodes;
                qd.chunkOff  = chunks;
                qd.regOff    = regs;
                qd.subOff    = subs;
                qd.pathOff   = (uint32_t)paths;
                samples += (uint64_t)qd.D * qd.cap;
                rows += (uint64_t)row_words(qd.D) * qd.cap;
                nodes += ((uint64_t)qd.cap + kNodeAlign - 1) / kNodeAlign * kNodeAlign;
                chunks += qd.nChunks + 1;
                regs += (uint32_t)qd.R1;
                subs += (uint32_t)qd.R2words;
                paths += (uint64_t)(cfgs[q].max_iter + 2) * qd.D;
                if(paths > 0xffffffffull) return fail(KGMT_E_CAPACITY, "path pool too large");
                if(cfgs[q].model != model) model = -1;
                if(qd.wdim != hq[0].wdim)
                    return fail(KGMT_E_ARG, "a batch cannot mix planar (unicycle) and 3D queries");
                maxIter = std::max(maxIter, cfgs[q].max_iter);
            }
        int rc = 0;
        if((rc = dev_alloc(dq, n, owned)) || (rc = dev_alloc(ds, n, owned)) ||
           (rc = dev_alloc(P.treeS, samples, owned)) || (rc = dev_alloc(P.childR, rows, owned)) ||
           (rc = dev_alloc(P.parent, nodes, owned)) || (rc = dev_alloc(P.grand, nodes, owned)) || (rc = dev_alloc(P.cost, nodes, owned)) ||
           (rc = dev_alloc(P.anc, nodes, owned)) ||
           (rc = dev_alloc(P.frontier, nodes, owned)) || (rc = dev_alloc(P.r1node, nodes, owned)) ||
           (rc = dev_alloc(P.r1child, nodes, owned)) ||
           (rc = dev_alloc(P.fnext, nodes, owned)) || (rc = dev_alloc(P.react, nodes, owned)) ||
           (rc = dev_alloc(P.repIdx, nodes, owned)) || (rc = dev_alloc(P.partA, chunks, owned)) ||
           (rc = dev_alloc(P.look, 4 * (size_t)chunks, owned)) || (rc = dev_alloc(P.total, regs, owned)) ||
           (rc = dev_alloc(P.validc, regs, owned)) || (rc = dev_alloc(P.score, regs, owned)) ||
           (rc = dev_alloc(P.actPrev, subs, owned)) || (rc = dev_alloc(P.actNext, subs, owned)) ||
           (rc = dev_alloc(P.path, paths, owned)) || (rc = dev_alloc(W.tileStart, 2 * (size_t)(n + 1), owned)) ||
           (rc = dev_alloc(W.jobStart, 2 * (size_t)(n + 1), owned)) || (rc = dev_alloc(W.active, 2, owned)) ||
           (rc = dev_alloc(W.counter, 1, owned)) || (rc = dev_alloc(W.claim, 2, owned)) ||
           (rc = dev_alloc(W.closedDev, 1, owned)))
            return rc;
        W.nq   = n;
        W.tile = L.propBlock / L.lanes;
        HIP_TRY(hipMemset(W.counter, 0, sizeof(uint32_t)));
        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        ownStream = true;
        HIP_TRY(hipHostMalloc((void**)&pinned, 3 * sizeof(uint32_t), hipHostMallocMapped));
        pinned[0] = pinned[1] = pinned[2] = 0;
        HIP_TRY(hipHostGetDevicePointer((void**)&W.allDone, pinned, 0));
        W.error  = W.allDone + 1;
        W.closed = W.allDone + 2;
        W.result   = nullptr;
        W.hostPath = nullptr;
        const size_t pathFloats = (size_t)(cfgs[0].max_iter + 2) * (size_t)hq[0].D;
        if(n == 1 && pathFloats <= kHostPathFloats)
            {
                // a single query publishes its result and path rows into
                // host-mapped memory (too long a possible path: the plan
                // synchronises and copies instead)
                HIP_TRY(hipHostMalloc((void**)&hostRes, sizeof(HostResult), hipHostMallocMapped));
                std::memset((void*)hostRes, 0, sizeof(HostResult));
                HIP_TRY(hipHostGetDevicePointer((void**)&W.result, hostRes, 0));
                HIP_TRY(hipHostMalloc((void**)&hostPath, pathFloats * sizeof(float), hipHostMallocMapped));
                HIP_TRY(hipHostGetDevicePointer((void**)&W.hostPath, hostPath, 0));
            }
        // launch shape: persistent workgroups sized to the device
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, device));
        const uint32_t cus = (uint32_t)std::max(prop.multiProcessorCount, 1);
        L.propGrid         = cus * 4;
        L.chunkGrid        = cus * 4;
        return KGMT_OK;
    }

    int set_stream(void* s)
    {
        if(stream) HIP_TRY(hipStreamSynchronize(stream));
        if(ownStream && stream) HIP_TRY(hipStreamDestroy(stream));
        ownStream = false;
        stream    = (hipStream_t)s;
        if(!stream)
            {
                HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
                ownStream = true;
            }
        drop_graph();
        return KGMT_OK;
    }

    void drop_graph()
    {
        if(graphExec && stream) (void)hipStreamSynchronize(stream);  // a replay may still drain
        if(graphExec) (void)hipGraphExecDestroy(graphExec);
        if(graph) (void)hipGraphDestroy(graph);
        graphExec = nullptr, graph = nullptr, graphObs = 0xffffffffu;
    }

    int ensure_obstacles(uint32_t pairs)
    {
        if(pairs <= obsCap && obsPool) return KGMT_OK;
        const uint32_t cap = std::max<uint32_t>(pairs, 64);
        float4* p          = nullptr;
        if(hipMalloc((void**)&p, (size_t)cap * 2 * sizeof(float4)) != hipSuccess)
            return fail(KGMT_E_NOMEM, "obstacle pool allocation failed");
        if(obsPool)
            {
                auto it = std::find(owned.begin(), owned.end(), (void*)obsPool);
                if(it != owned.end()) owned.erase(it);
                HIP_TRY(hipStreamSynchronize(stream));
                (void)hipFree(obsPool);
            }
        owned.push_back(p);
        obsPool = p, obsCap = cap, P.obs = p;
        drop_graph();
        return KGMT_OK;
    }

    int ensure_rows(uint32_t rows)
    {
        if(rows <= rowsCap && rowsBuf) return KGMT_OK;
        const uint32_t cap = std::max<uint32_t>(rows, 256);
        float* p           = nullptr;
        if(hipMalloc((void**)&p, (size_t)cap * 6 * sizeof(float)) != hipSuccess)
            return fail(KGMT_E_NOMEM, "obstacle row buffer allocation failed");
        if(rowsBuf)
            {
                auto it = std::find(owned.begin(), owned.end(), (void*)rowsBuf);
                if(it != owned.end()) owned.erase(it);
                HIP_TRY(hipStreamSynchronize(stream));
                (void)hipFree(rowsBuf);
            }
        owned.push_back(p);
        rowsBuf = p, rowsCap = cap;
        return KGMT_OK;
    }

    // One iteration: 2 launches for one query, 3 for a batch (k_finish).
    // marks: timing events around them (profiling).
    int enqueue_iteration(uint32_t parity, hipEvent_t* marks = nullptr)
    {
        const bool all = marks && profiling > 1;
        if(marks) HIP_TRY(hipEventRecord(marks[0], stream));
        HIP_TRY(launch_propagate(model, dq, ds, P, W, parity, maxPairs, L, stream));
        if(marks) HIP_TRY(hipEventRecord(marks[1], stream));
        HIP_TRY(launch_update(dq, ds, P, W, parity, L, true, stream));
        if(all) HIP_TRY(hipEventRecord(marks[2], stream));
        // a single query's last update job closes the iteration itself
        if(nq > 1)
            {
                HIP_TRY(launch_finish(dq, ds, W, parity, stream));
                if(all) HIP_TRY(hipEventRecord(marks[3], stream));
            }
        return KGMT_OK;
    }

    int set_profiling(int on)
    {
        if(on < 0 || on > 2) return fail(KGMT_E_ARG, "profiling level must be 0, 1 or 2");
        profiling = on;
        std::memset(&prof, 0, sizeof(prof));
        const size_t need = (size_t)(maxIter + 1) * (KGMT_PROFILE_KERNELS + 1);
        if(on && profEv.size() < need)
            {
                const size_t old = profEv.size();
                profEv.resize(need, nullptr);
                for(size_t i = old; i < need; ++i) HIP_TRY(hipEventCreate(&profEv[i]));
            }
        return KGMT_OK;
    }

    int collect_profile(uint32_t iters)
    {
        for(uint32_t it = 0; it < iters; ++it)
            {
                hipEvent_t* m = &profEv[(size_t)it * (KGMT_PROFILE_KERNELS + 1)];
                // propagate, update and (batches) finish; the last two
                // profile slots stay unused
                const int nk = profiling > 1 ? (nq > 1 ? 3 : 2) : 1;
                for(int k = 0; k < nk; ++k)
                    {
                        float ms = 0.0f;
                        HIP_TRY(hipEventElapsedTime(&ms, m[k], m[k + 1]));
                        prof.ms[k] += ms;
                        prof.launches[k] += 1;
                    }
            }
        prof.iterations += iters;
        for(uint32_t q = 0; q < nq; ++q) prof.children += hs[q].expanded, prof.flagged += hs[q].flagged;
        return KGMT_OK;
    }

    // graph of two iterations (parities 0 and 1): no memset or copy nodes
    int build_graph()
    {
        drop_graph();
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_iteration(0);
        if(!rc) rc = enqueue_iteration(1);
        hipGraph_t g  = nullptr;
        hipError_t e2 = hipStreamEndCapture(stream, &g);
        if(rc) return rc;
        if(e2 != hipSuccess) return fail(KGMT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
        graph = g;
        HIP_TRY(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0));
        graphObs    = maxPairs;
        graphLaunch = L;
        return KGMT_OK;
    }

    // The host-mapped words cleared, the descriptors to the device and k_init
    // enqueued: a single query's descriptor travels as k_init's argument (no
    // upload); a batch's is copied, and k_init's last workgroup leaves its
    // arrival counter at zero (after a plan that did not complete it is
    // cleared here).
    int start_plan()
    {
        __atomic_store_n(pinned, 0u, __ATOMIC_RELEASE);
        __atomic_store_n(pinned + 1, 0u, __ATOMIC_RELEASE);
        __atomic_store_n(pinned + 2, 0u, __ATOMIC_RELEASE);
        const float* rows = sceneBuild ? sceneRows : nullptr;
        if(nq == 1)
            {
                HIP_TRY(launch_init_one(hq[0], dq, ds, P, W, rows, stream));
                return KGMT_OK;
            }
        HIP_TRY(hipMemcpyAsync(dq, hq.data(), sizeof(QueryDesc) * nq, hipMemcpyHostToDevice, stream));
        if(counterDirty) HIP_TRY(hipMemsetAsync(W.counter, 0, sizeof(uint32_t), stream));
        counterDirty = true;
        HIP_TRY(launch_init(dq, ds, P, W, rows, stream));
        return KGMT_OK;
    }

    // Upload descriptors, run init + iterations until every query is done.
    // The host enqueues launch units (one iteration, or the two-iteration
    // graph) and stays at most kLag units ahead of the iterations the device
    // reports closed, so the GPU always has queued work; kernels of finished
    // queries exit at once.
    int run(bool useGraph)
    {
        const auto t0 = std::chrono::steady_clock::now();
        mirValid      = false;
        int rc0       = start_plan();
        if(rc0) return rc0;
        if(useGraph && (!graphExec || graphObs != maxPairs || graphLaunch.propGrid != L.propGrid ||
                        graphLaunch.chunkGrid != L.chunkGrid || graphLaunch.propBlock != L.propBlock ||
                        graphLaunch.lanes != L.lanes))
            {
                int rc = build_graph();
                if(rc) return rc;
            }
        const bool prof        = profiling && !useGraph;
        const uint32_t perUnit = useGraph ? 2 : 1;
        // one query, not profiling: the device publishes the result into
        // host-mapped memory before the completion word, and the host returns
        // as soon as it sees it (no path launch, copy or stream wait)
        const bool early = hostRes != nullptr && !prof;
        if(hostRes) hostRes->iterations = 0;
        bool published = false;
        auto seen      = [&]() { return early && __atomic_load_n(pinned, __ATOMIC_ACQUIRE) != 0; };
        // The closing job of every iteration (k_finish for a batch) writes the
        // number of iterations closed so far into host-mapped memory; the host
        // stays at most kLag launch units ahead of it. No event between the
        // units: an event record is a system-scope release on the stream.
        iterMs.assign(maxIter + 2, 0.0);
        uint32_t it = 0, noted = 0, spins = 0;
        auto lastSeen = std::chrono::steady_clock::now();
        bool finished = false;
        auto note = [&](uint32_t c) {  // when each iteration was seen closed
            if(c <= noted) return;
            const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            for(uint32_t k = noted; k < c && k < maxIter + 2; ++k) iterMs[k] = t;
            noted = c;
        };
        while(it < maxIter && !finished)
            {
                if(seen())
                    {
                        finished = published = true;
                        break;
                    }
                if(useGraph)
                    HIP_TRY(hipGraphLaunch(graphExec, stream));
                else
                    {
                        int rc = enqueue_iteration(it & 1u,
                                                   prof ? &profEv[(size_t)it * (KGMT_PROFILE_KERNELS + 1)] : nullptr);
                        if(rc) return rc;
                    }
                it += perUnit;
                // wait while the device is more than kLag units behind (after
                // the last unit: until it has closed every iteration)
                const uint32_t ahead = (prof ? kLagProfiled : it > kLongAfter ? kLagLong : kLag) * perUnit;
                const uint32_t need  = it < maxIter ? (it > ahead ? it - ahead : 0u) : maxIter;
                while(true)
                    {
                        if(seen())
                            {
                                finished = published = true;
                                break;
                            }
                        if(__atomic_load_n(pinned, __ATOMIC_ACQUIRE))
                            {
                                finished = true;
                                break;
                            }
                        const uint32_t c = __atomic_load_n(pinned + 2, __ATOMIC_ACQUIRE);
                        if(c != noted) lastSeen = std::chrono::steady_clock::now();
                        note(c);
                        if(c >= need) break;
                        // a failed launch (or a stream that drained without
                        // closing them) ends the wait: asked only after 200 us
                        // without a closed iteration (a stream query is not free)
                        if((++spins & 63u) == 0u &&
                           std::chrono::steady_clock::now() - lastSeen > std::chrono::microseconds(200))
                            {
                                const hipError_t q = hipStreamQuery(stream);
                                if(q == hipSuccess)
                                    {
                                        note(__atomic_load_n(pinned + 2, __ATOMIC_ACQUIRE));
                                        if(__atomic_load_n(pinned, __ATOMIC_ACQUIRE)) finished = true;
                                        break;
                                    }
                                if(q != hipErrorNotReady)
                                    return fail(KGMT_E_HIP, std::string("iteration: ") + hipGetErrorString(q));
                            }
                    }
            }
        if(published && hostRes->iterations != 0)
            {
                const volatile HostResult* h = hostRes;
                QueryState& s                = hs[0];
                s.done = 1, s.success = h->success;
                s.itr = s.finishItr = h->iterations;
                s.treeSize          = h->treeSize;
                s.goalIdx           = h->success ? h->goalIdx : kNoGoal;
                s.pathRows          = h->pathRows;
                s.pathCost = h->pathCost, s.pathLen = h->pathLen;
                s.expanded = h->expanded, s.flagged = h->flagged;
                counterDirty = false;
                itersRun     = std::min(it, maxIter);
                return device_error();
            }
        HIP_TRY(launch_paths(dq, ds, P, nq, stream));
        HIP_TRY(hipMemcpyAsync(hs.data(), ds, sizeof(QueryState) * nq, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        counterDirty = false;
        itersRun     = std::min(it, maxIter);
        if(int rc = device_error()) return rc;
        if(prof) return collect_profile(it);
        return KGMT_OK;
    }

    // An update job whose look-back wait ran out of its bound. Every wait
    // ends (jobs are taken in order by running workgroups); the bound only
    // keeps a kernel from spinning forever if that ever failed.
    int device_error() const
    {
        if(__atomic_load_n(pinned + 1, __ATOMIC_ACQUIRE))
            return fail(KGMT_E_HIP, "frontier update: a look-back wait ran out of its bound");
        return KGMT_OK;
    }

    // Step-by-step planning (planBench): begin() prepares the plan as run()
    // does, step() runs one iteration and waits for it, so that the state
    // after every iteration can be read.
    bool stepping     = false;
    uint32_t stepItr  = 0;
    int begin()
    {
        mirValid = false;
        if(hostRes) hostRes->iterations = 0;
        int rc = start_plan();
        if(rc) return rc;
        rc = pull_state();
        if(rc) return rc;
        stepping = true;
        stepItr  = 0;
        itersRun = 0;
        return KGMT_OK;
    }

    // the device query states, with each iteration's closing values (the
    // next-state fields) taken as current, as the next score kernel would
    int pull_state()
    {
        HIP_TRY(hipMemcpyAsync(hs.data(), ds, sizeof(QueryState) * nq, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        for(uint32_t q = 0; q < nq; ++q)
            {
                QueryState& st = hs[q];
                st.itr = st.itrN, st.treeSize = st.treeSizeN, st.nChildren = st.nChildrenN, st.per = st.perN;
                st.done = st.doneN, st.success = st.successN, st.finishItr = st.finishItrN;
                st.expanded = st.expandedN, st.flagged = st.flaggedN;
            }
        return KGMT_OK;
    }

    int step(int32_t* done)
    {
        if(!stepping) return fail(KGMT_E_ARG, "no plan in progress (kgmt_plan_begin)");
        mirValid = false;
        bool all = true;
        for(uint32_t q = 0; q < nq; ++q) all = all && hs[q].done;
        if(!all && stepItr < maxIter)
            {
                int rc = enqueue_iteration(stepItr & 1u);
                if(rc) return rc;
                ++stepItr;
                itersRun = stepItr;
                if((rc = pull_state()) || (rc = device_error())) return rc;
                all = true;
                for(uint32_t q = 0; q < nq; ++q) all = all && hs[q].done;
            }
        if(all || stepItr >= maxIter)
            {
                // goal paths of solved queries (a single query has walked its own)
                HIP_TRY(launch_paths(dq, ds, P, nq, stream));
                HIP_TRY(hipMemcpyAsync(hs.data(), ds, sizeof(QueryState) * nq, hipMemcpyDeviceToHost, stream));
                HIP_TRY(hipStreamSynchronize(stream));
                if(hostRes && hs[0].success && hostRes->pathRows) hs[0].pathRows = hostRes->pathRows;
                stepping     = false;
                counterDirty = false;
                all          = true;
            }
        *done = all ? 1 : 0;
        return KGMT_OK;
    }

    // frontier flags and repeat counts of query q's tree nodes after the
    // last iteration: 1 for an old node it re-activated, 15 or 1 for a node it
    // appended (its region's valid count, final for that iteration, below
    // sparse_valid_below or not), 0 otherwise (KGMT.cu:336-339, 373)
    int get_frontier(uint32_t q, uint8_t* frontier, uint32_t* repeat, uint32_t cap, uint32_t* n)
    {
        HIP_TRY(hipStreamSynchronize(stream));  // a plan returns before its trailing launches drain
        QueryState st;
        HIP_TRY(hipMemcpy(&st, ds + q, sizeof(QueryState), hipMemcpyDeviceToHost));
        const QueryDesc& qd = hq[q];
        const uint32_t ts   = st.treeSizeN;
        if(n) *n = ts;
        if(!frontier && !repeat) return KGMT_OK;
        if(cap < ts) return fail(KGMT_E_CAPACITY, "frontier output buffer too small");
        if(frontier) HIP_TRY(hipMemcpy(frontier, P.frontier + qd.nodeOff, ts, hipMemcpyDeviceToHost));
        if(repeat)
            {
                const uint32_t next = std::min(st.next, ts), old = ts - next;
                std::vector<uint8_t> react(old);
                std::vector<int32_t> r1(next), valid(qd.R1);
                if(old) HIP_TRY(hipMemcpy(react.data(), P.react + qd.nodeOff, old, hipMemcpyDeviceToHost));
                if(next)
                    HIP_TRY(hipMemcpy(r1.data(), P.r1node + qd.nodeOff + old, sizeof(int32_t) * next,
                                      hipMemcpyDeviceToHost));
                HIP_TRY(hipMemcpy(valid.data(), P.validc + qd.regOff, sizeof(int32_t) * qd.R1, hipMemcpyDeviceToHost));
                for(uint32_t i = 0; i < old; ++i) repeat[i] = react[i] ? 1u : 0u;
                for(uint32_t j = 0; j < next; ++j)
                    repeat[old + j] = valid[r1[j]] < qd.cfg.sparse_valid_below ? qd.cfg.repeat_sparse
                                                                                  : qd.cfg.repeat_dense;
            }
        return KGMT_OK;
    }

    void result(uint32_t q, kgmt_result* out, double wallMs) const
    {
        const QueryState& s = hs[q];
        out->success        = (int32_t)s.success;
        out->goal_index     = s.success ? s.goalIdx : -1;
        out->iterations     = s.done ? s.finishItr : s.itr;
        out->tree_size      = s.treeSize;
        out->path_rows      = s.pathRows;
        out->path_cost      = s.pathCost;
        out->path_length    = s.pathLen;
        out->time_ms        = (float)wallMs;
        out->expanded       = s.expanded;
    }

    int get_tree(uint32_t q, float* samples, int32_t* parents, float* costs, uint32_t cap, uint32_t* n)
    {
        HIP_TRY(hipStreamSynchronize(stream));  // a plan returns before its trailing launches drain
        const QueryDesc& qd = hq[q];
        const uint32_t ts   = hs[q].treeSize;
        if(n) *n = ts;
        if(!samples && !parents && !costs) return KGMT_OK;
        if(cap < ts) return fail(KGMT_E_CAPACITY, "tree output buffer too small");
        if(samples)
            {
                std::vector<float> soa((size_t)qd.D * ts);
                for(int d = 0; d < qd.D; ++d)
                    HIP_TRY(hipMemcpy(soa.data() + (size_t)d * ts, P.treeS + qd.sampleOff + (size_t)d * qd.cap,
                                      sizeof(float) * ts, hipMemcpyDeviceToHost));
                for(uint32_t i = 0; i < ts; ++i)
                    for(int d = 0; d < qd.D; ++d) samples[(size_t)i * qd.D + d] = soa[(size_t)d * ts + i];
            }
        if(parents)
            HIP_TRY(hipMemcpy(parents, P.parent + qd.nodeOff, sizeof(int32_t) * ts, hipMemcpyDeviceToHost));
        if(costs) HIP_TRY(hipMemcpy(costs, P.cost + qd.nodeOff, sizeof(float) * ts, hipMemcpyDeviceToHost));
        return KGMT_OK;
    }

    // query 0's tree in the reference's layout in library-owned device
    // memory, complete when this returns
    int tree_device(kgmt_device_tree* out)
    {
        const QueryDesc& qd = hq[0];
        int rc              = 0;
        if(!mirRows && ((rc = dev_alloc(mirRows, (size_t)qd.cap * qd.D, owned)) ||
                        (rc = dev_alloc(mirPar, qd.cap, owned)) || (rc = dev_alloc(mirCost, qd.cap, owned))))
            return rc;
        const uint32_t ts = hs[0].treeSize;
        if(!mirValid)
            {
                HIP_TRY(launch_tree_rows(P.treeS + qd.sampleOff, P.parent + qd.nodeOff, P.cost + qd.nodeOff, qd.cap,
                                         qd.D, ts, mirRows, mirPar, mirCost, stream));
                HIP_TRY(hipStreamSynchronize(stream));
                mirValid = true;
            }
        out->samples    = mirRows;
        out->parents    = mirPar;
        out->costs      = mirCost;
        out->tree_size  = ts;
        out->capacity   = qd.cap;
        out->sample_dim = qd.D;
        out->reserved   = 0;
        return KGMT_OK;
    }

    int get_path(uint32_t q, float* rows, uint32_t capRows, uint32_t* nRows)
    {
        const QueryDesc& qd = hq[q];
        const uint32_t r    = hs[q].success ? hs[q].pathRows : 0;
        if(nRows) *nRows = r;
        if(!rows || r == 0) return KGMT_OK;
        if(capRows < r) return fail(KGMT_E_CAPACITY, "path output buffer too small");
        if(hostPath)
            {
                // a single query's rows were written to host memory before
                // its result was published (or the plan synchronised)
                std::memcpy(rows, (const void*)hostPath, sizeof(float) * r * qd.D);
                return KGMT_OK;
            }
        HIP_TRY(hipStreamSynchronize(stream));  // a plan returns before its trailing launches drain
        HIP_TRY(hipMemcpy(rows, P.path + qd.pathOff, sizeof(float) * r * qd.D, hipMemcpyDeviceToHost));
        return KGMT_OK;
    }
};

struct kgmt_ctx
{
    Engine eng;
    bool useGraph = false;
};

struct kgmt_batch
{
    Engine eng;
};

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int kgmt_abi_version(void)
{
    return KGMT_ABI_VERSION;
}

const char* kgmt_last_error(void)
{
    return g_err.c_str();
}

int kgmt_model_dims(int32_t model, int32_t* s, int32_t* c, int32_t* d)
{
    int S, C;
    if(model_dims(model, S, C)) return fail(KGMT_E_UNSUPPORTED, "model not built");
    if(s) *s = S;
    if(c) *c = C;
    if(d) *d = S + C + 1;
    return KGMT_OK;
}

int kgmt_config_preset(int32_t preset, kgmt_config* o)
{
    if(!o) return fail(KGMT_E_ARG, "out is null");
    std::memset(o, 0, sizeof(*o));
    // shared planner constants (KGMT.cu:30, 67, 336-339; config.h:262)
    o->epsilon            = 1e-2f;
    o->children_per_entry = 32;
    o->root_repeat        = 5;
    o->repeat_sparse      = 15;
    o->repeat_dense       = 1;
    o->sparse_valid_below = 10;
    o->max_iter           = 300;
    o->quad_mass = 1.0f, o->quad_ix = 1.0f, o->quad_iy = 1.0f, o->quad_iz = 2.0f;
    o->quad_nu = 10e-3f, o->quad_mu = 2e-6f, o->gravity = -9.81f;
    o->pitch_min = -(float)M_PI / 3.0f, o->pitch_max = (float)M_PI / 3.0f;
    o->uni_length = 1.0f;  // UNI_LENGTH, config.h:270
    o->seed      = 1;
    const float pi = (float)M_PI;
    switch(preset)
        {
            case KGMT_PRESET_REFERENCE_QUAD:  // config.h:209-307
                o->model = KGMT_MODEL_QUADCOPTER;
                o->capacity = 400000, o->dt = 0.1f, o->steps_min = 1, o->steps_max = 10;
                o->goal_thresh = 5.0f, o->ws_min = 0.0f, o->ws_max = 100.0f;
                o->v_min = -30.0f, o->v_max = 30.0f;
                o->ctrl_min[0] = 0.0f, o->ctrl_max[0] = 30.0f;
                for(int i = 1; i < 4; ++i) o->ctrl_min[i] = -pi, o->ctrl_max[i] = pi;
                o->r1_len = 8, o->r2_len = 2;
                return KGMT_OK;
            case KGMT_PRESET_DI6D:
                o->model = KGMT_MODEL_DOUBLE_INTEGRATOR, o->integrator = KGMT_DI_REFERENCE;
                o->capacity = 1u << 17, o->dt = 0.05f, o->steps_min = 2, o->steps_max = 10;
                o->goal_thresh = 0.05f, o->ws_min = 0.0f, o->ws_max = 1.0f;
                o->v_min = -1.0f, o->v_max = 1.0f;
                for(int i = 0; i < 3; ++i) o->ctrl_min[i] = -1.0f, o->ctrl_max[i] = 1.0f;
                o->r1_len = 8, o->r2_len = 4;
                return KGMT_OK;
            case KGMT_PRESET_DUBINS5D:
                o->model    = KGMT_MODEL_DUBINS_5D;
                o->capacity = 1u << 18, o->dt = 0.02f, o->steps_min = 5, o->steps_max = 25;
                o->goal_thresh = 0.05f, o->ws_min = 0.0f, o->ws_max = 1.0f;
                o->v_min = 0.0f, o->v_max = 1.0f;
                o->ctrl_min[0] = -1.0f, o->ctrl_max[0] = 1.0f;
                o->ctrl_min[1] = -0.5f, o->ctrl_max[1] = 0.5f;
                o->ctrl_min[2] = 0.2f, o->ctrl_max[2] = 0.4f;
                o->r1_len = 16, o->r2_len = 2;
                return KGMT_OK;
            case KGMT_PRESET_QUAD12D:
                o->model    = KGMT_MODEL_QUADCOPTER;
                o->capacity = 1u << 19, o->dt = 0.01f, o->steps_min = 10, o->steps_max = 50;
                o->goal_thresh = 0.05f, o->ws_min = 0.0f, o->ws_max = 1.0f;
                o->v_min = -5.0f, o->v_max = 5.0f;
                o->ctrl_min[0] = 0.0f, o->ctrl_max[0] = 19.62f;
                for(int i = 1; i < 4; ++i) o->ctrl_min[i] = -0.05f, o->ctrl_max[i] = 0.05f;
                o->quad_ix = 0.01f, o->quad_iy = 0.01f, o->quad_iz = 0.02f;
                o->r1_len = 10, o->r2_len = 4;
                return KGMT_OK;
            case KGMT_PRESET_UNICYCLE2D:  // steering and UNI_LENGTH of config.h:266-270, A of :34-35
                o->model    = KGMT_MODEL_UNICYCLE;
                o->capacity = 1u << 17, o->dt = 0.1f, o->steps_min = 1, o->steps_max = 10;
                o->goal_thresh = 0.05f, o->ws_min = 0.0f, o->ws_max = 1.0f;
                o->ctrl_min[0] = -0.2f, o->ctrl_max[0] = 0.2f;
                o->ctrl_min[1] = -pi / 2.0f, o->ctrl_max[1] = pi / 2.0f;
                o->r1_len = 16, o->r2_len = 4;
                return KGMT_OK;
            default: return fail(KGMT_E_ARG, "unknown preset");
        }
}

int kgmt_config_check(const kgmt_config* cfg)
{
    return check_config(cfg);
}

int kgmt_create(const kgmt_config* cfg, kgmt_ctx** out)
{
    if(!cfg || !out) return fail(KGMT_E_ARG, "null argument");
    *out   = nullptr;
    auto* c = new(std::nothrow) kgmt_ctx;
    if(!c) return fail(KGMT_E_NOMEM, "context allocation failed");
    int rc = c->eng.init(cfg, 1);
    if(rc)
        {
            delete c;
            return rc;
        }
    *out = c;
    return KGMT_OK;
}

int kgmt_destroy(kgmt_ctx* ctx)
{
    delete ctx;
    return KGMT_OK;
}

int kgmt_set_seed(kgmt_ctx* ctx, uint64_t seed)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    ctx->eng.hq[0].cfg.seed = seed;
    return KGMT_OK;
}

static int set_launch(Engine& e, uint32_t propagate_block, uint32_t lanes_per_child, uint32_t propagate_grid,
                      uint32_t chunk_grid)
{
    const uint32_t blk = propagate_block ? propagate_block : e.L.propBlock;
    const uint32_t ln  = lanes_per_child ? lanes_per_child : e.L.lanes;
    if(!launch_supported(blk, ln))
        return fail(KGMT_E_UNSUPPORTED, "propagate shape must be 128/256/512 threads with 1 lane per child, "
                                        "or 256 threads with 2 or 4 lanes per child");
    if(propagate_grid > (1u << 20) || chunk_grid > (1u << 20)) return fail(KGMT_E_ARG, "grid too large");
    e.L.propBlock = blk, e.L.lanes = ln, e.W.tile = blk / ln;
    if(propagate_grid) e.L.propGrid = propagate_grid;
    if(chunk_grid) e.L.chunkGrid = chunk_grid;
    return KGMT_OK;
}

int kgmt_set_launch(kgmt_ctx* ctx, uint32_t propagate_block, uint32_t lanes_per_child, uint32_t propagate_grid,
                    uint32_t chunk_grid)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    return set_launch(ctx->eng, propagate_block, lanes_per_child, propagate_grid, chunk_grid);
}

int kgmt_set_graph(kgmt_ctx* ctx, int32_t on)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    ctx->useGraph = on != 0;
    return KGMT_OK;
}

int kgmt_set_stream(kgmt_ctx* ctx, void* s)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    return ctx->eng.set_stream(s);
}

int kgmt_synchronize(kgmt_ctx* ctx)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->eng.stream));
    return KGMT_OK;
}

// Obstacle rows of every query -> scenes (grouped, Morton-sorted) in the pool,
// built on the device. Rows hold 2 * W floats (W = the queries' workspace
// dimension); they may be device memory (the reference passes a device
// pointer), which the scene kernels read in place: reject a pointer that the
// device cannot read (for example host memory passed as device memory).
static int check_device_rows(const float* p)
{
    hipPointerAttribute_t attr{};
    if(hipPointerGetAttributes(&attr, p) != hipSuccess)
        {
            (void)hipGetLastError();
            return fail(KGMT_E_ARG, "obstacle pointer is not device memory");
        }
    if(attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged && attr.type != hipMemoryTypeUnified)
        return fail(KGMT_E_ARG, "obstacle pointer is not device memory");
    return KGMT_OK;
}

static int upload_obstacles(Engine& e, const float* obstacles, const uint32_t* offsets, const uint32_t* counts,
                            uint32_t total, int onDevice)
{
    if(total && !obstacles) return fail(KGMT_E_ARG, "obstacles pointer is null");
    uint32_t pairs = 0;
    for(uint32_t q = 0; q < e.nq; ++q)
        {
            const uint32_t c   = counts ? counts[q] : total;
            const uint32_t off = offsets ? offsets[q] : 0;
            if(c > kMaxObs) return fail(KGMT_E_CAPACITY, "more than 1500 obstacles in one query");
            if((uint64_t)off + c > (uint64_t)total) return fail(KGMT_E_ARG, "obstacle range out of bounds");
            pairs += scene_pairs(c);
        }
    int rc = e.ensure_obstacles(pairs);
    if(rc) return rc;
    // scenes are built on the device (run()) from the caller's device rows or
    // from a device copy of the host rows
    e.sceneRows  = nullptr;
    e.sceneBuild = total != 0;
    if(total)
        {
            if(onDevice)
                {
                    if((rc = check_device_rows(obstacles))) return rc;
                    e.sceneRows = obstacles;
                }
            else
                {
                    if((rc = e.ensure_rows(total))) return rc;
                    const size_t rowW = 2 * (size_t)e.hq[0].wdim;
                    HIP_TRY(hipMemcpyAsync(e.rowsBuf, obstacles, (size_t)total * rowW * sizeof(float),
                                           hipMemcpyHostToDevice, e.stream));
                    e.sceneRows = e.rowsBuf;
                }
        }
    uint32_t p = 0;
    e.maxPairs = 0;
    for(uint32_t q = 0; q < e.nq; ++q)
        {
            const uint32_t c   = counts ? counts[q] : total;
            const uint32_t off = offsets ? offsets[q] : 0;
            QueryDesc& qd      = e.hq[q];
            qd.obsOff          = p;
            qd.nObs            = c;
            qd.nGrp            = scene_groups(c);
            qd.rowOff          = off;
            p += scene_pairs(c);
            e.maxPairs = std::max(e.maxPairs, scene_pairs(c));
        }
    return KGMT_OK;
}

int kgmt_plan(kgmt_ctx* ctx, const float* start, const float* goal, const float* obstacles, uint32_t n_obs,
              int32_t on_device, kgmt_result* out)
{
    if(!ctx || !start || !goal || !out) return fail(KGMT_E_ARG, "null argument");
    const auto t0 = std::chrono::steady_clock::now();
    Engine& e     = ctx->eng;
    QueryDesc& qd = e.hq[0];
    std::memset(qd.start, 0, sizeof(qd.start));
    std::memcpy(qd.start, start, sizeof(float) * qd.D);
    qd.goal[0] = goal[0], qd.goal[1] = goal[1], qd.goal[2] = qd.wdim == 3 ? goal[2] : 0.0f;
    int rc = upload_obstacles(e, obstacles, nullptr, nullptr, n_obs, on_device);
    if(rc) return rc;
    rc = e.run(ctx->useGraph && !e.profiling);
    if(rc) return rc;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    e.result(0, out, ms);
    return KGMT_OK;
}

int kgmt_plan_begin(kgmt_ctx* ctx, const float* start, const float* goal, const float* obstacles, uint32_t n_obs,
                    int32_t on_device)
{
    if(!ctx || !start || !goal) return fail(KGMT_E_ARG, "null argument");
    Engine& e     = ctx->eng;
    QueryDesc& qd = e.hq[0];
    std::memset(qd.start, 0, sizeof(qd.start));
    std::memcpy(qd.start, start, sizeof(float) * qd.D);
    qd.goal[0] = goal[0], qd.goal[1] = goal[1], qd.goal[2] = qd.wdim == 3 ? goal[2] : 0.0f;
    int rc = upload_obstacles(e, obstacles, nullptr, nullptr, n_obs, on_device);
    if(rc) return rc;
    return e.begin();
}

int kgmt_plan_step(kgmt_ctx* ctx, kgmt_result* out, int32_t* done)
{
    if(!ctx || !out || !done) return fail(KGMT_E_ARG, "null argument");
    Engine& e     = ctx->eng;
    const auto t0 = std::chrono::steady_clock::now();
    int rc        = e.step(done);
    if(rc) return rc;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    e.result(0, out, ms);
    if(!*done) out->iterations = e.stepItr;
    return KGMT_OK;
}

int kgmt_get_frontier(kgmt_ctx* ctx, uint8_t* frontier, uint32_t* repeat, uint32_t cap, uint32_t* n)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    return ctx->eng.get_frontier(0, frontier, repeat, cap, n);
}

int kgmt_set_profiling(kgmt_ctx* ctx, int32_t on)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    return ctx->eng.set_profiling(on);
}

int kgmt_get_profile(kgmt_ctx* ctx, kgmt_profile* out)
{
    if(!ctx || !out) return fail(KGMT_E_ARG, "null argument");
    *out = ctx->eng.prof;
    return KGMT_OK;
}

int kgmt_batch_set_profiling(kgmt_batch* b, int32_t on)
{
    if(!b) return fail(KGMT_E_ARG, "null batch");
    return b->eng.set_profiling(on);
}

int kgmt_batch_get_profile(kgmt_batch* b, kgmt_profile* out)
{
    if(!b || !out) return fail(KGMT_E_ARG, "null argument");
    *out = b->eng.prof;
    return KGMT_OK;
}

int kgmt_get_tree(kgmt_ctx* ctx, float* samples, int32_t* parents, float* costs, uint32_t cap, uint32_t* n)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    return ctx->eng.get_tree(0, samples, parents, costs, cap, n);
}

int kgmt_get_tree_device(kgmt_ctx* ctx, kgmt_device_tree* out)
{
    if(!ctx || !out) return fail(KGMT_E_ARG, "null argument");
    return ctx->eng.tree_device(out);
}

int kgmt_get_path(kgmt_ctx* ctx, float* rows, uint32_t cap_rows, uint32_t* n_rows)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    return ctx->eng.get_path(0, rows, cap_rows, n_rows);
}

int kgmt_get_regions(kgmt_ctx* ctx, int32_t* total, int32_t* valid, float* score, uint32_t cap, uint32_t* n)
{
    if(!ctx) return fail(KGMT_E_ARG, "null context");
    const Engine& e     = ctx->eng;
    const QueryDesc& qd = e.hq[0];
    if(n) *n = (uint32_t)qd.R1;
    if(cap < (uint32_t)qd.R1 && (total || valid || score)) return fail(KGMT_E_CAPACITY, "region buffer too small");
    HIP_TRY(hipStreamSynchronize(e.stream));  // a plan returns before its trailing launches drain
    if(total) HIP_TRY(hipMemcpy(total, e.P.total + qd.regOff, sizeof(int32_t) * qd.R1, hipMemcpyDeviceToHost));
    if(valid) HIP_TRY(hipMemcpy(valid, e.P.validc + qd.regOff, sizeof(int32_t) * qd.R1, hipMemcpyDeviceToHost));
    if(score) HIP_TRY(hipMemcpy(score, e.P.score + qd.regOff, sizeof(float) * qd.R1, hipMemcpyDeviceToHost));
    return KGMT_OK;
}

// ---- batch ----------------------------------------------------------------
int kgmt_batch_create(const kgmt_config* cfgs, uint32_t n_q, kgmt_batch** out)
{
    if(!cfgs || !out || n_q == 0) return fail(KGMT_E_ARG, "bad argument");
    *out   = nullptr;
    auto* b = new(std::nothrow) kgmt_batch;
    if(!b) return fail(KGMT_E_NOMEM, "batch allocation failed");
    int rc = b->eng.init(cfgs, n_q);
    if(rc)
        {
            delete b;
            return rc;
        }
    *out = b;
    return KGMT_OK;
}

int kgmt_batch_destroy(kgmt_batch* b)
{
    delete b;
    return KGMT_OK;
}

int kgmt_batch_set_stream(kgmt_batch* b, void* s)
{
    if(!b) return fail(KGMT_E_ARG, "null batch");
    return b->eng.set_stream(s);
}

int kgmt_batch_set_launch(kgmt_batch* b, uint32_t propagate_block, uint32_t lanes_per_child, uint32_t propagate_grid,
                          uint32_t chunk_grid)
{
    if(!b) return fail(KGMT_E_ARG, "null batch");
    return set_launch(b->eng, propagate_block, lanes_per_child, propagate_grid, chunk_grid);
}

int kgmt_batch_plan(kgmt_batch* b, const float* starts, const float* goals, const float* obstacles,
                    const uint32_t* obstacle_offsets, const uint32_t* obstacle_counts, int32_t use_graph,
                    kgmt_result* out, float* finish_ms)
{
    if(!b || !starts || !goals || !out || !obstacle_offsets || !obstacle_counts)
        return fail(KGMT_E_ARG, "null argument");
    const auto t0 = std::chrono::steady_clock::now();
    Engine& e     = b->eng;
    uint32_t total = 0;
    for(uint32_t q = 0; q < e.nq; ++q)
        {
            QueryDesc& qd = e.hq[q];
            std::memset(qd.start, 0, sizeof(qd.start));
            std::memcpy(qd.start, starts + (size_t)q * KGMT_MAX_SAMPLE, sizeof(float) * qd.D);
            const float* g = goals + (size_t)q * KGMT_MAX_SAMPLE;
            qd.goal[0] = g[0], qd.goal[1] = g[1], qd.goal[2] = qd.wdim == 3 ? g[2] : 0.0f;
            const uint64_t end = (uint64_t)obstacle_offsets[q] + obstacle_counts[q];
            if(end > 0xffffffffull) return fail(KGMT_E_ARG, "obstacle range out of bounds");
            total = std::max(total, (uint32_t)end);
        }
    int rc = upload_obstacles(e, obstacles, obstacle_offsets, obstacle_counts, total, 0);
    if(rc) return rc;
    rc = e.run(use_graph != 0);
    if(rc) return rc;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for(uint32_t q = 0; q < e.nq; ++q)
        {
            e.result(q, &out[q], ms);
            if(finish_ms)
                {
                    const uint32_t fi = e.hs[q].done ? e.hs[q].finishItr : e.itersRun;
                    finish_ms[q]      = fi >= 1 && fi <= e.iterMs.size() && e.iterMs[fi - 1] > 0.0
                                          ? (float)e.iterMs[fi - 1]
                                          : (float)ms;
                }
        }
    return KGMT_OK;
}

int kgmt_batch_get_path(kgmt_batch* b, uint32_t q, float* rows, uint32_t cap_rows, uint32_t* n_rows)
{
    if(!b || q >= b->eng.nq) return fail(KGMT_E_ARG, "bad argument");
    return b->eng.get_path(q, rows, cap_rows, n_rows);
}

int kgmt_batch_get_tree(kgmt_batch* b, uint32_t q, float* samples, int32_t* parents, float* costs, uint32_t cap,
                        uint32_t* n)
{
    if(!b || q >= b->eng.nq) return fail(KGMT_E_ARG, "bad argument");
    return b->eng.get_tree(q, samples, parents, costs, cap, n);
}

// ---- stage-level ----------------------------------------------------------
extern "C++" {
namespace
{
struct StageDesc
{
    QueryDesc* d = nullptr;
    hipStream_t s;
    explicit StageDesc(hipStream_t st) : s(st) {}
    int make(const kgmt_config* cfg)
    {
        int rc = check_config(cfg);
        if(rc) return rc;
        QueryDesc h;
        fill_desc(h, *cfg);
        HIP_TRY(hipMallocAsync((void**)&d, sizeof(QueryDesc), s));
        HIP_TRY(hipMemcpyAsync(d, &h, sizeof(QueryDesc), hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        return KGMT_OK;
    }
    ~StageDesc()
    {
        if(d) (void)hipFreeAsync(d, s);
    }
};

// Stream-ordered scratch buffers of one stage call, released on every path
// (hipFreeAsync on the call's stream after the work enqueued on it).
struct AsyncBuffers
{
    hipStream_t s;
    std::vector<void*> ptrs;
    explicit AsyncBuffers(hipStream_t st) : s(st) {}
    AsyncBuffers(const AsyncBuffers&)            = delete;
    AsyncBuffers& operator=(const AsyncBuffers&) = delete;
    ~AsyncBuffers()
    {
        for(void* p : ptrs) (void)hipFreeAsync(p, s);
    }
    template <typename T>
    int get(T*& p, size_t count)
    {
        void* q            = nullptr;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if(hipMallocAsync(&q, bytes, s) != hipSuccess)
            {
                (void)hipGetLastError();
                return fail(KGMT_E_NOMEM, "hipMallocAsync of " + std::to_string(bytes) + " bytes failed");
            }
        ptrs.push_back(q);
        p = static_cast<T*>(q);
        return KGMT_OK;
    }
};
}  // namespace
}  // extern "C++"

int kgmt_draw_controls(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t u0, uint32_t n,
                       float* d_controls, int32_t* d_steps, float* d_ufront, void* stream)
{
    if(n && (!d_controls || !d_steps || !d_ufront)) return fail(KGMT_E_ARG, "null buffer");
    StageDesc sd((hipStream_t)stream);
    int rc = sd.make(cfg);
    if(rc) return rc;
    int S, C;
    model_dims(cfg->model, S, C);
    HIP_TRY(launch_stage_draw(sd.d, C, query, itr, u0, n, d_controls, d_steps, d_ufront, sd.s));
    return KGMT_OK;
}

int kgmt_build_scene(const float* d_rows, uint32_t n, float* d_out, uint32_t out_floats, uint32_t* n_groups,
                     void* stream)
{
    if(!n_groups) return fail(KGMT_E_ARG, "null argument");
    if(n > kMaxObs) return fail(KGMT_E_CAPACITY, "more than 1500 obstacles");
    const uint32_t nGrp = scene_groups(n);
    *n_groups           = nGrp;
    if(n == 0) return KGMT_OK;
    if(!d_rows || !d_out) return fail(KGMT_E_ARG, "null buffer");
    if((uint64_t)out_floats < (uint64_t)scene_pairs(n) * 8) return fail(KGMT_E_CAPACITY, "scene buffer too small");
    int rc = check_device_rows(d_rows);
    if(rc) return rc;
    HIP_TRY(launch_build_scene(d_rows, n, nGrp, reinterpret_cast<float4*>(d_out), 3, (hipStream_t)stream));
    return KGMT_OK;
}

int kgmt_propagate_collide(const kgmt_config* cfg, const int32_t* d_parent_idx, const float* d_parent_states,
                           uint32_t n_parents, const float* d_controls, const int32_t* d_steps, uint32_t n_children,
                           const float* d_obstacles, uint32_t n_obs, float* d_child_samples, uint8_t* d_valid,
                           int32_t* d_r1, int32_t* d_r2, void* stream)
{
    if(n_children && (!d_parent_idx || !d_parent_states || !d_controls || !d_steps || !d_child_samples || !d_valid ||
                      !d_r1 || !d_r2))
        return fail(KGMT_E_ARG, "null buffer");
    if(n_obs > kMaxObs) return fail(KGMT_E_CAPACITY, "more than 1500 obstacles");
    if(n_obs && !d_obstacles) return fail(KGMT_E_ARG, "obstacles pointer is null");
    StageDesc sd((hipStream_t)stream);
    int rc = sd.make(cfg);
    if(rc) return rc;
    if(n_obs && (rc = check_device_rows(d_obstacles))) return rc;
    const uint32_t nGrp = scene_groups(n_obs);
    float4* d_scene     = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_scene, (size_t)std::max<uint32_t>(scene_pairs(n_obs), 1) * 2 * sizeof(float4),
                           sd.s));
    hipError_t e = launch_build_scene(d_obstacles, n_obs, nGrp, d_scene, model_wdim(cfg->model), sd.s);
    if(e == hipSuccess)
        e = launch_stage_propagate(cfg->model, sd.d, d_parent_idx, d_parent_states, n_parents, d_controls, d_steps,
                                   n_children, d_scene, nGrp, n_obs, d_child_samples, d_valid, d_r1, d_r2, sd.s);
    (void)hipFreeAsync(d_scene, sd.s);
    if(e != hipSuccess) return fail(KGMT_E_HIP, std::string("propagate: ") + hipGetErrorString(e));
    return KGMT_OK;
}

int kgmt_propagate_collide_noalloc(const kgmt_config* cfg, const int32_t* d_parent_idx, const float* d_parent_states,
                                   uint32_t n_parents, const float* d_controls, const int32_t* d_steps,
                                   uint32_t n_children, const float* d_obstacles, uint32_t n_obs,
                                   float* d_child_samples, uint8_t* d_valid, int32_t* d_r1, int32_t* d_r2,
                                   void* stream)
{
    if(!cfg) return fail(KGMT_E_ARG, "null configuration");
    if(n_children && (!d_parent_idx || !d_parent_states || !d_controls || !d_steps || !d_child_samples || !d_valid ||
                      !d_r1 || !d_r2))
        return fail(KGMT_E_ARG, "null buffer");
    if(n_obs > kMaxObs || scene_pairs(n_obs) > scratch_pairs())
        return fail(KGMT_E_CAPACITY, "more than 1500 obstacles: the scratch scene holds at most 1500 boxes");
    if(n_obs && !d_obstacles) return fail(KGMT_E_ARG, "obstacles pointer is null");
    int rc = check_config(cfg);
    if(rc) return rc;
    if(n_obs && (rc = check_device_rows(d_obstacles))) return rc;
    float4* scene = nullptr;
    HIP_TRY(scratch_address(&scene));
    QueryDesc h;
    fill_desc(h, *cfg);
    const hipStream_t s = (hipStream_t)stream;
    const uint32_t nGrp = scene_groups(n_obs);
    hipError_t e        = launch_build_scene(d_obstacles, n_obs, nGrp, scene, model_wdim(cfg->model), s);
    if(e == hipSuccess)
        e = launch_stage_propagate_by_value(cfg->model, h, d_parent_idx, d_parent_states, n_parents, d_controls,
                                            d_steps, n_children, scene, nGrp, d_child_samples, d_valid, d_r1, d_r2,
                                            s);
    if(e != hipSuccess) return fail(KGMT_E_HIP, std::string("propagate: ") + hipGetErrorString(e));
    return KGMT_OK;
}

int kgmt_propagate_collide_planner(const kgmt_config* cfg, const int32_t* d_parent_idx,
                                   const float* d_parent_states, uint32_t n_parents, const float* d_controls,
                                   const int32_t* d_steps, uint32_t n_children, const float* d_obstacles,
                                   uint32_t n_obs, float* d_child_samples, uint8_t* d_valid, int32_t* d_r1,
                                   int32_t* d_r2, uint8_t* d_check, float* d_edge_box, void* stream)
{
    if(n_children && (!d_parent_idx || !d_parent_states || !d_controls || !d_steps || !d_child_samples || !d_valid ||
                      !d_r1 || !d_r2))
        return fail(KGMT_E_ARG, "null buffer");
    if(n_obs > kMaxObs) return fail(KGMT_E_CAPACITY, "more than 1500 obstacles");
    if(n_obs && !d_obstacles) return fail(KGMT_E_ARG, "obstacles pointer is null");
    StageDesc sd((hipStream_t)stream);
    int rc = sd.make(cfg);
    if(rc) return rc;
    if(n_obs && (rc = check_device_rows(d_obstacles))) return rc;
    const uint32_t nGrp = scene_groups(n_obs);
    float4* d_scene     = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_scene, (size_t)std::max<uint32_t>(scene_pairs(n_obs), 1) * 2 * sizeof(float4),
                           sd.s));
    hipError_t e = launch_build_scene(d_obstacles, n_obs, nGrp, d_scene, model_wdim(cfg->model), sd.s);
    if(e == hipSuccess)
        e = launch_stage_propagate_planner(cfg->model, sd.d, d_parent_idx, d_parent_states, n_parents, d_controls,
                                           d_steps, n_children, d_scene, nGrp, d_child_samples, d_valid, d_r1, d_r2,
                                           d_check, d_edge_box, sd.s);
    (void)hipFreeAsync(d_scene, sd.s);
    if(e != hipSuccess) return fail(KGMT_E_HIP, std::string("propagate: ") + hipGetErrorString(e));
    return KGMT_OK;
}

int kgmt_frontier_update(const kgmt_config* cfg, uint32_t itr, uint32_t tree_size, float* d_tree_samples,
                         int32_t* d_parents, float* d_costs, uint8_t* d_frontier, const float* d_child_samples,
                         const int32_t* d_child_parent, const uint8_t* d_fnext, uint32_t n_children,
                         const int32_t* d_valid_count, const float* d_score, const float* goal, uint32_t* d_rep_idx,
                         int32_t* d_info, void* stream)
{
    if(!cfg || !goal) return fail(KGMT_E_ARG, "null argument");
    if(!d_tree_samples || !d_parents || !d_costs || !d_frontier || !d_valid_count || !d_score || !d_rep_idx ||
       !d_info || (n_children && (!d_child_samples || !d_child_parent || !d_fnext)))
        return fail(KGMT_E_ARG, "null buffer");
    int rc = check_config(cfg);
    if(rc) return rc;
    if(itr < 1) return fail(KGMT_E_ARG, "itr must be >= 1");
    if(tree_size < 1 || tree_size > cfg->capacity) return fail(KGMT_E_ARG, "tree_size must be in [1, capacity]");
    if(n_children > cfg->capacity) return fail(KGMT_E_CAPACITY, "more children than capacity");
    if(n_children > 1024u * kChunk) return fail(KGMT_E_CAPACITY, "at most 2^20 children");
    hipStream_t s = (hipStream_t)stream;
    QueryDesc h;
    fill_desc(h, *cfg);
    h.goal[0] = goal[0], h.goal[1] = goal[1], h.goal[2] = h.wdim == 3 ? goal[2] : 0.0f;
    const uint32_t cap = h.cap;
    const size_t nodes = ((size_t)cap + kNodeAlign - 1) / kNodeAlign * kNodeAlign + kNodeAlign;
    AsyncBuffers a(s);
    QueryDesc* dq  = nullptr;
    QueryState* ds = nullptr;
    int32_t* info  = nullptr;
    Pools P{};
    WorkLists W{};
    if((rc = a.get(dq, 1)) || (rc = a.get(ds, 1)) || (rc = a.get(info, 4)) ||
       (rc = a.get(P.treeS, (size_t)h.D * cap)) || (rc = a.get(P.childR, (size_t)row_words(h.D) * cap)) ||
       (rc = a.get(P.parent, nodes)) || (rc = a.get(P.grand, nodes)) || (rc = a.get(P.cost, nodes)) ||
       (rc = a.get(P.anc, nodes)) ||
       (rc = a.get(P.frontier, nodes)) || (rc = a.get(P.r1node, nodes)) || (rc = a.get(P.r1child, nodes)) ||
       (rc = a.get(P.fnext, nodes)) || (rc = a.get(P.react, nodes)) || (rc = a.get(P.repIdx, nodes)) ||
       (rc = a.get(P.partA, h.nChunks + 1)) || (rc = a.get(P.look, 4 * ((size_t)h.nChunks + 1))) ||
       (rc = a.get(P.validc, h.R1)) || (rc = a.get(P.score, h.R1)) || (rc = a.get(W.tileStart, 4)) ||
       (rc = a.get(W.jobStart, 4)) || (rc = a.get(W.active, 2)) || (rc = a.get(W.counter, 1)) ||
       (rc = a.get(W.claim, 2)) || (rc = a.get(W.allDone, 1)) || (rc = a.get(W.error, 1)) ||
       (rc = a.get(W.closed, 1)) || (rc = a.get(W.closedDev, 1)))
        return rc;
    W.nq   = 1;
    W.tile = kTile;
    HIP_TRY(hipMemcpyAsync(dq, &h, sizeof(QueryDesc), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(P.treeS, d_tree_samples, sizeof(float) * h.D * cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(P.parent, d_parents, sizeof(int32_t) * cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(P.cost, d_costs, sizeof(float) * cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(P.frontier, d_frontier, cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(P.validc, d_valid_count, sizeof(int32_t) * h.R1, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(P.score, d_score, sizeof(float) * h.R1, hipMemcpyDeviceToDevice, s));
    hipError_t e = launch_stage_frontier_update(dq, ds, P, W, itr, tree_size, d_child_samples, d_child_parent,
                                                d_fnext, n_children, info, s);
    if(e != hipSuccess) return fail(KGMT_E_HIP, std::string("frontier update: ") + hipGetErrorString(e));
    HIP_TRY(hipMemcpyAsync(d_tree_samples, P.treeS, sizeof(float) * h.D * cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_parents, P.parent, sizeof(int32_t) * cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_costs, P.cost, sizeof(float) * cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_frontier, P.frontier, cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_rep_idx, P.repIdx, sizeof(uint32_t) * cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_info, info, sizeof(int32_t) * 4, hipMemcpyDeviceToDevice, s));
    uint32_t err = 0;
    HIP_TRY(hipMemcpyAsync(&err, W.error, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));  // h and err are released on return
    if(err) return fail(KGMT_E_HIP, "frontier update: a look-back wait ran out of its bound");
    return KGMT_OK;
}

int kgmt_region_update(const kgmt_config* cfg, const int32_t* d_r1, const int32_t* d_r2, const uint8_t* d_valid,
                       const float* d_ufront, uint32_t n, const float* d_score, const uint32_t* d_active_prev,
                       uint32_t* d_active_next, int32_t* d_total, int32_t* d_valid_count, uint8_t* d_fnext,
                       void* stream)
{
    if(n && (!d_r1 || !d_r2 || !d_valid || !d_ufront || !d_score || !d_active_prev || !d_active_next || !d_total ||
             !d_valid_count || !d_fnext))
        return fail(KGMT_E_ARG, "null buffer");
    StageDesc sd((hipStream_t)stream);
    int rc = sd.make(cfg);
    if(rc) return rc;
    HIP_TRY(launch_stage_region_update(sd.d, d_r1, d_r2, d_valid, d_ufront, n, d_score, d_active_prev, d_active_next,
                                       d_total, d_valid_count, d_fnext, sd.s));
    return KGMT_OK;
}

int kgmt_score(const kgmt_config* cfg, const int32_t* d_total, const int32_t* d_valid_count,
               const uint32_t* d_active, float* d_score, void* stream)
{
    if(!d_total || !d_valid_count || !d_active || !d_score) return fail(KGMT_E_ARG, "null buffer");
    StageDesc sd((hipStream_t)stream);
    int rc = sd.make(cfg);
    if(rc) return rc;
    HIP_TRY(launch_stage_score(sd.d, d_total, d_valid_count, d_active, d_score, sd.s));
    return KGMT_OK;
}

int kgmt_compact(const uint8_t* d_flags, uint32_t n, uint32_t* d_out, uint32_t* d_count, void* stream)
{
    if(!d_count || (n && (!d_flags || !d_out))) return fail(KGMT_E_ARG, "null buffer");
    hipStream_t s = (hipStream_t)stream;
    uint32_t* tmp = nullptr;
    HIP_TRY(hipMallocAsync((void**)&tmp, sizeof(uint32_t) * ((n + kChunk - 1) / kChunk + 1), s));
    hipError_t e = launch_stage_compact(d_flags, n, d_out, d_count, tmp, s);
    (void)hipFreeAsync(tmp, s);
    if(e != hipSuccess) return fail(KGMT_E_HIP, std::string("compact: ") + hipGetErrorString(e));
    return KGMT_OK;
}

int kgmt_repeat(const uint32_t* d_ids, const uint32_t* d_counts, uint32_t n, uint32_t* d_out, uint32_t out_cap,
                uint32_t* d_total, void* stream)
{
    if(!d_total || (n && (!d_ids || !d_counts)) || (out_cap && !d_out)) return fail(KGMT_E_ARG, "null buffer");
    hipStream_t s = (hipStream_t)stream;
    uint32_t* tmp = nullptr;
    HIP_TRY(hipMallocAsync((void**)&tmp, sizeof(uint32_t) * ((n + kChunk - 1) / kChunk + 1), s));
    hipError_t e = launch_stage_repeat(d_ids, d_counts, n, d_out, out_cap, d_total, tmp, s);
    (void)hipFreeAsync(tmp, s);
    if(e != hipSuccess) return fail(KGMT_E_HIP, std::string("repeat: ") + hipGetErrorString(e));
    return KGMT_OK;
}

// ---- trajectory refinement --------------------------------------------------
static int adjoint_common(bool refine, const kgmt_config* cfg, const kgmt_adjoint_params* ap, uint32_t n,
                          uint32_t max_steps, const int32_t* d_steps, const float* d_x0, const float* d_goal, float* d_u,
                          const float* d_obstacles, uint32_t n_obs, float* d_cost, float* d_grad, void* stream)
{
    if(!cfg || !ap) return fail(KGMT_E_ARG, "null argument");
    if(cfg->model != KGMT_MODEL_QUADCOPTER) return fail(KGMT_E_UNSUPPORTED, "refinement is built for the quadcopter");
    if(n && (!d_steps || !d_x0 || !d_goal || !d_u || !d_cost || (!refine && !d_grad)))
        return fail(KGMT_E_ARG, "null buffer");
    if(n_obs > kMaxObs) return fail(KGMT_E_CAPACITY, "more than 1500 obstacles");
    if(n_obs && !d_obstacles) return fail(KGMT_E_ARG, "obstacles pointer is null");
    if(!(ap->tau > 0.0f)) return fail(KGMT_E_ARG, "tau must be positive");
    if(refine && ap->iterations > 100000) return fail(KGMT_E_ARG, "too many iterations");
    if(n == 0) return KGMT_OK;
    StageDesc sd((hipStream_t)stream);
    int rc = sd.make(cfg);
    if(rc) return rc;
    const size_t ck = ((size_t)max_steps + 1) * 12 * n, mv = (size_t)max_steps * 4 * n;
    float *ckpt = nullptr, *mom = nullptr, *vel = nullptr;
    float4* boxes = nullptr;
    AsyncBuffers a(sd.s);  // released on every return path
    if((rc = a.get(ckpt, ck)) || (rc = a.get(boxes, 2 * (size_t)std::max<uint32_t>(n_obs, 1)))) return rc;
    if(refine)
        {
            if((rc = a.get(mom, mv)) || (rc = a.get(vel, mv))) return rc;
            HIP_TRY(hipMemsetAsync(mom, 0, mv * sizeof(float), sd.s));
            HIP_TRY(hipMemsetAsync(vel, 0, mv * sizeof(float), sd.s));
        }
    hipError_t e = launch_pack_boxes(d_obstacles, n_obs, boxes, sd.s);
    if(e == hipSuccess)
        e = launch_adjoint(refine, sd.d, *ap, n, max_steps, d_steps, d_x0, d_goal, d_u, d_grad, ckpt, mom, vel, boxes,
                           n_obs, d_cost, sd.s);
    if(e != hipSuccess) return fail(KGMT_E_HIP, std::string("adjoint: ") + hipGetErrorString(e));
    return KGMT_OK;
}

int kgmt_rk4_adjoint(const kgmt_config* cfg, const kgmt_adjoint_params* ap, uint32_t n, uint32_t max_steps,
                     const int32_t* d_steps, const float* d_x0, const float* d_goal, const float* d_u,
                     const float* d_obstacles, uint32_t n_obs, float* d_cost, float* d_grad, void* stream)
{
    return adjoint_common(false, cfg, ap, n, max_steps, d_steps, d_x0, d_goal, const_cast<float*>(d_u), d_obstacles,
                          n_obs, d_cost, d_grad, stream);
}

int kgmt_refine(const kgmt_config* cfg, const kgmt_adjoint_params* ap, uint32_t n, uint32_t max_steps,
                const int32_t* d_steps, const float* d_x0, const float* d_goal, float* d_u,
                const float* d_obstacles, uint32_t n_obs, float* d_cost_history, void* stream)
{
    return adjoint_common(true, cfg, ap, n, max_steps, d_steps, d_x0, d_goal, d_u, d_obstacles, n_obs,
                          d_cost_history, nullptr, stream);
}

}  // extern "C"
