// kgmt_noalloc.hip — the allocation-free variant of the propagate + collide
// stage entry (kgmt_propagate_collide_noalloc in kgmt.h), for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops). Its only scratch, the obstacle scene the kernel stages into LDS,
// lives in a file-scope device array instead of a per-call allocation, and the
// query descriptor travels as a kernel argument instead of a device copy.
#include "kgmt_launch.hpp"

#include <mutex>

namespace kgmt
{

// The largest scene a query may carry (kMaxObs boxes): its group bounds and
// boxes as (lo, hi) float4 pairs, 54 KB. Every call rewrites the scene it
// uses whole (group bounds, boxes and padding), so nothing in it has to start
// at zero. Each GPU holds one copy until the process ends; calls that overlap
// on one GPU share it and race.
constexpr uint32_t kScratchPairs =
  (kMaxObs + kGroup * kGroupBatch - 1) / (kGroup * kGroupBatch) * kGroupBatch * (1 + kGroup);
static __device__ __align__(256) float4 scratch[2 * kScratchPairs];

uint32_t scratch_pairs()
{
    return kScratchPairs;
}

// Device address of `scratch` on the current device, looked up on the first
// call with that device current and kept per device. Host code uses this
// address, never the array's name (in host code only a stand-in).
hipError_t scratch_address(float4** out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex lock;
    static float4* addr[kMaxDevices] = {};
    int dev             = 0;
    const hipError_t e0 = hipGetDevice(&dev);
    if(e0 != hipSuccess) return e0;
    if(dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(lock);
    if(!addr[dev])
        {
            void* p             = nullptr;
            const hipError_t e1 = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if(e1 != hipSuccess) return e1;
            addr[dev] = static_cast<float4*>(p);
        }
    *out = addr[dev];
    return hipSuccess;
}

// kgmt_propagate_collide's kernel with the descriptor passed by value
template <int MODEL>
__global__ __launch_bounds__(kBlock) void k_stage_propagate_by_value(
  const QueryDesc qd, const int32_t* __restrict__ parentIdx, const float* __restrict__ parentStates,
  uint32_t nParents, const float* __restrict__ controls, const int32_t* __restrict__ steps, uint32_t n,
  const float4* __restrict__ obs, uint32_t nGrp, float* child, uint8_t* valid, int32_t* r1o, int32_t* r2o)
{
    extern __shared__ float4 sObs[];
    stage_scene(obs, 0, nGrp, sObs);
    __syncthreads();
    stage_propagate_child<MODEL>(qd, parentIdx, parentStates, nParents, controls, steps, n, sObs, nGrp, child, valid,
                                 r1o, r2o);
}

hipError_t launch_stage_propagate_by_value(int model, const QueryDesc& qd, const int32_t* parentIdx,
                                           const float* parentStates, uint32_t nParents, const float* controls,
                                           const int32_t* steps, uint32_t n, const float4* obs, uint32_t nGrp,
                                           float* child, uint8_t* valid, int32_t* r1, int32_t* r2, hipStream_t s)
{
    if(n == 0) return hipSuccess;
    const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
    const size_t lds = (size_t)(nGrp ? nGrp * (1 + kGroup) : 1) * 2 * sizeof(float4);
#define KGMT_STAGE_PROP(M)                                                                                        \
    hipLaunchKernelGGL(k_stage_propagate_by_value<M>, g, b, lds, s, qd, parentIdx, parentStates, nParents, controls, \
                       steps, n, obs, nGrp, child, valid, r1, r2)
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

}  // namespace kgmt
