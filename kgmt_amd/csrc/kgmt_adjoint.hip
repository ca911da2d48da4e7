// kgmt_adjoint.hip — differentiable RK4 rollout of the quadcopter ODE on gfx950.
//
// Refinement of found control sequences (an extension beside the planner):
// one lane per trajectory integrates x_{t+1} = RK4(x_t, u_t) forward, keeping
// the state at every step as a checkpoint, then runs the discrete adjoint
// backwards: each RK4 step's four stages are recomputed from its checkpoint
// and their vector-Jacobian products chained (stage recomputation instead of
// storing 4x12 stage values per step). The cost is
//   J = w_e sum |u_t - u_ref|^2 + w_o sum_{t>=1} sum_b tau softplus((m - sd(p_t, b)) / tau)
//     + w_g |p_T - goal|^2.
// k_adjoint<false> writes J and dJ/du; k_adjoint<true> runs Adam for
// `iterations` steps in place and records J before every step and after the
// last. The dynamics are those of statePropagator.cu:325-370 (ode).
#include "kgmt_launch.hpp"

namespace kgmt
{

// dynamics constants of one configuration (quad parameters, dt)
struct QuadK
{
    float h, hh, h6, h3;
    float g, minv, nu, mu;
    float ax, bx, ay, by, az, bz;
};

// d(ode)/d(state)^T a and d(ode)/d(control)^T a, FP32 (same derivation as
// the oracle's FP64 orc_quad_vjp_f64)
__device__ __forceinline__ void quad_vjp(const QuadK& K, const float* s, const float* a, float* gs, float* gc)
{
    const float u = s[6], v = s[7], w = s[8], p = s[9], q = s[10], r = s[11];
    float sf, cf, st, ct, sp, cp;
    sincos_k(s[3], sf, cf);
    sincos_k(s[4], st, ct);
    sincos_k(s[5], sp, cp);
    const float tt = st / ct, ict = 1.0f / ct;
    const float R00 = ct * cp, R01 = sf * st * cp - cf * sp, R02 = cf * st * cp + sf * sp;
    const float R10 = ct * sp, R11 = sf * st * sp + cf * cp, R12 = cf * st * sp - sf * cp;
    const float R20 = -st, R21 = sf * ct, R22 = cf * ct;
    const float a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5];
    const float a6 = a[6], a7 = a[7], a8 = a[8], a9 = a[9], a10 = a[10], a11 = a[11];
    float g3 = 0.0f, g4 = 0.0f, g5 = 0.0f, g6, g7, g8, g9, g10, g11;
    g6 = R00 * a0 + R10 * a1 + R20 * a2;
    g7 = R01 * a0 + R11 * a1 + R21 * a2;
    g8 = R02 * a0 + R12 * a1 + R22 * a2;
    const float d0f = (cf * st * cp + sf * sp) * v + (-sf * st * cp + cf * sp) * w;
    const float d1f = (cf * st * sp - sf * cp) * v + (-sf * st * sp - cf * cp) * w;
    const float d2f = cf * ct * v - sf * ct * w;
    const float d0t = -st * cp * u + sf * ct * cp * v + cf * ct * cp * w;
    const float d1t = -st * sp * u + sf * ct * sp * v + cf * ct * sp * w;
    const float d2t = -ct * u - sf * st * v - cf * st * w;
    const float f0 = R00 * u + R01 * v + R02 * w, f1 = R10 * u + R11 * v + R12 * w;
    g3 += d0f * a0 + d1f * a1 + d2f * a2;
    g4 += d0t * a0 + d1t * a1 + d2t * a2;
    g5 += -f1 * a0 + f0 * a1;
    const float E = q * sf + r * cf, F = q * cf - r * sf;
    g3 += F * tt * a3 - E * a4 + F * ict * a5;
    g4 += E * ict * ict * a3 + E * st * ict * ict * a5;
    g9  = a3;
    g10 = sf * tt * a3 + cf * a4 + sf * ict * a5;
    g11 = cf * tt * a3 - sf * a4 + cf * ict * a5;
    g7 += r * a6 - p * a8;
    g8 += -q * a6 + p * a7;
    g6 += -r * a7 + q * a8;
    g9 += w * a7 - v * a8;
    g10 += -w * a6 + u * a8;
    g11 += v * a6 - u * a7;
    g4 += -K.g * ct * a6 - K.g * st * sf * a7 - K.g * st * cf * a8;
    g3 += K.g * ct * cf * a7 - K.g * ct * sf * a8;
    const float nv = __builtin_sqrtf(u * u + v * v + w * w);
    const float N  = -K.nu * nv;
    g6 += K.minv * N * a6;
    g7 += K.minv * N * a7;
    g8 += K.minv * N * a8;
    if(nv > 0.0f)
        {
            const float k = K.minv * (-K.nu) / nv * (u * a6 + v * a7 + w * a8);
            g6 += k * u, g7 += k * v, g8 += k * w;
        }
    g10 += K.ax * r * a9;
    g11 += K.ax * q * a9;
    g9 += K.ay * r * a10;
    g11 += K.ay * p * a10;
    g9 += K.az * q * a11;
    g10 += K.az * p * a11;
    const float nw = __builtin_sqrtf(p * p + q * q + r * r);
    const float M  = -K.mu * nw;
    const float b9 = K.bx * a9, b10 = K.by * a10, b11 = K.bz * a11;
    g9 += M * b9, g10 += M * b10, g11 += M * b11;
    if(nw > 0.0f)
        {
            const float k = (-K.mu) / nw * (p * b9 + q * b10 + r * b11);
            g9 += k * p, g10 += k * q, g11 += k * r;
        }
    gs[0] = 0.0f, gs[1] = 0.0f, gs[2] = 0.0f;
    gs[3] = g3, gs[4] = g4, gs[5] = g5, gs[6] = g6, gs[7] = g7, gs[8] = g8;
    gs[9] = g9, gs[10] = g10, gs[11] = g11;
    gc[0] = K.minv * a8;
    gc[1] = K.bx * a9;
    gc[2] = K.by * a10;
    gc[3] = K.bz * a11;
}

__device__ __forceinline__ void rk4_step(const QueryDesc& qd, const QuadK& K, float* x, const float* uc)
{
    float k1[12], k2[12], k3[12], k4[12], st[12];
    quad_ode(qd, x, uc, k1);
#pragma unroll
    for(int j = 0; j < 12; ++j) st[j] = x[j] + K.hh * k1[j];
    quad_ode(qd, st, uc, k2);
#pragma unroll
    for(int j = 0; j < 12; ++j) st[j] = x[j] + K.hh * k2[j];
    quad_ode(qd, st, uc, k3);
#pragma unroll
    for(int j = 0; j < 12; ++j) st[j] = x[j] + K.h * k3[j];
    quad_ode(qd, st, uc, k4);
#pragma unroll
    for(int j = 0; j < 12; ++j) x[j] = x[j] + K.h6 * (k1[j] + 2.0f * k2[j] + 2.0f * k3[j] + k4[j]);
}

// xbar = J_x^T lam, ubar = J_u^T lam of one RK4 step from checkpoint x
__device__ __forceinline__ void rk4_vjp(const QueryDesc& qd, const QuadK& K, const float* x, const float* uc,
                                        const float* lam, float* xbar, float* ubar)
{
    float k[12], s2[12], s3[12], s4[12];
    quad_ode(qd, x, uc, k);
#pragma unroll
    for(int j = 0; j < 12; ++j) s2[j] = x[j] + K.hh * k[j];
    quad_ode(qd, s2, uc, k);
#pragma unroll
    for(int j = 0; j < 12; ++j) s3[j] = x[j] + K.hh * k[j];
    quad_ode(qd, s3, uc, k);
#pragma unroll
    for(int j = 0; j < 12; ++j) s4[j] = x[j] + K.h * k[j];
    float kb[12], gs[12], gc[4];
#pragma unroll
    for(int j = 0; j < 12; ++j) xbar[j] = lam[j], kb[j] = K.h6 * lam[j];
    ubar[0] = ubar[1] = ubar[2] = ubar[3] = 0.0f;
    // stage 4: k4 = f(s4), s4 = x + h k3
    quad_vjp(K, s4, kb, gs, gc);
#pragma unroll
    for(int j = 0; j < 12; ++j) xbar[j] += gs[j], kb[j] = K.h3 * lam[j] + K.h * gs[j];
#pragma unroll
    for(int i = 0; i < 4; ++i) ubar[i] += gc[i];
    // stage 3: k3 = f(s3), s3 = x + h/2 k2
    quad_vjp(K, s3, kb, gs, gc);
#pragma unroll
    for(int j = 0; j < 12; ++j) xbar[j] += gs[j], kb[j] = K.h3 * lam[j] + K.hh * gs[j];
#pragma unroll
    for(int i = 0; i < 4; ++i) ubar[i] += gc[i];
    // stage 2: k2 = f(s2), s2 = x + h/2 k1
    quad_vjp(K, s2, kb, gs, gc);
#pragma unroll
    for(int j = 0; j < 12; ++j) xbar[j] += gs[j], kb[j] = K.h6 * lam[j] + K.hh * gs[j];
#pragma unroll
    for(int i = 0; i < 4; ++i) ubar[i] += gc[i];
    // stage 1: k1 = f(x)
    quad_vjp(K, x, kb, gs, gc);
#pragma unroll
    for(int j = 0; j < 12; ++j) xbar[j] += gs[j];
#pragma unroll
    for(int i = 0; i < 4; ++i) ubar[i] += gc[i];
}

// lanes per trajectory: they run the same rollout (same operations, same
// bits) and split the obstacle penalty's boxes
constexpr int kLanes = 16;

// v + v of the lane the DPP control selects (all 16 lanes of the row active)
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// sum over a row of 16 lanes, the same bits in every lane: lane ^ 1, lane ^ 2
// (quad permutations), then the half-row and row mirrors
__device__ __forceinline__ float group_sum(float v)
{
    static_assert(kLanes == 16, "one DPP row per trajectory");
    v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
    v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
    v = dpp_add<0x141>(v);  // row_half_mirror
    v = dpp_add<0x140>(v);  // row_mirror
    return v;
}

// sum over boxes of tau * softplus((margin - sd(p, box)) / tau), boxes
// sub, sub + kLanes, ... in this lane, summed over the trajectory's lanes;
// with gp, adds scale * gradient with respect to p
__device__ __forceinline__ float penalty(const kgmt_adjoint_params& ap, const float* p, const float4* sObs, int nObs,
                                         float scale, float* gp, int sub)
{
    float total = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    for(int o = sub; o < nObs; o += kLanes)
        {
            const float4 lo4 = sObs[2 * o], hi4 = sObs[2 * o + 1];
            const float lo[3] = {lo4.x, lo4.y, lo4.z}, hi[3] = {hi4.x, hi4.y, hi4.z};
            float qv[3], sgn[3];
#pragma unroll
            for(int d = 0; d < 3; ++d)
                {
                    const float l = lo[d] - p[d], h = p[d] - hi[d];
                    qv[d]  = l > h ? l : h;
                    sgn[d] = l > h ? -1.0f : 1.0f;
                }
            const float ox = qv[0] > 0.0f ? qv[0] : 0.0f, oy = qv[1] > 0.0f ? qv[1] : 0.0f;
            const float oz      = qv[2] > 0.0f ? qv[2] : 0.0f;
            const float outside = __builtin_sqrtf(ox * ox + oy * oy + oz * oz);
            const int am = qv[0] >= qv[1] ? (qv[0] >= qv[2] ? 0 : 2) : (qv[1] >= qv[2] ? 1 : 2);
            const float inside = qv[am] < 0.0f ? qv[am] : 0.0f;
            const float z      = (ap.margin - (outside + inside)) / ap.tau;
            total += ap.tau * (z > 30.0f ? z : log1pf(expf(z)));
            if(gp)
                {
                    const float dsd = -scale / (1.0f + expf(-z));
                    if(outside > 0.0f)
                        {
                            g0 += dsd * (ox / outside) * sgn[0];
                            g1 += dsd * (oy / outside) * sgn[1];
                            g2 += dsd * (oz / outside) * sgn[2];
                        }
                    else
                        {
                            const float gd = dsd * sgn[am];
                            g0 += am == 0 ? gd : 0.0f;
                            g1 += am == 1 ? gd : 0.0f;
                            g2 += am == 2 ? gd : 0.0f;
                        }
                }
        }
    if(gp)
        {
            gp[0] += group_sum(g0);
            gp[1] += group_sum(g1);
            gp[2] += group_sum(g2);
        }
    return group_sum(total);
}

// kLanes lanes per trajectory; every lane of a trajectory computes and writes
// the same values (the same address and value from each lane)
template <bool REFINE>
__global__ __launch_bounds__(64) void k_adjoint(const QueryDesc* __restrict__ qdp, kgmt_adjoint_params ap,
                                                uint32_t n, uint32_t maxT, const int32_t* __restrict__ steps,
                                                const float* __restrict__ x0g, const float* __restrict__ goalg,
                                                float* __restrict__ u, float* __restrict__ grad,
                                                float* __restrict__ ckpt, float* __restrict__ mom,
                                                float* __restrict__ vel, const float4* __restrict__ obs,
                                                uint32_t nObs, float* __restrict__ cost)
{
    extern __shared__ float4 sObs[];
    for(uint32_t k = threadIdx.x; k < 2 * nObs; k += blockDim.x) sObs[k] = obs[k];
    __syncthreads();
    const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / kLanes;
    const int sub    = (int)(threadIdx.x % kLanes);
    if(i >= n) return;  // whole groups of kLanes lanes leave together
    const QueryDesc& qd  = *qdp;
    const kgmt_config& c = qd.cfg;
    QuadK K;
    K.h = c.dt, K.hh = 0.5f * c.dt, K.h6 = c.dt / 6.0f, K.h3 = c.dt / 3.0f;
    K.g = c.gravity, K.minv = qd.qminv, K.nu = c.quad_nu, K.mu = c.quad_mu;
    K.ax = qd.qax, K.bx = qd.qbx, K.ay = qd.qay, K.by = qd.qby, K.az = qd.qaz, K.bz = qd.qbz;
    int T = steps[i];
    if(T < 0) T = 0;
    if(T > (int)maxT) T = (int)maxT;
    const float gx = goalg[i], gy = goalg[(size_t)n + i], gz = goalg[2 * (size_t)n + i];
    const uint32_t iters = REFINE ? ap.iterations : 0;
    float b1t = 1.0f, b2t = 1.0f;
    for(uint32_t it = 0; it <= iters; ++it)
        {
            // forward rollout with checkpoints
            float x[12];
#pragma unroll
            for(int j = 0; j < 12; ++j) x[j] = x0g[(size_t)j * n + i], ckpt[(size_t)j * n + i] = x[j];
            float J = 0.0f;
            for(int t = 0; t < T; ++t)
                {
                    float uc[4];
#pragma unroll
                    for(int k = 0; k < 4; ++k)
                        {
                            uc[k]         = u[((size_t)t * 4 + k) * n + i];
                            const float d = uc[k] - ap.u_ref[k];
                            J += ap.w_effort * d * d;
                        }
                    rk4_step(qd, K, x, uc);
#pragma unroll
                    for(int j = 0; j < 12; ++j) ckpt[(((size_t)t + 1) * 12 + j) * n + i] = x[j];
                    J += ap.w_obstacle * penalty(ap, x, sObs, (int)nObs, 1.0f, nullptr, sub);
                }
            const float ex = x[0] - gx, ey = x[1] - gy, ez = x[2] - gz;
            J += ap.w_goal * (ex * ex + ey * ey + ez * ez);
            if(REFINE)
                {
                    cost[(size_t)it * n + i] = J;
                    if(it == iters) break;
                    b1t *= ap.beta1;
                    b2t *= ap.beta2;
                }
            else
                cost[i] = J;
            // backward: lam = dJ/dx_t
            float lam[12];
#pragma unroll
            for(int j = 0; j < 12; ++j) lam[j] = 0.0f;
            lam[0] = 2.0f * ap.w_goal * ex, lam[1] = 2.0f * ap.w_goal * ey, lam[2] = 2.0f * ap.w_goal * ez;
            if(T >= 1) penalty(ap, x, sObs, (int)nObs, ap.w_obstacle, lam, sub);
            for(int t = T - 1; t >= 0; --t)
                {
                    float xt[12], uc[4], xbar[12], ubar[4];
#pragma unroll
                    for(int j = 0; j < 12; ++j) xt[j] = ckpt[((size_t)t * 12 + j) * n + i];
#pragma unroll
                    for(int k = 0; k < 4; ++k) uc[k] = u[((size_t)t * 4 + k) * n + i];
                    rk4_vjp(qd, K, xt, uc, lam, xbar, ubar);
#pragma unroll
                    for(int k = 0; k < 4; ++k)
                        {
                            const float gk   = ubar[k] + 2.0f * ap.w_effort * (uc[k] - ap.u_ref[k]);
                            const size_t idx = ((size_t)t * 4 + k) * n + i;
                            if(REFINE)
                                {
                                    const float m = ap.beta1 * mom[idx] + (1.0f - ap.beta1) * gk;
                                    const float v = ap.beta2 * vel[idx] + (1.0f - ap.beta2) * gk * gk;
                                    mom[idx]      = m;
                                    vel[idx]      = v;
                                    float un      = uc[k] - ap.lr * (m / (1.0f - b1t)) /
                                                           (__builtin_sqrtf(v / (1.0f - b2t)) + ap.adam_eps);
                                    if(ap.clamp)
                                        un = un < c.ctrl_min[k] ? c.ctrl_min[k] : (un > c.ctrl_max[k] ? c.ctrl_max[k] : un);
                                    u[idx] = un;
                                }
                            else
                                grad[idx] = gk;
                        }
                    if(t >= 1) penalty(ap, xt, sObs, (int)nObs, ap.w_obstacle, xbar, sub);
#pragma unroll
                    for(int j = 0; j < 12; ++j) lam[j] = xbar[j];
                }
        }
    if(!REFINE)
        for(uint32_t t = (uint32_t)T; t < maxT; ++t)
#pragma unroll
            for(int k = 0; k < 4; ++k) grad[((size_t)t * 4 + k) * n + i] = 0.0f;
}

__global__ void k_pack_boxes(const float* __restrict__ rows, uint32_t n, float4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const float* r = rows + 6 * (size_t)i;
    out[2 * i]     = make_float4(r[0], r[1], r[2], 0.0f);
    out[2 * i + 1] = make_float4(r[3], r[4], r[5], 0.0f);
}

hipError_t launch_pack_boxes(const float* rows, uint32_t n, float4* out, hipStream_t s)
{
    if(n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_boxes, dim3((n + 255) / 256), dim3(256), 0, s, rows, n, out);
    return hipGetLastError();
}

hipError_t launch_adjoint(bool refine, const QueryDesc* qd, const kgmt_adjoint_params& ap, uint32_t n, uint32_t maxT,
                          const int32_t* steps, const float* x0, const float* goal, float* u, float* grad, float* ckpt,
                          float* mom, float* vel, const float4* obs, uint32_t nObs, float* cost, hipStream_t s)
{
    if(n == 0) return hipSuccess;
    const dim3 g((uint32_t)(((uint64_t)n * kLanes + 63) / 64)), b(64);
    const size_t lds = (size_t)(nObs ? nObs : 1) * 2 * sizeof(float4);
    if(refine)
        hipLaunchKernelGGL(k_adjoint<true>, g, b, lds, s, qd, ap, n, maxT, steps, x0, goal, u, grad, ckpt, mom, vel, obs,
                           nObs, cost);
    else
        hipLaunchKernelGGL(k_adjoint<false>, g, b, lds, s, qd, ap, n, maxT, steps, x0, goal, u, grad, ckpt, mom, vel,
                           obs, nObs, cost);
    return hipGetLastError();
}

}  // namespace kgmt
