/*
 * kgmt_oracle.c — CPU restatement of KGMT's per-iteration expansion path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path. It is loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
 * the product library.
 *
 * Built with -ffp-contract=off so that every float expression is evaluated
 * as written: one IEEE operation per operator, in C's left-to-right order.
 * The HIP kernels are built the same way and follow the same operation order,
 * so the two agree bit for bit on identical inputs.
 *
 * Reference file:line citations are relative to the KGMT repository.
 * Deterministic semantics chosen where the reference is racy or undefined
 * (numbers are the quirk ids used in the design notes):
 *   Q3  subregion "active" flags are read as of iteration start (snapshot);
 *   Q4  an expanded parent's frontier flag is cleared (both launch regimes);
 *   Q5  expansion multiset = {i repeated repeat[i] times, repeat[i] > 0}, ascending i;
 *   Q6  region scores use the closed formula with a fixed summation order;
 *   Q8  the last frontierNext flag is counted;
 *   Q9  appended nodes are clamped to the remaining capacity, and a plan ends
 *       (unsuccessfully) once the tree is full;
 *   Q10 among goal hits of one iteration the smallest tree index wins;
 *   Q11/Q12 randomness is Philox4x32-10 keyed by (seed) with counter
 *       (index, iteration, stream, query) — no per-thread RNG state;
 *   Q16 steps are uniform integers in [steps_min, steps_max].
 */
#include "kgmt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11). Multipliers and Weyl key increments are the     */
/* published constants of that paper.                                        */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for(int round = 0; round < 10; ++round)
        {
            if(round > 0)
                {
                    k0 += PHILOX_W0;
                    k1 += PHILOX_W1;
                }
            uint64_t p0 = (uint64_t)PHILOX_M0 * (uint64_t)c0;
            uint64_t p1 = (uint64_t)PHILOX_M1 * (uint64_t)c2;
            uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
            uint32_t n1 = (uint32_t)p1;
            uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
            uint32_t n3 = (uint32_t)p0;
            c0 = n0, c1 = n1, c2 = n2, c3 = n3;
        }
    out[0] = c0, out[1] = c1, out[2] = c2, out[3] = c3;
}

/* Uniform in (0, 1], like curand_uniform's range: the top 24 bits plus one,
 * times 2^-24 (exact in FP32). */
float orc_u01(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f;
}

/* ------------------------------------------------------------------------ */
/* sin/cos in FP32: Cody–Waite reduction by pi/2 (three-part split whose     */
/* first two parts carry few bits, so k*P1 and k*P2 are exact for |k| < 2^13)*/
/* and Taylor polynomials of sin (to x^9) and cos (to x^12) on [-pi/4, pi/4].*/
/* Truncation error < 2e-9 there, below FP32 resolution.                    */
/* ------------------------------------------------------------------------ */
#define SC_TWO_OVER_PI 0x1.45f306p-1f
#define SC_P1 0x1.92p+0f
#define SC_P2 0x1.fb4p-12f
#define SC_P3 0x1.4442d2p-24f
#define SC_S1 -0x1.555556p-3f  /* -1/3!  */
#define SC_S2 0x1.111112p-7f   /*  1/5!  */
#define SC_S3 -0x1.a01a02p-13f /* -1/7!  */
#define SC_S4 0x1.71de3ap-19f  /*  1/9!  */
#define SC_C1 0x1.555556p-5f   /*  1/4!  */
#define SC_C2 -0x1.6c16c2p-10f /* -1/6!  */
#define SC_C3 0x1.a01a02p-16f  /*  1/8!  */
#define SC_C4 -0x1.27e4fcp-22f /* -1/10! */
#define SC_C5 0x1.1eed8ep-29f  /*  1/12! */

void orc_sincosf(float x, float* s, float* c)
{
    float fk = rintf(x * SC_TWO_OVER_PI);
    float r  = x - fk * SC_P1;
    r        = r - fk * SC_P2;
    r        = r - fk * SC_P3;
    int q    = (int)fk;
    float z  = r * r;
    float ps = SC_S4;
    ps       = ps * z + SC_S3;
    ps       = ps * z + SC_S2;
    ps       = ps * z + SC_S1;
    float sr = r + (r * z) * ps;
    float pc = SC_C5;
    pc       = pc * z + SC_C4;
    pc       = pc * z + SC_C3;
    pc       = pc * z + SC_C2;
    pc       = pc * z + SC_C1;
    float cr = (1.0f - 0.5f * z) + (z * z) * pc;
    switch(q & 3)
        {
            case 0: *s = sr, *c = cr; break;
            case 1: *s = cr, *c = -sr; break;
            case 2: *s = -sr, *c = -cr; break;
            default: *s = -cr, *c = sr; break;
        }
}

/* ------------------------------------------------------------------------ */
/* model dimensions                                                          */
/* ------------------------------------------------------------------------ */
static int dims(int model, int* S, int* C)
{
    switch(model)
        {
            case KGMT_MODEL_DOUBLE_INTEGRATOR: *S = 6, *C = 3; return 0;
            case KGMT_MODEL_DUBINS_AIRPLANE: *S = 6, *C = 3; return 0;
            case KGMT_MODEL_QUADCOPTER: *S = 12, *C = 4; return 0;
            case KGMT_MODEL_DUBINS_5D: *S = 5, *C = 3; return 0;
            case KGMT_MODEL_UNICYCLE: *S = 4, *C = 2; return 0;
            default: return -1;
        }
}

/* workspace dimension W_DIM: the unicycle is planar (x y), the others move in
 * x y z */
int orc_wdim(const kgmt_config* cfg)
{
    return cfg->model == KGMT_MODEL_UNICYCLE ? 2 : 3;
}

/* ------------------------------------------------------------------------ */
/* per-child draws. Reference draws: statePropagator.cu:78-81 (DI),          */
/* :149-152 (dubins), :248-253 (quad); frontier draw KGMT.cu:254,294;        */
/* re-activation draw KGMT.cu:371-372. Here: block 0 = Philox(u, itr, 0, q), */
/* block 1 = Philox(u, itr, 1, q); words d[0..7]. Control i <- d[i] (sample  */
/* control order), steps <- d[C], frontier uniform <- d[7].                  */
/* ------------------------------------------------------------------------ */
static void draw_words(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t idx, uint32_t stream,
                       uint32_t out[4])
{
    uint32_t ctr[4] = {idx, itr, stream, query};
    uint32_t key[2] = {(uint32_t)cfg->seed, (uint32_t)(cfg->seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

void orc_draw_child(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t u, float* ctrl, int32_t* steps,
                    float* ufront)
{
    int S, C;
    if(dims(cfg->model, &S, &C)) return;
    uint32_t d[8];
    draw_words(cfg, query, itr, u, 0, d);
    draw_words(cfg, query, itr, u, 1, d + 4);
    for(int i = 0; i < C; ++i)
        {
            float range = cfg->ctrl_max[i] - cfg->ctrl_min[i];
            ctrl[i]     = cfg->ctrl_min[i] + orc_u01(d[i]) * range;
        }
    int32_t n = cfg->steps_max - cfg->steps_min + 1;
    int32_t k = (int32_t)(orc_u01(d[C]) * (float)n);
    if(k > n - 1) k = n - 1;
    *steps  = cfg->steps_min + k;
    *ufront = orc_u01(d[7]);
}

float orc_draw_react(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t node)
{
    uint32_t d[4];
    draw_words(cfg, query, itr, node, 2, d);
    return orc_u01(d[0]);
}

/* ------------------------------------------------------------------------ */
/* AABB segment test: collisionCheck.cu:3-10 (isBroadPhaseValid) and :49-62  */
/* (isMotionValid); bounding box built as at statePropagator.cu:121-133.     */
/* obs rows: xmin ymin zmin xmax ymax zmax.                                  */
/* ------------------------------------------------------------------------ */
int orc_segment_valid(const float* p0, const float* p1, const float* obs, uint32_t n_obs)
{
    float bbMin[3], bbMax[3];
    for(int d = 0; d < 3; ++d)
        {
            if(p0[d] > p1[d])
                bbMin[d] = p1[d], bbMax[d] = p0[d];
            else
                bbMin[d] = p0[d], bbMax[d] = p1[d];
        }
    for(uint32_t o = 0; o < n_obs; ++o)
        {
            const float* b = obs + 6 * (size_t)o;
            int separated  = 0;
            for(int d = 0; d < 3; ++d)
                if(bbMax[d] <= b[d] || b[3 + d] <= bbMin[d]) separated = 1;
            if(!separated) return 0;
        }
    return 1;
}

/* The same test with W_DIM = 2: rows xmin ymin xmax ymax. */
int orc_segment_valid2(const float* p0, const float* p1, const float* obs, uint32_t n_obs)
{
    float bbMin[2], bbMax[2];
    for(int d = 0; d < 2; ++d)
        {
            if(p0[d] > p1[d])
                bbMin[d] = p1[d], bbMax[d] = p0[d];
            else
                bbMin[d] = p0[d], bbMax[d] = p1[d];
        }
    for(uint32_t o = 0; o < n_obs; ++o)
        {
            const float* b = obs + 4 * (size_t)o;
            int separated  = 0;
            for(int d = 0; d < 2; ++d)
                if(bbMax[d] <= b[d] || b[2 + d] <= bbMin[d]) separated = 1;
            if(!separated) return 0;
        }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* regions: getRegion Graph.cu:99-153 and getSubRegion :155-205 for a        */
/* workspace-only grid (last axis fastest) over W_DIM axes. Region minimum   */
/* corner as computed by initializeRegions_kernel Graph.cu:58-70:            */
/* W_MIN + index * W_R1_SIZE.                                                */
/* ------------------------------------------------------------------------ */
/* (int)f clamped to [0, L-1], as Graph.cu:107-109. The clamp is applied in
 * float first so that NaN, infinities and |f| >= 2^31 (undefined for a C
 * cast) give the same answer on every target; finite in-range values give
 * the reference's result. */
static int grid_index(float f, int L)
{
    if(!(f >= 0.0f)) return 0;
    if(f >= (float)L) return L - 1;
    int i = (int)f;
    return i >= L ? L - 1 : i;
}

int orc_num_regions(const kgmt_config* cfg, int32_t* r1, int32_t* r2, int32_t* per_r1)
{
    int32_t L = cfg->r1_len, L2 = cfg->r2_len;
    *r1 = 1, *per_r1 = 1;
    for(int i = 0; i < orc_wdim(cfg); ++i) *r1 *= L, *per_r1 *= L2;
    *r2 = *r1 * *per_r1;
    return 0;
}

int orc_region(const kgmt_config* cfg, const float* pos)
{
    int L = cfg->r1_len, region = 0, factor = 1;
    for(int i = orc_wdim(cfg) - 1; i >= 0; --i)
        {
            int index = grid_index(((float)L * (pos[i] - cfg->ws_min)) / (cfg->ws_max - cfg->ws_min), L);
            region += factor * index;
            factor *= L;
        }
    return region;
}

int orc_subregion(const kgmt_config* cfg, const float* pos, int r1)
{
    int L = cfg->r1_len, L2 = cfg->r2_len, W = orc_wdim(cfg);
    float size = (cfg->ws_max - cfg->ws_min) / (float)L;
    int idx[3], t = r1;
    for(int i = W - 1; i >= 0; --i)
        {
            idx[i] = t % L;
            t /= L;
        }
    int local = 0, factor = 1;
    for(int i = W - 1; i >= 0; --i)
        {
            float mn  = cfg->ws_min + (float)idx[i] * size;
            int index = grid_index(((float)L2 * (pos[i] - mn)) / size, L2);
            local += factor * index;
            factor *= L2;
        }
    return r1 * factor + local;
}

/* ------------------------------------------------------------------------ */
/* dynamics (FP32, reference operation order)                                */
/* ------------------------------------------------------------------------ */
static int ws_out(const kgmt_config* cfg, float x, float y, float z)
{
    return x < cfg->ws_min || x > cfg->ws_max || y < cfg->ws_min || y > cfg->ws_max || z < cfg->ws_min ||
           z > cfg->ws_max;
}

/* statePropagator.cu:13-71: Euler steps of the planar unicycle, workspace
 * check on x y, then the planar segment test. tan(steering) is sin/cos of
 * the sincos above (the reference calls tan; one correctly rounded quotient
 * of two faithfully rounded values). */
static int prop_unicycle(const kgmt_config* cfg, const float* x0, const float* ctrl, int32_t steps, const float* obs,
                         uint32_t n_obs, float* x1)
{
    float h = cfg->dt, a = ctrl[0], steer = ctrl[1];
    float x = x0[0], y = x0[1], theta = x0[2], v = x0[3];
    float ss, cs;
    orc_sincosf(steer, &ss, &cs);
    float tn  = ss / cs;
    int valid = 1;
    for(int i = 0; i < steps; ++i)
        {
            float p0[2] = {x, y};
            float st, ct;
            orc_sincosf(theta, &st, &ct);
            x     = x + (v * ct) * h;
            y     = y + (v * st) * h;
            theta = theta + ((v / cfg->uni_length) * tn) * h;
            v     = v + a * h;
            if(x < cfg->ws_min || x > cfg->ws_max || y < cfg->ws_min || y > cfg->ws_max)
                {
                    valid = 0;
                    break;
                }
            float p1[2] = {x, y};
            if(!orc_segment_valid2(p0, p1, obs, n_obs))
                {
                    valid = 0;
                    break;
                }
        }
    x1[0] = x, x1[1] = y, x1[2] = theta, x1[3] = v;
    x1[4] = a, x1[5] = steer;
    x1[6] = cfg->dt * (float)steps;
    return valid;
}

/* statePropagator.cu:76-142. Reference-literal position update (Q1) or exact. */
static int prop_di(const kgmt_config* cfg, const float* x0, const float* ctrl, int32_t steps, const float* obs,
                   uint32_t n_obs, float* x1)
{
    float h = cfg->dt, ax = ctrl[0], ay = ctrl[1], az = ctrl[2];
    float x = x0[0], y = x0[1], z = x0[2], vx = x0[3], vy = x0[4], vz = x0[5];
    int valid = 1;
    for(int i = 0; i < steps; ++i)
        {
            float p0[3] = {x, y, z};
            if(cfg->integrator == KGMT_DI_EXACT)
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
            if(vx < cfg->v_min || vx > cfg->v_max || vy < cfg->v_min || vy > cfg->v_max || vz < cfg->v_min ||
               vz > cfg->v_max)
                {
                    valid = 0;
                    break;
                }
            if(ws_out(cfg, x, y, z))
                {
                    valid = 0;
                    break;
                }
            float p1[3] = {x, y, z};
            if(!orc_segment_valid(p0, p1, obs, n_obs))
                {
                    valid = 0;
                    break;
                }
        }
    x1[0] = x, x1[1] = y, x1[2] = z, x1[3] = vx, x1[4] = vy, x1[5] = vz;
    x1[6] = ax, x1[7] = ay, x1[8] = az;
    x1[9] = cfg->dt * (float)steps;
    return valid;
}

/* sin/cos of angle + increment from sin/cos of the angle (s0, c0) and of the
 * increment (rs, rc): angle addition, the second product rounded, then one
 * fused multiply-add (C99 fmaf, correctly rounded like the GPU's v_fma_f32) */
static void rotate(float s0, float c0, float rs, float rc, float* s, float* c)
{
    *s = fmaf(s0, rc, c0 * rs);
    *c = fmaf(c0, rc, -(s0 * rs));
}

/* statePropagator.cu:147-241 (6D: speed v is a state, a its control) and the
 * 5D variant (speed is a control, a = 0, no speed bound). The reference takes
 * sin/cos of the start, mid (+h/2 rate) and end (+h rate) angles of every
 * step; here yaw and pitch advance by constant increments within the edge,
 * so the mid and end values are the step-start values rotated by sin/cos of
 * the increments (evaluated once per edge), and the end values start the
 * next step. */
static int prop_dubins(const kgmt_config* cfg, const float* x0, const float* ctrl, int32_t steps, const float* obs,
                       uint32_t n_obs, float* x1, int five_d)
{
    float h  = cfg->dt;
    float hh = 0.5f * h;
    float yr = ctrl[0], pr = ctrl[1];
    float a  = five_d ? 0.0f : ctrl[2];
    float x = x0[0], y = x0[1], z = x0[2], yaw = x0[3], pitch = x0[4];
    float v   = five_d ? ctrl[2] : x0[5];
    int valid = 1;
    float sp0, cp0, sy0, cy0, phs, phc, pes, pec, yhs, yhc, yes, yec;
    orc_sincosf(pitch, &sp0, &cp0);
    orc_sincosf(yaw, &sy0, &cy0);
    orc_sincosf(hh * pr, &phs, &phc);
    orc_sincosf(h * pr, &pes, &pec);
    orc_sincosf(hh * yr, &yhs, &yhc);
    orc_sincosf(h * yr, &yes, &yec);
    for(int i = 0; i < steps; ++i)
        {
            float p0[3] = {x, y, z};
            /* fused multiply-adds (C99 fmaf: one rounding) at fixed places */
            float vm = fmaf(hh, a, v);
            float ve = fmaf(h, a, v);
            float spm, cpm, sym, cym, spe, cpe, sye, cye;
            rotate(sp0, cp0, phs, phc, &spm, &cpm);
            rotate(sy0, cy0, yhs, yhc, &sym, &cym);
            rotate(sp0, cp0, pes, pec, &spe, &cpe);
            rotate(sy0, cy0, yes, yec, &sye, &cye);
            float mx = vm * cpm * cym;
            float my = vm * cpm * sym;
            float mz = vm * spm;
            float h6 = h / 6.0f;
            x        = fmaf(h6, fmaf(ve * cpe, cye, fmaf(2.0f, mx + mx, v * cp0 * cy0)), x);
            y        = fmaf(h6, fmaf(ve * cpe, sye, fmaf(2.0f, my + my, v * cp0 * sy0)), y);
            z        = fmaf(h6, fmaf(ve, spe, fmaf(2.0f, mz + mz, v * sp0)), z);
            yaw      = fmaf(h, yr, yaw);
            pitch    = fmaf(h, pr, pitch);
            v        = fmaf(h6, fmaf(2.0f, a + a, a) + a, v);
            sp0 = spe, cp0 = cpe, sy0 = sye, cy0 = cye;
            if(!five_d && (v < cfg->v_min || v > cfg->v_max))
                {
                    valid = 0;
                    break;
                }
            if(pitch < cfg->pitch_min || pitch > cfg->pitch_max)
                {
                    valid = 0;
                    break;
                }
            if(ws_out(cfg, x, y, z))
                {
                    valid = 0;
                    break;
                }
            float p1[3] = {x, y, z};
            if(!orc_segment_valid(p0, p1, obs, n_obs))
                {
                    valid = 0;
                    break;
                }
        }
    x1[0] = x, x1[1] = y, x1[2] = z, x1[3] = yaw, x1[4] = pitch;
    if(five_d)
        {
            x1[5] = yr, x1[6] = pr, x1[7] = v;
            x1[8] = cfg->dt * (float)steps;
        }
    else
        {
            x1[5] = v;
            x1[6] = yr, x1[7] = pr, x1[8] = a;
            x1[9] = cfg->dt * (float)steps;
        }
    return valid;
}

/* ode, statePropagator.cu:325-370 */
typedef struct quad_k {
    float minv, ax, bx, ay, by, az, bz; /* 1/m, (Iy-Iz)/Ix, 1/Ix, ... */
} quad_k;

static void quad_consts(const kgmt_config* cfg, quad_k* k)
{
    float ix = cfg->quad_ix, iy = cfg->quad_iy, iz = cfg->quad_iz;
    k->minv = 1.0f / cfg->quad_mass;
    k->ax = (iy - iz) / ix, k->bx = 1.0f / ix;
    k->ay = (iz - ix) / iy, k->by = 1.0f / iy;
    k->az = (ix - iy) / iz, k->bz = 1.0f / iz;
}

static void quad_ode(const kgmt_config* cfg, const quad_k* K, const float* s, const float* u4, float* xd)
{
    float phi = s[3], theta = s[4], psi = s[5];
    float u = s[6], v = s[7], w = s[8], p = s[9], q = s[10], r = s[11];
    float sph, cph, sth, cth, sps, cps;
    orc_sincosf(phi, &sph, &cph);
    orc_sincosf(theta, &sth, &cth);
    orc_sincosf(psi, &sps, &cps);
    float tth = sth / cth;
    float g   = cfg->gravity;
    xd[0]     = cth * cps * u + (sph * sth * cps - cph * sps) * v + (cph * sth * cps + sph * sps) * w;
    xd[1]     = cth * sps * u + (sph * sth * sps + cph * cps) * v + (cph * sth * sps - sph * cps) * w;
    xd[2]     = -sth * u + sph * cth * v + cph * cth * w;
    xd[3]     = p + (q * sph + r * cph) * tth;
    xd[4]     = q * cph - r * sph;
    xd[5]     = (q * sph + r * cph) / cth;
    float XYZ = -cfg->quad_nu * sqrtf(u * u + v * v + w * w);
    xd[6]     = (r * v - q * w) - g * sth + K->minv * XYZ * u;
    xd[7]     = (p * w - r * u) + g * cth * sph + K->minv * XYZ * v;
    xd[8]     = (q * u - p * v) + g * cth * cph + K->minv * XYZ * w + K->minv * u4[0];
    float LMN = -cfg->quad_mu * sqrtf(p * p + q * q + r * r);
    xd[9]     = K->ax * q * r + K->bx * LMN * p + K->bx * u4[1];
    xd[10]    = K->ay * p * r + K->by * LMN * q + K->by * u4[2];
    xd[11]    = K->az * p * q + K->bz * LMN * r + K->bz * u4[3];
}

/* statePropagator.cu:246-323 */
static int prop_quad(const kgmt_config* cfg, const float* x0, const float* ctrl, int32_t steps, const float* obs,
                     uint32_t n_obs, float* x1)
{
    quad_k K;
    quad_consts(cfg, &K);
    float h = cfg->dt, hh = 0.5f * h;
    float s[12], k1[12], k2[12], k3[12], k4[12], st[12];
    memcpy(s, x0, sizeof(s));
    int valid = 1;
    for(int i = 0; i < steps; ++i)
        {
            float p0[3] = {s[0], s[1], s[2]};
            quad_ode(cfg, &K, s, ctrl, k1);
            for(int j = 0; j < 12; ++j) st[j] = s[j] + hh * k1[j];
            quad_ode(cfg, &K, st, ctrl, k2);
            for(int j = 0; j < 12; ++j) st[j] = s[j] + hh * k2[j];
            quad_ode(cfg, &K, st, ctrl, k3);
            for(int j = 0; j < 12; ++j) st[j] = s[j] + h * k3[j];
            quad_ode(cfg, &K, st, ctrl, k4);
            for(int j = 0; j < 12; ++j) s[j] = s[j] + (h / 6.0f) * (k1[j] + 2.0f * k2[j] + 2.0f * k3[j] + k4[j]);
            if(s[6] < cfg->v_min || s[6] > cfg->v_max || s[7] < cfg->v_min || s[7] > cfg->v_max || s[8] < cfg->v_min ||
               s[8] > cfg->v_max)
                {
                    valid = 0;
                    break;
                }
            if(ws_out(cfg, s[0], s[1], s[2]))
                {
                    valid = 0;
                    break;
                }
            float p1[3] = {s[0], s[1], s[2]};
            if(!orc_segment_valid(p0, p1, obs, n_obs))
                {
                    valid = 0;
                    break;
                }
        }
    memcpy(x1, s, sizeof(s));
    x1[12] = ctrl[0], x1[13] = ctrl[1], x1[14] = ctrl[2], x1[15] = ctrl[3];
    x1[16] = cfg->dt * (float)steps;
    return valid;
}

int orc_propagate(const kgmt_config* cfg, const float* x0, const float* ctrl, int32_t steps, const float* obs,
                  uint32_t n_obs, float* x1)
{
    switch(cfg->model)
        {
            case KGMT_MODEL_DOUBLE_INTEGRATOR: return prop_di(cfg, x0, ctrl, steps, obs, n_obs, x1);
            case KGMT_MODEL_DUBINS_AIRPLANE: return prop_dubins(cfg, x0, ctrl, steps, obs, n_obs, x1, 0);
            case KGMT_MODEL_DUBINS_5D: return prop_dubins(cfg, x0, ctrl, steps, obs, n_obs, x1, 1);
            case KGMT_MODEL_QUADCOPTER: return prop_quad(cfg, x0, ctrl, steps, obs, n_obs, x1);
            case KGMT_MODEL_UNICYCLE: return prop_unicycle(cfg, x0, ctrl, steps, obs, n_obs, x1);
            default: return -1;
        }
}

/* ------------------------------------------------------------------------ */
/* FP64 twin of the dynamics (libm sin/cos, double arithmetic): the pin of   */
/* the FP32 restatement. Same step structure and checks.                     */
/* ------------------------------------------------------------------------ */
static int seg_valid_f64(const double* p0, const double* p1, const double* obs, uint32_t n_obs, int W)
{
    for(uint32_t o = 0; o < n_obs; ++o)
        {
            const double* b = obs + 2 * W * (size_t)o;
            int separated   = 0;
            for(int d = 0; d < W; ++d)
                {
                    double lo = p0[d] < p1[d] ? p0[d] : p1[d];
                    double hi = p0[d] < p1[d] ? p1[d] : p0[d];
                    if(hi <= b[d] || b[W + d] <= lo) separated = 1;
                }
            if(!separated) return 0;
        }
    return 1;
}

static int ws_out_f64(const kgmt_config* cfg, const double* p, int W)
{
    for(int d = 0; d < W; ++d)
        if(p[d] < cfg->ws_min || p[d] > cfg->ws_max) return 1;
    return 0;
}

void orc_quad_ode_f64(const kgmt_config* cfg, const double* s, const double* u4, double* xd)
{
    double phi = s[3], theta = s[4], psi = s[5];
    double u = s[6], v = s[7], w = s[8], p = s[9], q = s[10], r = s[11];
    double sph = sin(phi), cph = cos(phi), sth = sin(theta), cth = cos(theta), sps = sin(psi), cps = cos(psi);
    double ix = cfg->quad_ix, iy = cfg->quad_iy, iz = cfg->quad_iz, g = cfg->gravity, minv = 1.0 / cfg->quad_mass;
    xd[0]      = cth * cps * u + (sph * sth * cps - cph * sps) * v + (cph * sth * cps + sph * sps) * w;
    xd[1]      = cth * sps * u + (sph * sth * sps + cph * cps) * v + (cph * sth * sps - sph * cps) * w;
    xd[2]      = -sth * u + sph * cth * v + cph * cth * w;
    xd[3]      = p + (q * sph + r * cph) * tan(theta);
    xd[4]      = q * cph - r * sph;
    xd[5]      = (q * sph + r * cph) / cth;
    double XYZ = -(double)cfg->quad_nu * sqrt(u * u + v * v + w * w);
    xd[6]      = (r * v - q * w) - g * sth + minv * XYZ * u;
    xd[7]      = (p * w - r * u) + g * cth * sph + minv * XYZ * v;
    xd[8]      = (q * u - p * v) + g * cth * cph + minv * XYZ * w + minv * u4[0];
    double LMN = -(double)cfg->quad_mu * sqrt(p * p + q * q + r * r);
    xd[9]      = (iy - iz) / ix * q * r + (1.0 / ix) * LMN * p + (1.0 / ix) * u4[1];
    xd[10]     = (iz - ix) / iy * p * r + (1.0 / iy) * LMN * q + (1.0 / iy) * u4[2];
    xd[11]     = (ix - iy) / iz * p * q + (1.0 / iz) * LMN * r + (1.0 / iz) * u4[3];
}

int orc_propagate_f64(const kgmt_config* cfg, const double* x0, const double* ctrl, int32_t steps, const double* obs,
                      uint32_t n_obs, double* x1)
{
    double h  = (double)cfg->dt;
    int valid = 1, S, C, W = orc_wdim(cfg);
    if(dims(cfg->model, &S, &C)) return -1;
    double s[12];
    memcpy(s, x0, sizeof(double) * (size_t)S);
    for(int i = 0; i < steps; ++i)
        {
            double p0[3] = {s[0], s[1], s[2]};
            int bad      = 0;
            if(cfg->model == KGMT_MODEL_UNICYCLE)
                {
                    double v = s[3], th = s[2];
                    s[0] += v * cos(th) * h;
                    s[1] += v * sin(th) * h;
                    s[2] += v / (double)cfg->uni_length * tan(ctrl[1]) * h;
                    s[3] = v + ctrl[0] * h;
                }
            else if(cfg->model == KGMT_MODEL_DOUBLE_INTEGRATOR)
                {
                    for(int d = 0; d < 3; ++d)
                        {
                            double a = ctrl[d], v = s[3 + d];
                            if(cfg->integrator == KGMT_DI_EXACT)
                                s[d] += v * h + a * h * h / 2.0;
                            else
                                s[d] += (5.0 * v + 2.0 * a * h) * h / 6.0;
                            s[3 + d] = v + a * h;
                        }
                    for(int d = 0; d < 3; ++d)
                        if(s[3 + d] < cfg->v_min || s[3 + d] > cfg->v_max) bad = 1;
                }
            else if(cfg->model == KGMT_MODEL_DUBINS_AIRPLANE || cfg->model == KGMT_MODEL_DUBINS_5D)
                {
                    int five   = cfg->model == KGMT_MODEL_DUBINS_5D;
                    double yr  = ctrl[0], pr = ctrl[1], a = five ? 0.0 : ctrl[2];
                    double v   = five ? ctrl[2] : s[5];
                    double yaw = s[3], pitch = s[4];
                    double vm = v + 0.5 * h * a, pm = pitch + 0.5 * h * pr, ym = yaw + 0.5 * h * yr;
                    double ve = v + h * a, pe = pitch + h * pr, ye = yaw + h * yr;
                    s[0] += h / 6.0 * (v * cos(pitch) * cos(yaw) + 4.0 * vm * cos(pm) * cos(ym) + ve * cos(pe) * cos(ye));
                    s[1] += h / 6.0 * (v * cos(pitch) * sin(yaw) + 4.0 * vm * cos(pm) * sin(ym) + ve * cos(pe) * sin(ye));
                    s[2] += h / 6.0 * (v * sin(pitch) + 4.0 * vm * sin(pm) + ve * sin(pe));
                    s[3] = yaw + h * yr;
                    s[4] = pitch + h * pr;
                    if(!five)
                        {
                            s[5] = v + h * a;
                            if(s[5] < cfg->v_min || s[5] > cfg->v_max) bad = 1;
                        }
                    if(!bad && (s[4] < cfg->pitch_min || s[4] > cfg->pitch_max)) bad = 1;
                }
            else
                {
                    double k1[12], k2[12], k3[12], k4[12], st[12];
                    orc_quad_ode_f64(cfg, s, ctrl, k1);
                    for(int j = 0; j < 12; ++j) st[j] = s[j] + 0.5 * h * k1[j];
                    orc_quad_ode_f64(cfg, st, ctrl, k2);
                    for(int j = 0; j < 12; ++j) st[j] = s[j] + 0.5 * h * k2[j];
                    orc_quad_ode_f64(cfg, st, ctrl, k3);
                    for(int j = 0; j < 12; ++j) st[j] = s[j] + h * k3[j];
                    orc_quad_ode_f64(cfg, st, ctrl, k4);
                    for(int j = 0; j < 12; ++j) s[j] += h / 6.0 * (k1[j] + 2.0 * k2[j] + 2.0 * k3[j] + k4[j]);
                    for(int d = 6; d < 9; ++d)
                        if(s[d] < cfg->v_min || s[d] > cfg->v_max) bad = 1;
                }
            if(!bad && ws_out_f64(cfg, s, W)) bad = 1;
            if(!bad && !seg_valid_f64(p0, s, obs, n_obs, W)) bad = 1;
            if(bad)
                {
                    valid = 0;
                    break;
                }
        }
    memcpy(x1, s, sizeof(double) * (size_t)S);
    for(int i = 0; i < C; ++i) x1[S + i] = ctrl[i];
    x1[S + C] = h * steps;
    return valid;
}

/* ------------------------------------------------------------------------ */
/* region bookkeeping of propagateFrontier_kernel1/2 (KGMT.cu:249-256,       */
/* 289-296) with snapshot semantics for the subregion flag (Q3).             */
/* ------------------------------------------------------------------------ */
void orc_region_update(const kgmt_config* cfg, const int32_t* r1, const int32_t* r2, const uint8_t* valid,
                       const float* ufront, uint32_t n, const float* score, const uint8_t* active_prev,
                       uint8_t* active_next, int32_t* total, int32_t* valid_count, uint8_t* fnext)
{
    (void)cfg;
    for(uint32_t i = 0; i < n; ++i)
        {
            total[r1[i]] += 1;
            fnext[i] = 0;
            if(valid[i])
                {
                    valid_count[r1[i]] += 1;
                    fnext[i]          = (ufront[i] < score[r1[i]]) || !active_prev[r2[i]];
                    active_next[r2[i]] = 1;
                }
        }
}

/* ------------------------------------------------------------------------ */
/* scores: Graph.cu:239-257 (per-region score), :297-304 / :349-357          */
/* (normalisation). Q6: formula over all regions; fv^4 = (fv*fv)*(fv*fv);    */
/* the sum of scores is a pairwise tree over regions padded with zeros to a  */
/* power of two: at stride s = P/2, P/4, ..., 1: a[i] += a[i + s], i < s.    */
/* ------------------------------------------------------------------------ */
void orc_score(const kgmt_config* cfg, const int32_t* total, const int32_t* valid_count, const uint8_t* active,
               float* score)
{
    int32_t R1, R2, P;
    orc_num_regions(cfg, &R1, &R2, &P);
    float size = (cfg->ws_max - cfg->ws_min) / (float)cfg->r1_len;
    /* W_R1_VOL: the region's measure, W_R1_SIZE^W_DIM (config.h:251 for W_DIM 3) */
    float vol = orc_wdim(cfg) == 2 ? size * size : size * size * size;
    float eps = cfg->epsilon;
    int32_t P2 = 1;
    while(P2 < R1) P2 <<= 1;
    float* a = (float*)calloc((size_t)P2, sizeof(float));
    for(int32_t r = 0; r < R1; ++r)
        {
            if(valid_count[r] > 0)
                {
                    int32_t cnt = 0;
                    for(int32_t k = 0; k < P; ++k) cnt += active[(size_t)r * P + k] ? 1 : 0;
                    float cov = (float)cnt / (float)P;
                    float vf  = (float)valid_count[r];
                    float fv  = (eps + vf) / ((eps + vf) + (float)(total[r] - valid_count[r])) * vol;
                    float f2  = fv * fv;
                    float f4  = f2 * f2;
                    float t   = (float)total[r];
                    a[r]      = f4 / ((1.0f + cov) * (1.0f + t * t));
                }
            score[r] = a[r];
        }
    for(int32_t s = P2 / 2; s >= 1; s >>= 1)
        for(int32_t i = 0; i < s; ++i) a[i] = a[i] + a[i + s];
    float sum = a[0];
    for(int32_t r = 0; r < R1; ++r) score[r] = valid_count[r] == 0 ? 1.0f : eps + score[r] / sum;
    free(a);
}

/* KGMT.cu:387-388 with the clamped accepted count (Q9). The reference's
 * host computes (h_itr_ * EPSILON) in float times pow(treeAddSize, 5) in
 * double (std::pow of a float promotes to double) and stores the product as
 * float: here the fifth power is (a^2)^2 * a in double (a^2 is exact) and the
 * product is rounded to float once. */
float orc_faccept(const kgmt_config* cfg, uint32_t itr, uint32_t tree_size, uint32_t next)
{
    float a   = 1.0f - (float)(tree_size + next) / (float)cfg->capacity;
    double ad = (double)a;
    double a2 = ad * ad;
    double a5 = (a2 * a2) * ad;
    return (float)((double)((float)itr * cfg->epsilon) * a5);
}

/* ------------------------------------------------------------------------ */
/* full planner: KGMT::plan KGMT.cu:41-105, propagateFrontier :167-211,      */
/* propagateFrontier_kernel1/2 :218-299, Graph::updateVertices               */
/* Graph.cu:207-227, updateFrontier :380-403 and updateFrontier_kernel       */
/* :305-378.                                                                 */
/* ------------------------------------------------------------------------ */
/* helper.cuh:138-152: Euclidean distance of the W_DIM position coordinates */
static float pos_dist(const kgmt_config* cfg, const float* a, const float* b)
{
    float dx = a[0] - b[0], dy = a[1] - b[1];
    if(orc_wdim(cfg) == 2) return sqrtf(dx * dx + dy * dy);
    float dz = a[2] - b[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

/* updateFrontier (KGMT.cu:380-403) and updateFrontier_kernel (:305-378) of
 * iteration itr, after the iteration's propagate and score update:
 * compaction of the flagged children in child order, clamped to the free
 * capacity (Q9); fAccept; every old node with frontier 0 re-activated when
 * its draw U <= score[region] + fAccept (repeat 1); accepted child j appended
 * as node tree_size + j (parent, cost = goal distance, frontier 1, repeat 15
 * while its region's valid count < 10, else 1); the smallest new index
 * within the goal threshold; the next expansion multiset (re-activated nodes
 * ascending, then each new node repeated, Q5), its first capacity entries in
 * M. Layouts: samples [capacity][D], child [n_children][D]. */
int orc_frontier_update(const kgmt_config* cfg, uint32_t itr, uint32_t tree_size, float* samples, int32_t* parents,
                        float* costs, uint8_t* frontier, const float* child, const int32_t* parentU,
                        const uint8_t* fnext, uint32_t nC, const int32_t* validc, const float* score,
                        const float* goal, uint32_t* M, uint32_t* nM_out, uint64_t* R_out, uint32_t* next_out,
                        int32_t* goal_out, uint32_t* nreact_out)
{
    int S, C;
    if(dims(cfg->model, &S, &C)) return KGMT_E_UNSUPPORTED;
    const int D          = S + C + 1;
    const uint32_t N     = cfg->capacity;
    const uint32_t query = cfg->query_id;
    if(tree_size > N) return KGMT_E_ARG;
    uint32_t* accept = (uint32_t*)malloc(sizeof(uint32_t) * (nC ? nC : 1));
    uint32_t* react  = (uint32_t*)malloc(sizeof(uint32_t) * (tree_size ? tree_size : 1));
    uint32_t* newrep = (uint32_t*)malloc(sizeof(uint32_t) * (nC ? nC : 1));
    if(!accept || !react || !newrep)
        {
            free(accept), free(react), free(newrep);
            return KGMT_E_NOMEM;
        }
    const uint32_t avail = N - tree_size;
    uint32_t nAcc        = 0;
    for(uint32_t u = 0; u < nC; ++u)
        if(fnext[u]) accept[nAcc++] = u;
    const uint32_t next = nAcc < avail ? nAcc : avail;
    const float fA      = orc_faccept(cfg, itr, tree_size, next);
    uint32_t nReact     = 0;
    for(uint32_t i = 0; i < tree_size; ++i)
        {
            if(frontier[i] == 0)
                {
                    float U = orc_draw_react(cfg, query, itr, i);
                    if(U <= score[orc_region(cfg, samples + (size_t)i * D)] + fA)
                        {
                            frontier[i]     = 1;
                            react[nReact++] = i;
                        }
                }
        }
    int32_t best    = -1;
    uint64_t sumRep = 0;
    for(uint32_t j = 0; j < next; ++j)
        {
            uint32_t t = tree_size + j, u = accept[j];
            const float* x1 = child + (size_t)u * D;
            memcpy(samples + (size_t)t * D, x1, sizeof(float) * D);
            parents[t]  = parentU[u];
            float dg    = pos_dist(cfg, x1, goal);
            costs[t]    = dg;
            frontier[t] = 1;
            int r1      = orc_region(cfg, x1);
            newrep[j]   = validc[r1] < cfg->sparse_valid_below ? cfg->repeat_sparse : cfg->repeat_dense;
            sumRep += newrep[j];
            if(dg < cfg->goal_thresh && best < 0) best = (int32_t)t;
        }
    uint32_t nM = 0;
    for(uint32_t k = 0; k < nReact && nM < N; ++k) M[nM++] = react[k];
    for(uint32_t j = 0; j < next && nM < N; ++j)
        for(uint32_t k = 0; k < newrep[j] && nM < N; ++k) M[nM++] = tree_size + j;
    *nM_out     = nM;
    *R_out      = (uint64_t)nReact + sumRep;
    *next_out   = next;
    *goal_out   = best;
    *nreact_out = nReact;
    free(accept), free(react), free(newrep);
    return KGMT_OK;
}

int orc_plan(const kgmt_config* cfg, const float* start, const float* goal, const float* obs,
             uint32_t n_obs, float* samples, int32_t* parents, float* costs, float* path, int32_t* total_out,
             int32_t* valid_out, float* score_out, uint32_t* iter_log, orc_stats* st)
{
    int S, C;
    if(dims(cfg->model, &S, &C)) return KGMT_E_UNSUPPORTED;
    const int D          = S + C + 1;
    const uint32_t N     = cfg->capacity, B = cfg->children_per_entry;
    const uint32_t query = cfg->query_id;
    int32_t R1, R2, P;
    orc_num_regions(cfg, &R1, &R2, &P);

    float* child      = (float*)malloc(sizeof(float) * (size_t)N * D);
    int32_t* parentU  = (int32_t*)malloc(sizeof(int32_t) * N);
    uint8_t* fnext    = (uint8_t*)malloc(N);
    uint8_t* frontier = (uint8_t*)calloc(N, 1);
    uint32_t* M       = (uint32_t*)malloc(sizeof(uint32_t) * N); /* expansion list, first N entries */
    int32_t* total    = (int32_t*)calloc((size_t)R1, sizeof(int32_t));
    int32_t* validc   = (int32_t*)calloc((size_t)R1, sizeof(int32_t));
    float* score      = (float*)calloc((size_t)R1, sizeof(float));
    uint8_t* actPrev  = (uint8_t*)calloc((size_t)R2, 1);
    uint8_t* actNext  = (uint8_t*)calloc((size_t)R2, 1);
    int rc            = KGMT_OK;
    if(!child || !parentU || !fnext || !frontier || !M || !total || !validc || !score || !actPrev || !actNext)
        {
            rc = KGMT_E_NOMEM;
            goto done;
        }

    /* init, KGMT.cu:50-78: root = start row, parent -1, frontier 1, repeat 5 */
    memcpy(samples, start, sizeof(float) * D);
    parents[0]   = -1;
    costs[0]     = 0.0f;
    frontier[0]  = 1;
    uint64_t R   = cfg->root_repeat;
    uint32_t mroot;
    for(mroot = 0; mroot < cfg->root_repeat && mroot < N; ++mroot) M[mroot] = 0;
    uint32_t treeSize = 1, itr = 0;
    int32_t goalIdx   = -1;
    uint64_t expanded = 0;

    while(itr < cfg->max_iter)
        {
            ++itr;
            /* -- propagateFrontier: children per entry (KGMT.cu:183-201) -- */
            uint32_t avail = N - treeSize;
            uint32_t per;
            if(R * B <= avail)
                per = B;
            else
                {
                    uint64_t q = R ? avail / R : 0;
                    per        = q > B ? B : (uint32_t)q;
                    if(per == 0) per = 1;
                }
            uint64_t nc64 = R * per;
            uint32_t nC   = nc64 > N ? N : (uint32_t)nc64;
            uint64_t Rnow = R;
            expanded += nC;
            memcpy(actNext, actPrev, (size_t)R2);
            for(uint32_t u = 0; u < nC; ++u)
                {
                    uint32_t p = M[u / per];
                    float ctrl[4], uf;
                    int32_t steps;
                    orc_draw_child(cfg, query, itr, u, ctrl, &steps, &uf);
                    float* x1 = child + (size_t)u * D;
                    int valid = orc_propagate(cfg, samples + (size_t)p * D, ctrl, steps, obs, n_obs, x1);
                    parentU[u] = (int32_t)p;
                    frontier[p] = 0;
                    int r1      = orc_region(cfg, x1);
                    int r2      = orc_subregion(cfg, x1, r1);
                    total[r1] += 1;
                    fnext[u] = 0;
                    if(valid)
                        {
                            validc[r1] += 1;
                            fnext[u]    = (uf < score[r1]) || !actPrev[r2];
                            actNext[r2] = 1;
                        }
                }
            /* -- Graph::updateVertices -- */
            orc_score(cfg, total, validc, actNext, score);
            memcpy(actPrev, actNext, (size_t)R2);

            /* -- updateFrontier: compaction, fAccept, re-activation, append -- */
            uint32_t next = 0, nReact = 0, nM32 = 0;
            int32_t best  = -1;
            rc = orc_frontier_update(cfg, itr, treeSize, samples, parents, costs, frontier, child, parentU, fnext, nC,
                                     validc, score, goal, M, &nM32, &R, &next, &best, &nReact);
            if(rc) goto done;
            (void)nM32;
            if(iter_log)
                {
                    iter_log[4 * (itr - 1) + 0] = (uint32_t)Rnow;
                    iter_log[4 * (itr - 1) + 1] = nC;
                    iter_log[4 * (itr - 1) + 2] = next;
                    iter_log[4 * (itr - 1) + 3] = nReact;
                }
            treeSize += next;
            if(best >= 0)
                {
                    goalIdx = best;
                    break;
                }
            /* a full tree can never take another node, so the goal is out of
             * reach: the plan ends unsuccessfully */
            if(treeSize == N) break;
        }

    /* path goal->root (KGMT.cu:342-361), exact depth (Q13) */
    st->path_rows   = 0;
    st->path_cost   = 0.0f;
    st->path_length = 0.0f;
    if(goalIdx >= 0)
        {
            int32_t n = goalIdx, rows = 0;
            const float* prev = NULL;
            while(n != -1)
                {
                    const float* row = samples + (size_t)n * D;
                    memcpy(path + (size_t)rows * D, row, sizeof(float) * D);
                    if(parents[n] != -1) st->path_cost += row[D - 1];
                    if(prev) st->path_length += pos_dist(cfg, prev, row);
                    prev = row;
                    ++rows;
                    n = parents[n];
                }
            st->path_rows = (uint32_t)rows;
        }
    st->iterations = itr;
    st->tree_size  = treeSize;
    st->goal_index = goalIdx;
    st->expanded   = expanded;
    if(total_out) memcpy(total_out, total, sizeof(int32_t) * (size_t)R1);
    if(valid_out) memcpy(valid_out, validc, sizeof(int32_t) * (size_t)R1);
    if(score_out) memcpy(score_out, score, sizeof(float) * (size_t)R1);

done:
    free(child), free(parentU), free(fnext), free(frontier), free(M), free(total), free(validc), free(score);
    free(actPrev), free(actNext);
    return rc;
}

int orc_plan_many(const kgmt_config* cfgs, uint32_t n_q, const float* starts, const float* goals, const float* obs,
                  const uint32_t* obs_off, const uint32_t* obs_cnt, int threads, orc_stats* st_out, double* ms_out)
{
    size_t maxSamples = 0, maxNodes = 0, maxPath = 0;
    for(uint32_t q = 0; q < n_q; ++q)
        {
            int S, C;
            if(dims(cfgs[q].model, &S, &C)) return KGMT_E_UNSUPPORTED;
            size_t D = (size_t)(S + C + 1);
            if((size_t)cfgs[q].capacity * D > maxSamples) maxSamples = (size_t)cfgs[q].capacity * D;
            if(cfgs[q].capacity > maxNodes) maxNodes = cfgs[q].capacity;
            if(((size_t)cfgs[q].max_iter + 2) * D > maxPath) maxPath = ((size_t)cfgs[q].max_iter + 2) * D;
        }
    int rc = KGMT_OK;
#ifdef _OPENMP
    if(threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
#else
    (void)threads;
#endif
    {
        float* samples   = (float*)malloc(sizeof(float) * maxSamples);
        int32_t* parents = (int32_t*)malloc(sizeof(int32_t) * maxNodes);
        float* costs     = (float*)malloc(sizeof(float) * maxNodes);
        float* path      = (float*)malloc(sizeof(float) * maxPath);
        if(!samples || !parents || !costs || !path) rc = KGMT_E_NOMEM;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for(int64_t q = 0; q < (int64_t)n_q; ++q)
            {
                if(rc != KGMT_OK) continue;
                struct timespec t0, t1;
                clock_gettime(CLOCK_MONOTONIC, &t0);
                int r = orc_plan(&cfgs[q], starts + (size_t)q * KGMT_MAX_SAMPLE,
                                 goals + (size_t)q * KGMT_MAX_SAMPLE, obs + 2 * orc_wdim(&cfgs[q]) * (size_t)obs_off[q], obs_cnt[q], samples,
                                 parents, costs, path, NULL, NULL, NULL, NULL, &st_out[q]);
                if(r != KGMT_OK) rc = r;
                clock_gettime(CLOCK_MONOTONIC, &t1);
                ms_out[q] = (double)(t1.tv_sec - t0.tv_sec) * 1e3 + (double)(t1.tv_nsec - t0.tv_nsec) * 1e-6;
            }
        free(samples), free(parents), free(costs), free(path);
    }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* batch conveniences over the functions above, in the SoA layouts of the    */
/* stage-level C-ABI (element (d, i) of an [X][n] array at d * n + i)        */
/* ------------------------------------------------------------------------ */
void orc_sincosf_n(const float* x, uint32_t n, float* s, float* c)
{
    for(uint32_t i = 0; i < n; ++i) orc_sincosf(x[i], &s[i], &c[i]);
}

void orc_philox_n(const uint32_t* ctr /* n x 4 */, const uint32_t* key /* n x 2 */, uint32_t n, uint32_t* out)
{
    for(uint32_t i = 0; i < n; ++i) orc_philox4x32_10(ctr + 4 * (size_t)i, key + 2 * (size_t)i, out + 4 * (size_t)i);
}

void orc_draw_children(const kgmt_config* cfg, uint32_t query, uint32_t itr, uint32_t u0, uint32_t n, float* controls,
                       int32_t* steps, float* ufront)
{
    int S, C;
    if(dims(cfg->model, &S, &C)) return;
    for(uint32_t i = 0; i < n; ++i)
        {
            float ctrl[4];
            orc_draw_child(cfg, query, itr, u0 + i, ctrl, &steps[i], &ufront[i]);
            for(int k = 0; k < C; ++k) controls[(size_t)k * n + i] = ctrl[k];
        }
}

int orc_propagate_children(const kgmt_config* cfg, const int32_t* parent_idx, const float* parent_states,
                           uint32_t n_parents, const float* controls, const int32_t* steps, uint32_t n,
                           const float* obs, uint32_t n_obs, float* child, uint8_t* valid, int32_t* r1, int32_t* r2)
{
    int S, C;
    if(dims(cfg->model, &S, &C)) return KGMT_E_UNSUPPORTED;
    const int D = S + C + 1;
    for(uint32_t i = 0; i < n; ++i)
        {
            float x0[12], u[4], x1[17];
            for(int d = 0; d < S; ++d) x0[d] = parent_states[(size_t)d * n_parents + parent_idx[i]];
            for(int k = 0; k < C; ++k) u[k] = controls[(size_t)k * n + i];
            valid[i] = (uint8_t)orc_propagate(cfg, x0, u, steps[i], obs, n_obs, x1);
            for(int d = 0; d < D; ++d) child[(size_t)d * n + i] = x1[d];
            r1[i] = orc_region(cfg, x1);
            r2[i] = orc_subregion(cfg, x1, r1[i]);
        }
    return KGMT_OK;
}

int orc_propagate_children_f64(const kgmt_config* cfg, const double* x0s /* n x S */, const double* ctrls /* n x C */,
                               const int32_t* steps, uint32_t n, const double* obs, uint32_t n_obs,
                               double* x1s /* n x D */, uint8_t* valid)
{
    int S, C;
    if(dims(cfg->model, &S, &C)) return KGMT_E_UNSUPPORTED;
    const int D = S + C + 1;
    for(uint32_t i = 0; i < n; ++i)
        valid[i] = (uint8_t)orc_propagate_f64(cfg, x0s + (size_t)i * S, ctrls + (size_t)i * C, steps[i], obs, n_obs,
                                              x1s + (size_t)i * D);
    return KGMT_OK;
}

void orc_segment_valid_n(const float* p0 /* n x 3 */, const float* p1 /* n x 3 */, uint32_t n, const float* obs,
                         uint32_t n_obs, uint8_t* out)
{
    for(uint32_t i = 0; i < n; ++i) out[i] = (uint8_t)orc_segment_valid(p0 + 3 * (size_t)i, p1 + 3 * (size_t)i, obs, n_obs);
}

void orc_region_n(const kgmt_config* cfg, const float* pos /* n x 3 */, uint32_t n, int32_t* r1, int32_t* r2)
{
    for(uint32_t i = 0; i < n; ++i)
        {
            r1[i] = orc_region(cfg, pos + 3 * (size_t)i);
            r2[i] = orc_subregion(cfg, pos + 3 * (size_t)i, r1[i]);
        }
}

/* ------------------------------------------------------------------------ */
/* Trajectory cost and adjoint (FP64). Not part of the reference planner:    */
/* the refinement extension differentiates the reference's quadcopter RK4    */
/* rollout (statePropagator.cu:246-370). VJP of ode() derived by hand:       */
/* position rates R(phi,theta,psi) V, Euler-angle rates, -w x V + R^T g +    */
/* drag + thrust, Euler's equations with drag and torques.                   */
/* ------------------------------------------------------------------------ */
void orc_quad_vjp_f64(const kgmt_config* cfg, const double* s, const double* c, const double* a, double* gs,
                      double* gc)
{
    double phi = s[3], th = s[4], psi = s[5];
    double u = s[6], v = s[7], w = s[8], p = s[9], q = s[10], r = s[11];
    double sf = sin(phi), cf = cos(phi), st = sin(th), ct = cos(th), sp = sin(psi), cp = cos(psi);
    double tt = st / ct;
    double g = cfg->gravity, mi = 1.0 / cfg->quad_mass;
    double ix = cfg->quad_ix, iy = cfg->quad_iy, iz = cfg->quad_iz;
    double ax = (iy - iz) / ix, ay = (iz - ix) / iy, az = (ix - iy) / iz;
    double bx = 1.0 / ix, by = 1.0 / iy, bz = 1.0 / iz;
    (void)c;
    for(int i = 0; i < 12; ++i) gs[i] = 0.0;
    /* rotation rows */
    double R00 = ct * cp, R01 = sf * st * cp - cf * sp, R02 = cf * st * cp + sf * sp;
    double R10 = ct * sp, R11 = sf * st * sp + cf * cp, R12 = cf * st * sp - sf * cp;
    double R20 = -st, R21 = sf * ct, R22 = cf * ct;
    double a0 = a[0], a1 = a[1], a2 = a[2];
    gs[6] += R00 * a0 + R10 * a1 + R20 * a2;
    gs[7] += R01 * a0 + R11 * a1 + R21 * a2;
    gs[8] += R02 * a0 + R12 * a1 + R22 * a2;
    /* d/dphi, d/dtheta, d/dpsi of R V */
    double d0f = (cf * st * cp + sf * sp) * v + (-sf * st * cp + cf * sp) * w;
    double d1f = (cf * st * sp - sf * cp) * v + (-sf * st * sp - cf * cp) * w;
    double d2f = cf * ct * v - sf * ct * w;
    double d0t = -st * cp * u + sf * ct * cp * v + cf * ct * cp * w;
    double d1t = -st * sp * u + sf * ct * sp * v + cf * ct * sp * w;
    double d2t = -ct * u - sf * st * v - cf * st * w;
    double f0 = R00 * u + R01 * v + R02 * w, f1 = R10 * u + R11 * v + R12 * w;
    gs[3] += d0f * a0 + d1f * a1 + d2f * a2;
    gs[4] += d0t * a0 + d1t * a1 + d2t * a2;
    gs[5] += -f1 * a0 + f0 * a1;
    /* Euler-angle rates */
    double E = q * sf + r * cf, F = q * cf - r * sf;
    double a3 = a[3], a4 = a[4], a5 = a[5];
    gs[3] += F * tt * a3 - E * a4 + F / ct * a5;
    gs[4] += E / (ct * ct) * a3 + E * st / (ct * ct) * a5;
    gs[9] += a3;
    gs[10] += sf * tt * a3 + cf * a4 + sf / ct * a5;
    gs[11] += cf * tt * a3 - sf * a4 + cf / ct * a5;
    /* translational dynamics */
    double a6 = a[6], a7 = a[7], a8 = a[8];
    double nv = sqrt(u * u + v * v + w * w), N = -cfg->quad_nu * nv;
    gs[7] += r * a6 - p * a8;
    gs[8] += -q * a6 + p * a7;
    gs[6] += -r * a7 + q * a8;
    gs[9] += w * a7 - v * a8;
    gs[10] += -w * a6 + u * a8;
    gs[11] += v * a6 - u * a7;
    gs[4] += -g * ct * a6 - g * st * sf * a7 - g * st * cf * a8;
    gs[3] += g * ct * cf * a7 - g * ct * sf * a8;
    double dotV = u * a6 + v * a7 + w * a8;
    gs[6] += mi * N * a6;
    gs[7] += mi * N * a7;
    gs[8] += mi * N * a8;
    if(nv > 0.0)
        {
            double k = mi * (-cfg->quad_nu) / nv * dotV;
            gs[6] += k * u, gs[7] += k * v, gs[8] += k * w;
        }
    /* rotational dynamics */
    double a9 = a[9], a10 = a[10], a11 = a[11];
    double nw = sqrt(p * p + q * q + r * r), M = -cfg->quad_mu * nw;
    gs[10] += ax * r * a9;
    gs[11] += ax * q * a9;
    gs[9] += ay * r * a10;
    gs[11] += ay * p * a10;
    gs[9] += az * q * a11;
    gs[10] += az * p * a11;
    double b9 = bx * a9, b10 = by * a10, b11 = bz * a11;
    gs[9] += M * b9, gs[10] += M * b10, gs[11] += M * b11;
    if(nw > 0.0)
        {
            double k = (-cfg->quad_mu) / nw * (p * b9 + q * b10 + r * b11);
            gs[9] += k * p, gs[10] += k * q, gs[11] += k * r;
        }
    gc[0] = mi * a8;
    gc[1] = bx * a9;
    gc[2] = by * a10;
    gc[3] = bz * a11;
}

static void rk4_f64(const kgmt_config* cfg, const double* x, const double* u, double* out)
{
    double h = (double)cfg->dt;
    double k1[12], k2[12], k3[12], k4[12], st[12];
    orc_quad_ode_f64(cfg, x, u, k1);
    for(int j = 0; j < 12; ++j) st[j] = x[j] + 0.5 * h * k1[j];
    orc_quad_ode_f64(cfg, st, u, k2);
    for(int j = 0; j < 12; ++j) st[j] = x[j] + 0.5 * h * k2[j];
    orc_quad_ode_f64(cfg, st, u, k3);
    for(int j = 0; j < 12; ++j) st[j] = x[j] + h * k3[j];
    orc_quad_ode_f64(cfg, st, u, k4);
    for(int j = 0; j < 12; ++j) out[j] = x[j] + h / 6.0 * (k1[j] + 2.0 * k2[j] + 2.0 * k3[j] + k4[j]);
}

/* x-bar += J_Phi^T lam, u-bar += J_Phi,u^T lam for one RK4 step from x */
static void rk4_vjp_f64(const kgmt_config* cfg, const double* x, const double* u, const double* lam, double* xbar,
                        double* ubar)
{
    double h = (double)cfg->dt;
    double k1[12], k2[12], k3[12], s2[12], s3[12], s4[12];
    orc_quad_ode_f64(cfg, x, u, k1);
    for(int j = 0; j < 12; ++j) s2[j] = x[j] + 0.5 * h * k1[j];
    orc_quad_ode_f64(cfg, s2, u, k2);
    for(int j = 0; j < 12; ++j) s3[j] = x[j] + 0.5 * h * k2[j];
    orc_quad_ode_f64(cfg, s3, u, k3);
    for(int j = 0; j < 12; ++j) s4[j] = x[j] + h * k3[j];
    double kb1[12], kb2[12], kb3[12], kb4[12], gs[12], gc[4];
    for(int j = 0; j < 12; ++j)
        {
            xbar[j] += lam[j];
            kb1[j] = h / 6.0 * lam[j], kb2[j] = h / 3.0 * lam[j], kb3[j] = h / 3.0 * lam[j], kb4[j] = h / 6.0 * lam[j];
        }
    orc_quad_vjp_f64(cfg, s4, u, kb4, gs, gc);
    for(int j = 0; j < 12; ++j) xbar[j] += gs[j], kb3[j] += h * gs[j];
    for(int i = 0; i < 4; ++i) ubar[i] += gc[i];
    orc_quad_vjp_f64(cfg, s3, u, kb3, gs, gc);
    for(int j = 0; j < 12; ++j) xbar[j] += gs[j], kb2[j] += 0.5 * h * gs[j];
    for(int i = 0; i < 4; ++i) ubar[i] += gc[i];
    orc_quad_vjp_f64(cfg, s2, u, kb2, gs, gc);
    for(int j = 0; j < 12; ++j) xbar[j] += gs[j], kb1[j] += 0.5 * h * gs[j];
    for(int i = 0; i < 4; ++i) ubar[i] += gc[i];
    orc_quad_vjp_f64(cfg, x, u, kb1, gs, gc);
    for(int j = 0; j < 12; ++j) xbar[j] += gs[j];
    for(int i = 0; i < 4; ++i) ubar[i] += gc[i];
}

/* tau * softplus((margin - sd(p, box)) / tau) summed over boxes, and its
 * gradient with respect to p (added into gp when not NULL) */
static double penalty_f64(const kgmt_adjoint_params* ap, const double* p, const double* obs, uint32_t n_obs, double scale,
                          double* gp)
{
    double total = 0.0;
    for(uint32_t o = 0; o < n_obs; ++o)
        {
            const double* b = obs + 6 * (size_t)o;
            double qv[3], sgn[3];
            for(int d = 0; d < 3; ++d)
                {
                    double lo = b[d] - p[d], hi = p[d] - b[3 + d];
                    if(lo > hi)
                        qv[d] = lo, sgn[d] = -1.0;
                    else
                        qv[d] = hi, sgn[d] = 1.0;
                }
            double ox = qv[0] > 0 ? qv[0] : 0, oy = qv[1] > 0 ? qv[1] : 0, oz = qv[2] > 0 ? qv[2] : 0;
            double outside = sqrt(ox * ox + oy * oy + oz * oz);
            int am         = qv[0] >= qv[1] ? (qv[0] >= qv[2] ? 0 : 2) : (qv[1] >= qv[2] ? 1 : 2);
            double inside  = qv[am] < 0 ? qv[am] : 0;
            double sd      = outside + inside;
            double z       = ((double)ap->margin - sd) / (double)ap->tau;
            double sp      = z > 30.0 ? z : log1p(exp(z));
            total += (double)ap->tau * sp;
            if(gp)
                {
                    double sig  = 1.0 / (1.0 + exp(-z));
                    double dsd  = -sig * scale; /* d(tau softplus)/d sd = -sigmoid(z) */
                    if(outside > 0)
                        {
                            gp[0] += dsd * (ox / outside) * sgn[0];
                            gp[1] += dsd * (oy / outside) * sgn[1];
                            gp[2] += dsd * (oz / outside) * sgn[2];
                        }
                    else
                        gp[am] += dsd * sgn[am];
                }
        }
    return total;
}

double orc_traj_cost_f64(const kgmt_config* cfg, const kgmt_adjoint_params* ap, int32_t T, const double* x0,
                         const double* u, const double* goal, const double* obs, uint32_t n_obs, double* xT)
{
    double x[12], y[12], J = 0.0;
    memcpy(x, x0, sizeof(x));
    for(int32_t t = 0; t < T; ++t)
        {
            const double* ut = u + 4 * (size_t)t;
            for(int i = 0; i < 4; ++i)
                {
                    double d = ut[i] - (double)ap->u_ref[i];
                    J += (double)ap->w_effort * d * d;
                }
            rk4_f64(cfg, x, ut, y);
            memcpy(x, y, sizeof(x));
            J += (double)ap->w_obstacle * penalty_f64(ap, x, obs, n_obs, 1.0, NULL);
        }
    for(int d = 0; d < 3; ++d)
        {
            double e = x[d] - goal[d];
            J += (double)ap->w_goal * e * e;
        }
    if(xT) memcpy(xT, x, sizeof(x));
    return J;
}

double orc_traj_grad_f64(const kgmt_config* cfg, const kgmt_adjoint_params* ap, int32_t T, const double* x0,
                         const double* u, const double* goal, const double* obs, uint32_t n_obs, double* grad)
{
    double* xs = (double*)malloc(sizeof(double) * 12 * ((size_t)T + 1));
    if(!xs) return NAN;
    memcpy(xs, x0, sizeof(double) * 12);
    for(int32_t t = 0; t < T; ++t) rk4_f64(cfg, xs + 12 * (size_t)t, u + 4 * (size_t)t, xs + 12 * ((size_t)t + 1));
    double J = orc_traj_cost_f64(cfg, ap, T, x0, u, goal, obs, n_obs, NULL);
    double lam[12] = {0};
    const double* xT = xs + 12 * (size_t)T;
    for(int d = 0; d < 3; ++d) lam[d] = 2.0 * (double)ap->w_goal * (xT[d] - goal[d]);
    if(T >= 1) penalty_f64(ap, xT, obs, n_obs, (double)ap->w_obstacle, lam);
    for(int32_t t = T - 1; t >= 0; --t)
        {
            const double* xt = xs + 12 * (size_t)t;
            const double* ut = u + 4 * (size_t)t;
            double xbar[12] = {0}, ubar[4] = {0};
            rk4_vjp_f64(cfg, xt, ut, lam, xbar, ubar);
            for(int i = 0; i < 4; ++i) grad[4 * (size_t)t + i] = ubar[i] + 2.0 * (double)ap->w_effort * (ut[i] - (double)ap->u_ref[i]);
            if(t >= 1) penalty_f64(ap, xt, obs, n_obs, (double)ap->w_obstacle, xbar);
            memcpy(lam, xbar, sizeof(lam));
        }
    free(xs);
    return J;
}
