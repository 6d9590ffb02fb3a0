// kernels.hip -- gfx950 kernels of the MD step.
//
// Step stages (one launch each, covering every system of a batch):
//   k_integrate      BAOAB: kick (B), drift (A), Ornstein-Uhlenbeck (O), drift
//                    (A); refreshes the bead's pair records and checks list age
//   list build       [k_open_grid] -> k_cell_assign -> cell scan -> k_place ->
//                    k_tile_items -> k_nl_tiles (every rebuild interval)
//   k_prune          inner list: the outer entries within cutoff + margin
//                    (every prune interval)
//   k_pair           list pair forces, a team of lanes per member row, with the
//                    bonded terms as leading work items of the same launch
//   k_half_kick      closing half kick at the end of a chunk
//   k_pack           pair-record refresh outside a step (after set_positions,
//                    a restore, the minimizer's moves)
//   k_snapshot / k_status_out   chunk start snapshot and end-of-chunk flags
// Forces are computed in FP32 from FP64 displacements and accumulated as
// 64-bit fixed point (2^-32 kJ/mol/nm); positions/velocities are FP64.
// The energy pass (k_pair / k_bonded with ENERGY=true) evaluates terms in FP64.
#include "device_common.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

namespace ocm {

// ------------------------------------------------------------ small helpers
// Out-of-range or non-finite contributions are flagged and replaced by 0, so
// the float-to-integer conversion is only ever applied to representable values.
//
// FP32 values go to fixed point as floor(v * 2^32): floor(v) is the high word
// and the fraction, scaled by 2^32, the low word.  v - floor(v) is exact except
// for negative |v| < 1, where it is rounded once (by at most 2^-25); the low
// word is clamped below 2^32 for the case it rounds up to 1.  Requires |v| < 2^30.
__device__ inline long long fix_floor(float v) {
    const float h = floorf(v);
    const float l = fminf((v - h) * 4294967296.0f, 4294967040.0f);
    return (static_cast<long long>(static_cast<int>(h)) << 32) + static_cast<long long>(static_cast<unsigned>(l));
}
__device__ inline long long to_fix(float v, int& bad) {
    const bool ok = fabsf(v) < kFixLimitF;
    bad |= !ok;
    return fix_floor(ok ? v : 0.0f);
}
__device__ inline long long to_fix_d(double v, int& bad) {
    const bool ok = fabs(v) < 1073741824.0;
    bad |= !ok;
    return ok ? static_cast<long long>(v * kFixScale) : 0LL;
}

// Minimum-image component: d - L * rint(d / L), written with one fused
// multiply-add so the CPU oracle reproduces the same bits.
__device__ inline double min_image(double d, double L, double invL) {
    return __fma_rn(-L, rint(d * invL), d);
}
__device__ inline double dist2(double dx, double dy, double dz) {
    return __fma_rn(dz, dz, __fma_rn(dy, dy, dx * dx));
}

template <typename T>
__device__ inline T warp_sum(T v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, width);
    return v;
}

// XCD-sliced walk of the pair kernel's block table (batches of many
// systems).  Workgroups are dispatched to the device's 8 XCDs round-robin,
// and each XCD has its own L2; workgroup b takes entries of slice b % 8, one
// contiguous eighth of the table (the host gives every slice the same mix of
// work), so table-adjacent work stays on one XCD.  Every entry is taken
// exactly once (a grid smaller than 8 walks the whole table with a plain
// stride).
constexpr int kXcds = 8;
struct XcdWalk {
    int first, end, stride;
};
__device__ __forceinline__ XcdWalk xcd_walk(int n) {
    const int g = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    if (g < kXcds) return {b, n, g};
    const int x = b % kXcds, j = b / kXcds;
    const int nbx = (g - x + kXcds - 1) / kXcds;   // workgroups on XCD x
    const int lo = static_cast<int>(static_cast<long long>(x) * n / kXcds);
    const int hi = static_cast<int>(static_cast<long long>(x + 1) * n / kXcds);
    return {lo + j, hi, nbx};
}

// ------------------------------------------------------------------ fills
__global__ void k_fill_i32(int* p, int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i64(long long* p, long long v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------- chunk bracket
// First kernel of a chunk: copies the chunk-start state (positions,
// velocities and force words, 3*nb each, then the two step slots) into the
// snapshot buffer a rollback restores, and clears the batch flag word.
__global__ void k_snapshot(int n3, const double* __restrict__ pos, const double* __restrict__ vel,
                           const long long* __restrict__ frc, const long long* __restrict__ gstep,
                           double* __restrict__ snap, int* __restrict__ flags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        reinterpret_cast<long long*>(snap)[3 * static_cast<size_t>(n3)] = gstep[0];
        reinterpret_cast<long long*>(snap)[3 * static_cast<size_t>(n3) + 1] = gstep[1];
        *flags = 0;
    }
    if (t >= n3) return;
    snap[t] = pos[t];
    snap[n3 + t] = vel[t];
    reinterpret_cast<long long*>(snap)[2 * static_cast<size_t>(n3) + t] = frc[t];
}

// Last kernel of a chunk: copies the flag word and the per-system error
// words into pinned host memory (host[0] flags, host[1 + s] system s) and
// clears the error words for the next chunk.
__global__ void k_status_out(int nsys, const int* __restrict__ flags, int* __restrict__ err,
                             int* __restrict__ host) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) host[0] = *flags;
    for (int s = t; s < nsys; s += gridDim.x * blockDim.x) {
        host[1 + s] = err[s];
        err[s] = 0;
    }
}

// ---------------------------------------------------- record refresh
// The integrator refreshes the coordinates of every pair record of bead b
// (type and bead id were written by the build) and, when requested, flags
// the system if the bead moved more than skin/2 since the last list build
// (the list is then rebuilt by the host's rollback).  Reference positions and
// the (skin/2)^2 bound are per bead, so every input is loaded at once.
struct PackArgs {
    int enabled;
    int check;
    int maxm;              // member slots per bead in bead_slot (unused ones hold -1)
    const int* bead_slot;
    const double* bead_hs2;   // per bead: smallest (skin/2)^2 of its groups
    const double* refx;       // per bead: position at the last build
    const double* refy;
    const double* refz;
    PairRec* rec;
    float4* rec32;         // written on prune steps for periodic systems
    double* prx;           // positions at the last inner-list prune, by bead
    double* pry;
    double* prz;
    double prune_hm2;
    int prune;             // kPruneNone / kPruneWrite / kPruneCheck
};

// Inputs of one bead's record refresh and checks, loaded before its step is
// computed so the loads overlap the arithmetic.
struct PackIn {
    int slot0, slot1;
    double rx, ry, rz, hs2;
    double qx, qy, qz;
};

__device__ __forceinline__ PackIn pack_load(const PackArgs& pk, int b) {
    PackIn in{-1, -1, 0, 0, 0, 0, 0, 0, 0};
    if (!pk.enabled && pk.prune == kPruneNone) return in;
    in.slot0 = pk.bead_slot[static_cast<size_t>(b) * pk.maxm];
    in.slot1 = pk.maxm > 1 ? pk.bead_slot[static_cast<size_t>(b) * pk.maxm + 1] : -1;
    if (pk.enabled && pk.check) {
        in.rx = pk.refx[b]; in.ry = pk.refy[b]; in.rz = pk.refz[b];
        in.hs2 = pk.bead_hs2[b];
    }
    if (pk.prune == kPruneCheck) {
        in.qx = pk.prx[b]; in.qy = pk.pry[b]; in.qz = pk.prz[b];
    }
    return in;
}

__device__ __forceinline__ void put_xyz(PairRec* rec, int slot, double x, double y, double z) {
    double* p = reinterpret_cast<double*>(rec + slot);
    p[0] = x; p[1] = y; p[2] = z;
}

// Position wrapped into [0, L) on every axis, rounded to FP32: the prune's
// copy for periodic groups (one 16-byte gather instead of 24 bytes in two).
__device__ __forceinline__ float4 wrap32(double x, double y, double z, const double* box, const double* ib) {
    return make_float4(static_cast<float>(x - box[0] * floor(x * ib[0])),
                       static_cast<float>(y - box[1] * floor(y * ib[1])),
                       static_cast<float>(z - box[2] * floor(z * ib[2])), 0.f);
}

__device__ inline void pack_bead(const PackArgs& pk, const PackIn& in, int b, int sys, const DevSys& S, double x,
                                 double y, double z, int* err) {
    // the prune reads FP32 copies of periodic systems' positions
    const bool w32 = pk.prune == kPruneWrite && S.periodic && in.slot0 >= 0;
    const float4 q = w32 ? wrap32(x, y, z, S.box, S.inv_box) : make_float4(0.f, 0.f, 0.f, 0.f);
    if (w32) {
        pk.rec32[in.slot0] = q;
        if (in.slot1 >= 0) pk.rec32[in.slot1] = q;
        for (int w = 2; in.slot1 >= 0 && w < pk.maxm; ++w) {
            const int slot = pk.bead_slot[static_cast<size_t>(b) * pk.maxm + w];
            if (slot < 0) break;
            pk.rec32[slot] = q;
        }
    }
    if (pk.enabled && in.slot0 >= 0) {
        put_xyz(pk.rec, in.slot0, x, y, z);
        if (in.slot1 >= 0) {
            put_xyz(pk.rec, in.slot1, x, y, z);
            for (int w = 2; w < pk.maxm; ++w) {
                const int slot = pk.bead_slot[static_cast<size_t>(b) * pk.maxm + w];
                if (slot < 0) break;
                put_xyz(pk.rec, slot, x, y, z);
            }
        }
        if (pk.check) {
            const double dx = x - in.rx, dy = y - in.ry, dz = z - in.rz;
            if (!(dist2(dx, dy, dz) <= in.hs2)) atomicOr(&err[sys], kErrStale);
        }
    }
}

// ------------------------------------------------------------- integrator
// Three N(0,1) samples for the step: the Philox words of gauss3, Box-Muller on
// the top 24 bits of each word with the hardware FP32 log2, sqrt and sin/cos
// (v_sin_f32 / v_cos_f32 take the angle in revolutions).  The tails end at
// 5.8 sigma (u >= 2^-25) instead of 6.8; samples agree with gauss3 to ~1e-6.
__device__ inline void gauss3_fast(uint64_t seed, uint64_t step, uint32_t bead, double g[3]) {
    uint32_t c[4] = {static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32), bead, 0u};
    philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    const float s = 5.9604644775390625e-8f;  // 2^-24
    const float u0 = (static_cast<float>(c[0] >> 8) + 0.5f) * s, u1 = (static_cast<float>(c[1] >> 8) + 0.5f) * s;
    const float u2 = (static_cast<float>(c[2] >> 8) + 0.5f) * s, u3 = (static_cast<float>(c[3] >> 8) + 0.5f) * s;
    const float m2ln2 = -1.3862943611198906f;   // -2 ln 2: -2 ln u = -2 ln2 log2 u
    const float r0 = sqrtf(m2ln2 * __builtin_amdgcn_logf(u0));
    const float r1 = sqrtf(m2ln2 * __builtin_amdgcn_logf(u2));
    g[0] = static_cast<double>(r0 * __builtin_amdgcn_cosf(u1));
    g[1] = static_cast<double>(r0 * __builtin_amdgcn_sinf(u1));
    g[2] = static_cast<double>(r1 * __builtin_amdgcn_cosf(u3));
}

// One half kick v + (dt/2) F/m, as one FMA.  The closing kick of a step and the
// opening kick of the next are applied as two such roundings whether they run
// in one launch or in two, so the state after any step is the same bits
// however the steps are grouped into calls.
__device__ __forceinline__ double half_kick(double v, double h, double f, double im) {
    return fma(h * f, im, v);
}

// One bead's BAOAB step from its force words F (fixed point), position X and
// velocity V: the closing half kick of the previous step first when pending,
// the opening half kick, half a drift, the Ornstein-Uhlenbeck update with the
// step's noise, half a drift.  Fixed (massless) beads keep X and V.  Returns
// false when a mobile bead's new position is not finite.
// g holds the step's three N(0,1) samples (gauss3_fast).
__device__ __forceinline__ bool baoab_bead_g(const DevSys& S, double im, long long F0, long long F1, long long F2,
                                             const double (&X)[3], const double (&V)[3], int pending,
                                             const double (&g)[3], double (&x)[3], double (&v)[3]) {
    const double f0 = F0 * kFixInv, f1 = F1 * kFixInv, f2 = F2 * kFixInv;
    double v0 = V[0], v1 = V[1], v2 = V[2];
    if (pending) {
        v0 = half_kick(v0, S.half_dt, f0, im);
        v1 = half_kick(v1, S.half_dt, f1, im);
        v2 = half_kick(v2, S.half_dt, f2, im);
    }
    v0 = half_kick(v0, S.half_dt, f0, im);
    v1 = half_kick(v1, S.half_dt, f1, im);
    v2 = half_kick(v2, S.half_dt, f2, im);
    // every product-sum is an explicit fma, so the bits do not depend on
    // how the compiler contracts the expression in the kernel it is inlined in
    double x0 = fma(S.half_dt, v0, X[0]);
    double x1 = fma(S.half_dt, v1, X[1]);
    double x2 = fma(S.half_dt, v2, X[2]);
    const double sig = sqrt(S.noise_kT * im);
    v0 = fma(S.c1, v0, sig * g[0]);
    v1 = fma(S.c1, v1, sig * g[1]);
    v2 = fma(S.c1, v2, sig * g[2]);
    x0 = fma(S.half_dt, v0, x0);
    x1 = fma(S.half_dt, v1, x1);
    x2 = fma(S.half_dt, v2, x2);
    if (im == 0.0) {
        x0 = X[0]; x1 = X[1]; x2 = X[2];
        v0 = V[0]; v1 = V[1]; v2 = V[2];
    }
    x[0] = x0; x[1] = x1; x[2] = x2;
    v[0] = v0; v[1] = v1; v[2] = v2;
    return im == 0.0 || (isfinite(x0) && isfinite(x1) && isfinite(x2));
}

__device__ __forceinline__ bool baoab_bead(const DevSys& S, double im, long long F0, long long F1, long long F2,
                                           const double (&X)[3], const double (&V)[3], int pending, long long step,
                                           uint32_t lid, double (&x)[3], double (&v)[3]) {
    double g[3];
    gauss3_fast(S.seed, static_cast<uint64_t>(step - S.step_origin), lid, g);
    return baoab_bead_g(S, im, F0, F1, F2, X, V, pending, g, x, v);
}

// pending != 0: the closing half kick of the previous step is still owed and
// is applied first.  gstep[pin] is this step's global step; thread 0 writes
// gstep[pout] = gstep[pin] + 1 for the next step.
__global__ void k_integrate(int nb, const DevSys* __restrict__ sys, const int* __restrict__ sys_of,
                            const int* __restrict__ local_id, const double* __restrict__ inv_m,
                            double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
                            double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                            long long* __restrict__ fx, long long* __restrict__ fy,
                            long long* __restrict__ fz, long long* gstep, int pin, int pout,
                            int pending, int* __restrict__ err, PackArgs pk, DevSys sys0, int single) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const long long step = gstep[pin];
    if (b == 0) gstep[pout] = step + 1;
    if (b >= nb) return;
    // a batch of one passes its system record by value: one dependent load less
    const int s = single ? 0 : sys_of[b];
    const DevSys S = single ? sys0 : sys[s];
    // Every input is loaded before anything is stored or branched on, so the
    // step is one memory round trip: the bead's state, its force words, and
    // the record-refresh inputs.  Fixed (massless) beads run the same
    // arithmetic and keep their position and velocity by selection.
    const double im = inv_m[b];
    const PackIn pin_ = pack_load(pk, b);
    const long long F0 = fx[b], F1 = fy[b], F2 = fz[b];
    const double X0 = px[b], X1 = py[b], X2 = pz[b];
    const double V0 = vx[b], V1 = vy[b], V2 = vz[b];
    const uint32_t lid = static_cast<uint32_t>(local_id[b]);
    // an empty statement that needs every loaded value: the loads are all
    // issued before it and waited for once (otherwise the compiler sinks the
    // mobile-bead inputs behind a wait on the mass and a branch)
    asm volatile("" ::"v"(im), "v"(F0), "v"(F1), "v"(F2), "v"(X0), "v"(X1), "v"(X2), "v"(V0), "v"(V1), "v"(V2),
                 "v"(lid));
    double xn[3], vn[3];
    const bool finite = baoab_bead(S, im, F0, F1, F2, {X0, X1, X2}, {V0, V1, V2}, pending, step, lid, xn, vn);
    if (!finite) atomicOr(&err[s], kErrNumeric);
    const bool fixed = im == 0.0;
    const double x0 = xn[0], x1 = xn[1], x2 = xn[2], v0 = vn[0], v1 = vn[1], v2 = vn[2];
    fx[b] = 0; fy[b] = 0; fz[b] = 0;
    px[b] = x0; py[b] = x1; pz[b] = x2;
    vx[b] = v0; vy[b] = v1; vz[b] = v2;
    // inner list: a pair dropped at the prune (r >= cutoff + margin) cannot be
    // inside the cutoff while both beads stay within margin/2 of where they were
    if (pk.prune == kPruneWrite) {
        pk.prx[b] = x0; pk.pry[b] = x1; pk.prz[b] = x2;
    } else if (pk.prune == kPruneCheck && pin_.slot0 >= 0 && !fixed) {
        const double dx = x0 - pin_.qx, dy = x1 - pin_.qy, dz = x2 - pin_.qz;
        if (!(dist2(dx, dy, dz) <= pk.prune_hm2)) atomicOr(&err[s], kErrPrune);
    }
    pack_bead(pk, pin_, b, s, S, x0, x1, x2, err);
}

// Closing half kick; also resets both global-step slots to the value reached.
__global__ void k_half_kick(int nb, const DevSys* __restrict__ sys, const int* __restrict__ sys_of,
                            const double* __restrict__ inv_m, double* __restrict__ vx,
                            double* __restrict__ vy, double* __restrict__ vz,
                            const long long* __restrict__ fx, const long long* __restrict__ fy,
                            const long long* __restrict__ fz, long long* gstep, int pcur) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) {
        const long long s = gstep[pcur];
        gstep[0] = s;
        gstep[1] = s;
    }
    if (b >= nb) return;
    const double im = inv_m[b];
    if (im == 0.0) return;
    const double h = sys[sys_of[b]].half_dt;
    vx[b] = half_kick(vx[b], h, fx[b] * kFixInv, im);
    vy[b] = half_kick(vy[b], h, fy[b] * kFixInv, im);
    vz[b] = half_kick(vz[b], h, fz[b] * kFixInv, im);
}

// ------------------------------------------------------------ record pack
__global__ void k_pack(int nm, const DevPairGroup* __restrict__ groups,
                       const int* __restrict__ member_group, const int* __restrict__ member_bead,
                       const int* __restrict__ member_type, const int* __restrict__ member_slot,
                       const double* __restrict__ px, const double* __restrict__ py,
                       const double* __restrict__ pz, const double* __restrict__ refx,
                       const double* __restrict__ refy, const double* __restrict__ refz,
                       PairRec* __restrict__ rec, float4* __restrict__ rec32, int* __restrict__ err) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const int b = member_bead[m];
    const int slot = member_slot[m];
    PairRec r;
    r.x = px[b]; r.y = py[b]; r.z = pz[b];
    r.type = member_type[m];
    r.bead = b;
    rec[slot] = r;
    const DevPairGroup& G = groups[member_group[m]];
    if (rec32 && G.periodic) rec32[slot] = wrap32(r.x, r.y, r.z, G.box, G.inv_box);
    const double dx = r.x - refx[b], dy = r.y - refy[b], dz = r.z - refz[b];
    if (!(dist2(dx, dy, dz) <= G.half_skin2)) atomicOr(&err[G.sys], kErrStale);
}

// ------------------------------------------------------------ list build
// Open-boundary groups with a cell grid: one workgroup per group finds the
// members' bounding box and lays the grid over it, cells of rl = cutoff +
// skin (widened while the grid would exceed the cells reserved for it).
// Beads that leave the box before the next build are clamped into the edge
// cells, which keeps every pair within rl in neighbouring cells.
constexpr int kOpenGridThreads = 256;
__global__ __launch_bounds__(kOpenGridThreads) void k_open_grid(DevPairGroup* __restrict__ groups,
                                                                const int* __restrict__ member_bead,
                                                                const double* __restrict__ px,
                                                                const double* __restrict__ py,
                                                                const double* __restrict__ pz) {
    DevPairGroup& G = groups[blockIdx.x];
    if (!G.open_grid) return;   // uniform over the workgroup
    __shared__ double red[6][kOpenGridThreads];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int m = G.member_begin + threadIdx.x; m < G.member_begin + G.member_count; m += blockDim.x) {
        const int b = member_bead[m];
        const double p[3] = {px[b], py[b], pz[b]};
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (isfinite(p[d])) { lo[d] = fmin(lo[d], p[d]); hi[d] = fmax(hi[d], p[d]); }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) { red[d][threadIdx.x] = lo[d]; red[3 + d][threadIdx.x] = hi[d]; }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + w]);
                red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double ext[3], l0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const bool any = red[d][0] <= red[3 + d][0];
        l0[d] = any ? red[d][0] : 0.0;
        ext[d] = any ? red[3 + d][0] - red[d][0] : 0.0;
    }
    double h = sqrt(G.rl2);
    int n[3] = {1, 1, 1};
    bool fits = false;
    for (int it = 0; it < 512 && !fits; ++it) {
        long long total = 1;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double c = floor(ext[d] / h) + 1.0;
            n[d] = c < 1.0 ? 1 : (c > 1.0e6 ? 1000000 : static_cast<int>(c));
            total *= n[d];
        }
        fits = total <= G.grid_cap;
        if (!fits) h *= 1.25;
    }
    if (!fits) {   // an extent no finite widening covers: one cell
        n[0] = n[1] = n[2] = 1;
        h = 1.0;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        G.ncell[d] = n[d];
        G.reach[d] = n[d] > 1 ? 1 : 0;
        G.cell_scale[d] = 1.0 / h;
        G.cell_lo[d] = l0[d];
    }
}

__global__ void k_cell_assign(int nm, const DevPairGroup* __restrict__ groups,
                              const int* __restrict__ member_group, const int* __restrict__ member_bead,
                              const double* __restrict__ px, const double* __restrict__ py,
                              const double* __restrict__ pz, int* __restrict__ member_cell,
                              int* __restrict__ member_rank, int* __restrict__ cell_count,
                              int* __restrict__ n_items) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m == 0) *n_items = 0;   // the tile work list of this build starts empty
    if (m >= nm) return;
    const DevPairGroup& G = groups[member_group[m]];
    const int b = member_bead[m];
    int c[3] = {0, 0, 0};
    if (G.periodic) {
        const double p[3] = {px[b], py[b], pz[b]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double u = p[d] * G.inv_box[d];
            u -= floor(u);
            int ci = isfinite(u) ? static_cast<int>(u * G.ncell[d]) : 0;
            c[d] = ci < 0 ? 0 : (ci >= G.ncell[d] ? G.ncell[d] - 1 : ci);
        }
    } else if (G.open_grid) {
        const double p[3] = {px[b], py[b], pz[b]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            // clamped into the grid (in FP64, before the conversion)
            const double u = floor((p[d] - G.cell_lo[d]) * G.cell_scale[d]);
            c[d] = isfinite(u) ? static_cast<int>(fmin(fmax(u, 0.0), static_cast<double>(G.ncell[d] - 1))) : 0;
        }
    }
    const int cell = G.cell_begin + (c[2] * G.ncell[1] + c[1]) * G.ncell[0] + c[0];
    member_cell[m] = cell;
    // consecutive members (beads along a chain) mostly share a cell: each run
    // of equal cells in the wavefront takes its ranks with one atomic, and all
    // runs issue theirs together.  Active lanes are a prefix of the wave (only
    // m >= nm leave), so lane - 1 is active whenever lane is.
    const int lane = threadIdx.x & 63;
    const unsigned long long active = __builtin_amdgcn_ballot_w64(true);
    const int prev = __shfl_up(cell, 1, 64);
    const bool lead = lane == 0 || prev != cell;
    const unsigned long long leaders = __builtin_amdgcn_ballot_w64(lead);
    // this lane's run starts at the highest leader at or below it and ends
    // before the next leader (or the last active lane)
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int first = 63 - __clzll(static_cast<long long>(leaders & upto));
    const unsigned long long after = leaders & ~upto;
    const int end = after ? __ffsll(static_cast<long long>(after)) - 1 : 64 - __clzll(static_cast<long long>(active));
    int base = 0;
    if (lead) base = atomicAdd(&cell_count[cell], end - lane);
    base = __shfl(base, first, 64);
    member_rank[m] = base + lane - first;
}

__global__ void k_place(int nm, const DevPairGroup* __restrict__ groups, const int* __restrict__ member_group,
                        const int* __restrict__ member_bead, const int* __restrict__ member_type,
                        const int* __restrict__ member_cell, const int* __restrict__ member_rank,
                        const int* __restrict__ cell_start, const double* __restrict__ px,
                        const double* __restrict__ py, const double* __restrict__ pz,
                        int* __restrict__ member_slot, int* __restrict__ slot_member,
                        const int* __restrict__ member_kpos, int* __restrict__ bead_slot,
                        PairRec* __restrict__ rec, float4* __restrict__ rec32, double* __restrict__ refx,
                        double* __restrict__ refy, double* __restrict__ refz,
                        int* __restrict__ cell_count, int ncell) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    // the counts were consumed by the scan; leave them zero for the next build
    for (int c = m; c < ncell; c += gridDim.x * blockDim.x) cell_count[c] = 0;
    if (m >= nm) return;
    const int slot = cell_start[member_cell[m]] + member_rank[m];
    member_slot[m] = slot;
    slot_member[slot] = m;
    bead_slot[member_kpos[m]] = slot;
    const int b = member_bead[m];
    PairRec r;
    r.x = px[b]; r.y = py[b]; r.z = pz[b];
    r.type = member_type[m];
    r.bead = b;
    rec[slot] = r;
    refx[b] = r.x; refy[b] = r.y; refz[b] = r.z;   // every member of b writes the same values
    if (rec32) {
        const DevPairGroup& G = groups[member_group[m]];
        if (G.periodic) rec32[slot] = wrap32(r.x, r.y, r.z, G.box, G.inv_box);
    }
}

__device__ inline bool excluded(const int* __restrict__ excl, int beg, int cnt, int mj) {
    // exclusion rows hold partner bead ids, sorted
    int lo = beg, hi = beg + cnt;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = excl[mid];
        if (v == mj) return true;
        if (v < mj) lo = mid + 1; else hi = mid;
    }
    return false;
}

// LDS-staged cell tiles.  One workgroup per home cell: the records of the
// stencil cells are staged into LDS in windows of kTileCap, and every row of
// the home cell (one wavefront per row) sweeps the window with its 64 lanes.
// A candidate passes when r^2 < rl^2 at the build positions, its category is
// allowed and it is not excluded; passes are counted with a ballot and, in
// the FILL pass, written in window order at base + popcount of lower lanes.
constexpr int kTileCap = 1024;
constexpr int kMaxStencil = 27;

constexpr int kTileChunkMax = 64;  // rows per work item: 16, or 64 for large lists (staging amortised)

// Work items of the tile build: one per (non-empty cell, 16-row chunk),
// appended by k_tile_items after the cell scan; the tile kernels run on a
// fixed grid and loop over them.
__global__ void k_tile_items(int ncell, const int* __restrict__ cell_start, int2* __restrict__ items,
                             int* __restrict__ n_items, int chunk) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int n = (cell_start[c + 1] - cell_start[c] + chunk - 1) / chunk;
    if (n == 0) return;
    const int at = atomicAdd(n_items, n);
    for (int q = 0; q < n; ++q) items[at + q] = make_int2(c, q);
}

// One pass: every row owns G.stride slots of col starting at row_start (fixed
// per row), so rows are written as they are found.  A row longer than its
// slots writes only what fits, reports its length through group_need and the
// overflow flag, and gets row_count 0; the host then widens the rows of that
// group and builds again.
//
// The list only has to hold every pair within rc + skin (the pair kernel
// applies the exact FP64 cutoff, and the neighbor-pair dump its own FP64
// test), so candidates are tested in FP32 against (rc + skin)^2 (1 + 1e-3):
// a superset.  Candidates are staged as FP32 positions relative to the item's
// first row, already shifted to the image the stencil cell stands for; on a
// periodic axis with a single cell the image is chosen per pair instead.
constexpr float kListMargin = 1.001f;

__device__ __forceinline__ double wrap_coord(double x, double L, double invL) { return x - L * floor(x * invL); }

constexpr int kRowBlock = 4;   // rows a wave tests against each staged candidate
constexpr int kExclReg = 2;    // exclusions per row tested during the sweep; rows with more are filtered after

// One wave's sweep of a staged window for kRowBlock rows whose values are
// wave-uniform (scalar registers).  Window entries past the staged count are
// padding with no category bit, so every lane tests its candidate without a
// bounds branch.  A passing candidate is stored through the row's buffer
// resource at run + (passing lanes below it); the resource's range is the
// row's capacity, so a position past it is dropped by the hardware, and lanes
// that do not pass store to an offset past every range.
struct NlRow {
    float x, y, z;
    uint32_t amask;        // categories this row's category pairs with
    int self;              // the row's own slot
    int ex[kExclReg];      // excluded partner beads (-2: none; a row with more is filtered after)
    __amdgpu_buffer_rsrc_t out;
};

// CAT: the group has more than one category (test the allow mask; with one,
// only padding, whose category bit is 0, fails).
template <bool MI, bool CAT>
__device__ __forceinline__ void nl_sweep(const float4* __restrict__ cpos, const int2* __restrict__ cinfo, int nw,
                                         int lane, NlRow (&rw)[kRowBlock], int (&run)[kRowBlock],
                                         float rl2f, float Lf0, float Lf1, float Lf2, float iLf0, float iLf1,
                                         float iLf2, bool mi0, bool mi1, bool mi2) {
    for (int kb = 0; kb < nw; kb += 64) {
        const int k = kb + lane;
        const float4 c = cpos[k];
        const int2 inf = cinfo[k];   // {category bit, bead}
        const int sl = __float_as_int(c.w);
#pragma unroll
        for (int q = 0; q < kRowBlock; ++q) {
            float dx = c.x - rw[q].x, dy = c.y - rw[q].y, dz = c.z - rw[q].z;
            if (MI) {
                if (mi0) dx -= Lf0 * rintf(dx * iLf0);
                if (mi1) dy -= Lf1 * rintf(dy * iLf1);
                if (mi2) dz -= Lf2 * rintf(dz * iLf2);
            }
            const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            // the wave mask is the AND of the tests' own masks (each ballot is
            // its compare's result), so no per-lane flag is rebuilt for it
            const bool near = d2 < rl2f, other = sl != rw[q].self;
            const bool cat = CAT ? (rw[q].amask & static_cast<uint32_t>(inf.x)) != 0u : inf.x != 0;
            const bool ex0 = inf.y != rw[q].ex[0], ex1 = inf.y != rw[q].ex[1];
            static_assert(kExclReg == 2, "the sweep tests two exclusions");
            const unsigned long long mask = __builtin_amdgcn_ballot_w64(near) & __builtin_amdgcn_ballot_w64(other) &
                                            __builtin_amdgcn_ballot_w64(cat) & __builtin_amdgcn_ballot_w64(ex0) &
                                            __builtin_amdgcn_ballot_w64(ex1);
            const bool pass = near & other & cat & ex0 & ex1;
            const int below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
            const int off = pass ? (run[q] + below) * 4 : static_cast<int>(0x80000000u);
            __builtin_amdgcn_raw_buffer_store_b32(sl, rw[q].out, off, 0, 0);
            run[q] += __popcll(mask);
        }
    }
}

__global__ __launch_bounds__(256) void k_nl_tiles(
    const int2* __restrict__ items, const int* __restrict__ n_items, const DevPairGroup* __restrict__ groups,
    const int* __restrict__ cell_group,
    const int* __restrict__ cell_start, const PairRec* __restrict__ rec,
    const int* __restrict__ slot_member, const int* __restrict__ excl_start, const int* __restrict__ excl_count,
    const int* __restrict__ excl, const int* __restrict__ row_start, int* __restrict__ row_count,
    int* __restrict__ group_need, int* __restrict__ col, int* __restrict__ flags, int chunk) {
    __shared__ float4 cpos[kTileCap];      // x, y, z relative to the item's first row; w = slot (bits)
    __shared__ int2 cinfo[kTileCap];       // {1 << category, bead}; padding {0, -1}
    __shared__ int row_run[kTileChunkMax];
    __shared__ int st_start[kMaxStencil], st_pre[kMaxStencil + 1];
    __shared__ float st_shift[kMaxStencil][3];
    const int tid = threadIdx.x, lane = tid & 63, nwave = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nit = *n_items;
    for (int item = blockIdx.x; item < nit; item += gridDim.x) {
    __syncthreads();  // LDS of the previous item is no longer read
    const int home = items[item].x;
    const int h0 = cell_start[home] + items[item].y * chunk;
    const int nr = min(chunk, cell_start[home + 1] - h0);
    const int gid = cell_group[home];
    const DevPairGroup& G = groups[gid];
    const int cap = G.stride;
    const bool periodic = G.periodic;
    const double L[3] = {G.box[0], G.box[1], G.box[2]};
    const double iL[3] = {G.inv_box[0], G.inv_box[1], G.inv_box[2]};
    const int nc[3] = {G.ncell[0], G.ncell[1], G.ncell[2]};
    const int rch[3] = {G.reach[0], G.reach[1], G.reach[2]};
    // per-pair minimum image on periodic axes that have one cell
    const bool mi0 = periodic && nc[0] == 1, mi1 = periodic && nc[1] == 1, mi2 = periodic && nc[2] == 1;
    const bool any_mi = mi0 || mi1 || mi2;
    const float Lf0 = static_cast<float>(L[0]), Lf1 = static_cast<float>(L[1]), Lf2 = static_cast<float>(L[2]);
    const float iLf0 = static_cast<float>(iL[0]), iLf1 = static_cast<float>(iL[1]), iLf2 = static_cast<float>(iL[2]);
    const float rl2f = static_cast<float>(G.rl2) * kListMargin;
    const bool multi_cat = !(G.n_cat <= 1 && (G.allow[0] & 1u));   // else every member pair is allowed
    const int cbeg = G.cell_begin;
    // origin: the item's first row, wrapped into the box on periodic axes
    const PairRec r0 = rec[h0];
    const double o0 = periodic ? wrap_coord(r0.x, L[0], iL[0]) : r0.x;
    const double o1 = periodic ? wrap_coord(r0.y, L[1], iL[1]) : r0.y;
    const double o2 = periodic ? wrap_coord(r0.z, L[2], iL[2]) : r0.z;
    const int cell = home - cbeg;
    const int cx = cell % nc[0];
    const int cy = (cell / nc[0]) % nc[1];
    const int cz = cell / (nc[0] * nc[1]);
    const int w0 = 2 * rch[0] + 1, w1 = 2 * rch[1] + 1, w2 = 2 * rch[2] + 1;
    const int nst = w0 * w1 * w2;
    if (tid < nst) {
        const int dx = tid % w0 - rch[0], dy = (tid / w0) % w1 - rch[1], dz = tid / (w0 * w1) - rch[2];
        int x = cx + dx, y = cy + dy, z = cz + dz;
        float sx = 0.f, sy = 0.f, sz = 0.f;   // image shift of the cell's beads
        bool inside = true;
        if (periodic) {
            if (x < 0) { x += nc[0]; sx = -Lf0; } else if (x >= nc[0]) { x -= nc[0]; sx = Lf0; }
            if (y < 0) { y += nc[1]; sy = -Lf1; } else if (y >= nc[1]) { y -= nc[1]; sy = Lf1; }
            if (z < 0) { z += nc[2]; sz = -Lf2; } else if (z >= nc[2]) { z -= nc[2]; sz = Lf2; }
        } else {
            // open grid: stencil cells past its edge hold nothing
            inside = x >= 0 && x < nc[0] && y >= 0 && y < nc[1] && z >= 0 && z < nc[2];
        }
        const int c = cbeg + (z * nc[1] + y) * nc[0] + x;
        st_start[tid] = inside ? cell_start[c] : 0;
        st_pre[tid + 1] = inside ? cell_start[c + 1] - cell_start[c] : 0;
        st_shift[tid][0] = sx; st_shift[tid][1] = sy; st_shift[tid][2] = sz;
    }
    __syncthreads();
    if (tid == 0) {
        st_pre[0] = 0;
        for (int k = 0; k < nst; ++k) st_pre[k + 1] += st_pre[k];
    }
    __syncthreads();
    const int ntot = st_pre[nst];
    for (int r = tid; r < nr; r += blockDim.x) row_run[r] = 0;
    for (int wbeg = 0; wbeg < ntot; wbeg += kTileCap) {
        const int nw = min(kTileCap, ntot - wbeg);
        const int nw64 = (nw + 63) & ~63;
        __syncthreads();
        for (int k = tid; k < nw64; k += blockDim.x) {
            if (k >= nw) {   // padding: no category bit, never passes
                cpos[k] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
                cinfo[k] = make_int2(0, -1);
                continue;
            }
            const int flat = wbeg + k;
            int lo = 0, hi = nst - 1;  // last stencil cell with st_pre <= flat
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (st_pre[mid] <= flat) lo = mid; else hi = mid - 1;
            }
            const int sl = st_start[lo] + flat - st_pre[lo];
            const PairRec r = rec[sl];
            double x = r.x, y = r.y, z = r.z;
            if (periodic) {
                x = wrap_coord(x, L[0], iL[0]);
                y = wrap_coord(y, L[1], iL[1]);
                z = wrap_coord(z, L[2], iL[2]);
            }
            cpos[k] = make_float4(static_cast<float>(x - o0) + st_shift[lo][0],
                                  static_cast<float>(y - o1) + st_shift[lo][1],
                                  static_cast<float>(z - o2) + st_shift[lo][2], __int_as_float(sl));
            cinfo[k] = make_int2(static_cast<int>(1u << (r.type >> 16)), r.bead);
        }
        __syncthreads();
        // a wave takes kRowBlock rows at once: each staged candidate is read
        // from LDS once and tested against all of them
        for (int rb = wave * kRowBlock; rb < nr; rb += nwave * kRowBlock) {
            NlRow rw[kRowBlock];
            int run[kRowBlock];
#pragma unroll
            for (int q = 0; q < kRowBlock; ++q) {
                const int r = rb + q;
                if (r < nr) {
                    const int i = h0 + r;
                    const PairRec ri = rec[i];
                    rw[q].x = static_cast<float>((periodic ? wrap_coord(ri.x, L[0], iL[0]) : ri.x) - o0);
                    rw[q].y = static_cast<float>((periodic ? wrap_coord(ri.y, L[1], iL[1]) : ri.y) - o1);
                    rw[q].z = static_cast<float>((periodic ? wrap_coord(ri.z, L[2], iL[2]) : ri.z) - o2);
                    rw[q].amask = G.allow[ri.type >> 16];
                    rw[q].self = i;
                    rw[q].out = __builtin_amdgcn_make_buffer_rsrc(col + row_start[i], 0, cap * 4, 0x00020000);
                    const int m = slot_member[i];
                    const int eb = excl_start[m], ec = excl_count[m];
#pragma unroll
                    for (int e = 0; e < kExclReg; ++e) rw[q].ex[e] = e < ec ? excl[eb + e] : -2;
                    run[q] = row_run[r];
                } else {
                    rw[q].x = rw[q].y = rw[q].z = 0.f;
                    rw[q].amask = 0u;   // no candidate passes
                    rw[q].self = -1;
                    rw[q].out = __builtin_amdgcn_make_buffer_rsrc(col, 0, 0, 0x00020000);
#pragma unroll
                    for (int e = 0; e < kExclReg; ++e) rw[q].ex[e] = -2;
                    run[q] = 0;
                }
            }
#define OCM_SWEEP(MI, CAT) \
    nl_sweep<MI, CAT>(cpos, cinfo, nw, lane, rw, run, rl2f, Lf0, Lf1, Lf2, iLf0, iLf1, iLf2, mi0, mi1, mi2)
            if (any_mi) {
                if (multi_cat) OCM_SWEEP(true, true); else OCM_SWEEP(true, false);
            } else {
                if (multi_cat) OCM_SWEEP(false, true); else OCM_SWEEP(false, false);
            }
#undef OCM_SWEEP
#pragma unroll
            for (int q = 0; q < kRowBlock; ++q) {
                const int r = rb + q;
                if (r >= nr) continue;
                if (lane == 0) row_run[r] = run[q];
                // last window: pad the row to a multiple of 4 entries with -1
                if (wbeg + kTileCap >= ntot) {
                    const int padded = (run[q] + 3) & ~3;
                    if (padded <= cap && lane < padded - run[q]) col[row_start[h0 + r] + run[q] + lane] = -1;
                }
            }
        }
    }
    __syncthreads();
    // rows with more than kExclReg excluded partners: the written entries whose
    // bead is in the row's exclusion list become -1, which every reader skips
    for (int r = wave; r < nr; r += nwave) {
        const int i = h0 + r;
        const int m = slot_member[i];
        const int eb = excl_start[m], ec = excl_count[m];
        const int n = min(row_run[r], cap);
        if (ec <= kExclReg || n == 0) continue;
        const int base = row_start[i];
        const int mine = lane < ec ? excl[eb + lane] : -1;   // one partner per lane when <= 64
        for (int e = lane; e < n; e += 64) {
            const int sl = col[base + e];
            const int bj = rec[sl].bead;
            bool hit = false;
            if (ec <= 64) {
                for (int k = 0; k < ec; ++k) hit |= __builtin_amdgcn_readlane(mine, k) == bj;
            } else {
                hit = excluded(excl, eb, ec, bj);
            }
            if (hit) col[base + e] = -1;
        }
    }
    for (int r = tid; r < nr; r += blockDim.x) {
        const int padded = (row_run[r] + 3) & ~3;
        if (padded <= cap) {
            row_count[h0 + r] = padded;
        } else {
            row_count[h0 + r] = 0;  // the pair kernel reads nothing; the chunk is rebuilt wider
            atomicMax(&group_need[gid], padded);
            atomicOr(flags, kFlagListOverflow);
        }
    }
    }  // items
}

// -------------------------------------------------------------------- scan
// Exclusive scan of n ints in three launches: per-block scan (kScanItems
// items per block), scan of the block sums (one block), add-back.
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanItems = kScanThreads * kScanPerThread;

__device__ inline int block_exclusive_scan(int v, int* sh, int* block_total) {
    // v is this thread's sum; returns exclusive prefix among threads
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) sh[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) { const int t = sh[w]; sh[w] = run; run += t; }
        sh[kScanThreads / 64] = run;
    }
    __syncthreads();
    const int res = incl - v + sh[wid];
    *block_total = sh[kScanThreads / 64];
    __syncthreads();
    return res;
}

__global__ void k_scan_blocks(const int* __restrict__ in, int* __restrict__ out, int n,
                              int* __restrict__ partial) {
    __shared__ int sh[kScanThreads / 64 + 1];
    const int base = blockIdx.x * kScanItems + threadIdx.x * kScanPerThread;
    int vals[kScanPerThread];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        vals[k] = (base + k < n) ? in[base + k] : 0;
        sum += vals[k];
    }
    int tot;
    int run = block_exclusive_scan(sum, sh, &tot);
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        if (base + k < n) out[base + k] = run;
        run += vals[k];
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single block; scans nb partial sums in place and writes the grand total.
__global__ void k_scan_partials(int* __restrict__ partial, int nb, int* __restrict__ total) {
    __shared__ int sh[kScanThreads / 64 + 1];
    int carry = 0;
    for (int base = 0; base < nb; base += kScanItems) {
        const int b0 = base + threadIdx.x * kScanPerThread;
        int vals[kScanPerThread];
        int sum = 0;
#pragma unroll
        for (int k = 0; k < kScanPerThread; ++k) {
            vals[k] = (b0 + k < nb) ? partial[b0 + k] : 0;
            sum += vals[k];
        }
        int tot;
        int run = block_exclusive_scan(sum, sh, &tot) + carry;
#pragma unroll
        for (int k = 0; k < kScanPerThread; ++k) {
            if (b0 + k < nb) partial[b0 + k] = run;
            run += vals[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(int* __restrict__ out, int n, const int* __restrict__ partial) {
    const int add = partial[blockIdx.x];
    const int base = blockIdx.x * kScanItems + threadIdx.x * kScanPerThread;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k)
        if (base + k < n) out[base + k] += add;
}

// ----------------------------------------------------------------- bonded
template <typename R> struct V3T { R x, y, z; };

template <typename R>
__device__ inline V3T<R> bdisp(const double* __restrict__ px, const double* __restrict__ py,
                               const double* __restrict__ pz, int i, int j, const DevSys& S, bool pbc) {
    double dx = px[j] - px[i], dy = py[j] - py[i], dz = pz[j] - pz[i];
    if (pbc && S.periodic) {
        dx = min_image(dx, S.box[0], S.inv_box[0]);
        dy = min_image(dy, S.box[1], S.inv_box[1]);
        dz = min_image(dz, S.box[2], S.inv_box[2]);
    }
    return {static_cast<R>(dx), static_cast<R>(dy), static_cast<R>(dz)};
}
template <typename R> __device__ inline V3T<R> cross3(V3T<R> a, V3T<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> __device__ inline R dot3(V3T<R> a, V3T<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename R> __device__ inline R r_sqrt(R v);
template <> __device__ inline float r_sqrt<float>(float v) { return __builtin_amdgcn_sqrtf(v); }
template <> __device__ inline double r_sqrt<double>(double v) { return sqrt(v); }
template <typename R> __device__ inline R r_exp(R v);
template <> __device__ inline float r_exp<float>(float v) { return __expf(v); }
template <> __device__ inline double r_exp<double>(double v) { return exp(v); }
template <typename R> __device__ inline R r_atan2(R a, R b);
template <> __device__ inline float r_atan2<float>(float a, float b) { return atan2f(a, b); }
template <> __device__ inline double r_atan2<double>(double a, double b) { return atan2(a, b); }
// The force pass (R = float) takes the hardware's square root, reciprocal,
// exponential and sine/cosine (V_SQRT_F32, V_RCP_F32: 1 ulp; V_EXP_F32 on
// x log2(e); V_SIN/COS_F32 on x / 2pi, for the angles of dihedral terms) in
// place of the correctly rounded library sequences: a few instructions
// instead of 10-120, in a chain that sets the resident step's latency.  The
// energy pass (R = double) keeps the library functions and true division.
template <typename R> __device__ inline R r_div(R a, R b);
template <> __device__ inline float r_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <> __device__ inline double r_div<double>(double a, double b) { return a / b; }
template <typename R> __device__ inline R r_floor(R v);
template <> __device__ inline float r_floor<float>(float v) { return floorf(v); }
template <> __device__ inline double r_floor<double>(double v) { return floor(v); }
template <typename R> __device__ inline void r_sincos(R v, R* s, R* c);
template <> __device__ inline void r_sincos<float>(float v, float* s, float* c) { __sincosf(v, s, c); }
template <> __device__ inline void r_sincos<double>(double v, double* s, double* c) { sincos(v, s, c); }

struct ForceOut {
    long long* fx;
    long long* fy;
    long long* fz;
};

// A force contribution in fixed point, and its addition to bead b.  The last
// bead of a term takes the negated sum of the others' fixed-point words, so
// each term adds exactly zero net force (and converts one bead fewer).
struct FixV {
    long long x, y, z;
};
template <typename R>
__device__ inline FixV fix_v(R gx, R gy, R gz, int& bad) {
    if constexpr (sizeof(R) == 4) return {to_fix(gx, bad), to_fix(gy, bad), to_fix(gz, bad)};
    else return {to_fix_d(gx, bad), to_fix_d(gy, bad), to_fix_d(gz, bad)};
}
__device__ inline void add_fix(const ForceOut& o, int b, long long x, long long y, long long z) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&o.fx[b]), static_cast<unsigned long long>(x));
    atomicAdd(reinterpret_cast<unsigned long long*>(&o.fy[b]), static_cast<unsigned long long>(y));
    atomicAdd(reinterpret_cast<unsigned long long*>(&o.fz[b]), static_cast<unsigned long long>(z));
}

// One bonded term.  R = float for the force pass, double for the energy pass.
template <bool ENERGY, typename R>
__device__ inline void bonded_term(const BTerm& t, const DevSys& S, const double* __restrict__ px,
                                   const double* __restrict__ py, const double* __restrict__ pz,
                                   const FlaTable* __restrict__ fla, const FldTable* __restrict__ fld,
                                   const ForceOut& out, long long* __restrict__ efix, int sys,
                                   int& bad) {
    const bool pbc = (t.meta >> 5) & 1;
    const int group = t.meta & 31;
    const int table = t.meta >> 8;
    R e = 0;
    switch (t.kind) {
    case 1: case 2: case 3: case 8: {  // harmonic bond, Gaussian bond, Go contact, capped Go contact
        const V3T<double> dd = bdisp<double>(px, py, pz, t.idx[0], t.idx[1], S, pbc);
        const V3T<R> d = {static_cast<R>(dd.x), static_cast<R>(dd.y), static_cast<R>(dd.z)};
        const double r2d = fma(dd.z, dd.z, fma(dd.y, dd.y, dd.x * dd.x));
        const R r = r_sqrt<R>(static_cast<R>(r2d));
        // r - r0 of the two stiff kinds.  In FP32, r - r0 from an FP32 r
        // would carry an absolute error of ~r0 * 2^-24 (2e-8 nm), which a
        // 92,000 kJ/mol/nm^2 bond turns into 2e-3 kJ/mol/nm; (r^2 - r0^2) in
        // FP64 over (r + r0) keeps the relative error of r - r0 itself ~1e-7.
        auto stretch = [&](double r0d) -> R {
            if (sizeof(R) == sizeof(double)) return static_cast<R>(sqrt(r2d) - r0d);
            return r_div<R>(static_cast<R>(fma(-r0d, r0d, r2d)), r + static_cast<R>(r0d));
        };
        R dedr;
        if (t.kind == 1) {
            const R k = static_cast<R>(t.p[1]);
            const R dr = stretch(t.p[0]);
            e = R(0.5) * k * dr * dr;
            dedr = k * dr;
        } else if (t.kind == 2) {
            const R k = static_cast<R>(t.p[0]), s = static_cast<R>(t.p[2]);
            const R dr = stretch(t.p[1]);
            const R is2 = r_div<R>(R(1), s * s);
            e = k * r_exp<R>(-R(0.5) * dr * dr * is2);
            dedr = -e * dr * is2;
        } else if (t.kind == 3 || static_cast<R>(t.p[2]) - r_div<R>(r, static_cast<R>(t.p[1])) < R(0)) {
            // Go 12-10 (the capped kind beyond its cap, r / r0 > capping ratio)
            const R k = static_cast<R>(t.p[0]), r0 = static_cast<R>(t.p[1]);
            const R q = r_div<R>(r0, r);
            const R q2 = q * q, q4 = q2 * q2, q10 = q4 * q4 * q2, q12 = q10 * q2;
            e = k * (R(5) * q12 - R(6) * q10);
            dedr = r_div<R>(R(60) * k * (q10 - q12), r);
        } else {
            // capped Go contact inside its cap: the line through the 12-10
            // value and slope at r = capping_ratio * r0
            const R k = static_cast<R>(t.p[0]), r0 = static_cast<R>(t.p[1]);
            const R ic = r_div<R>(R(1), static_cast<R>(t.p[2]));
            const R ic2 = ic * ic, ic10 = ic2 * ic2 * ic2 * ic2 * ic2, ic12 = ic10 * ic2;
            const R slope = R(-60) * (ic12 * ic - ic10 * ic);
            const R intercept = R(65) * ic12 - R(66) * ic10;
            e = k * (slope * r_div<R>(r, r0) + intercept);
            dedr = r_div<R>(k * slope, r0);
        }
        if (!ENERGY) {
            const R s = r_div<R>(dedr, r);
            const FixV fi = fix_v<R>(d.x * s, d.y * s, d.z * s, bad);
            add_fix(out, t.idx[0], fi.x, fi.y, fi.z);
            add_fix(out, t.idx[1], -fi.x, -fi.y, -fi.z);
        }
        break;
    }
    case 4: case 5: {  // harmonic angle, flexible local angle
        const V3T<R> a = bdisp<R>(px, py, pz, t.idx[1], t.idx[0], S, pbc);  // x_i - x_j
        const V3T<R> c = bdisp<R>(px, py, pz, t.idx[1], t.idx[2], S, pbc);  // x_k - x_j
        const V3T<R> p = cross3(a, c);
        const R pn = r_sqrt<R>(dot3(p, p));
        const R th = r_atan2<R>(pn, dot3(a, c));
        R dedt;
        if (t.kind == 4) {
            const R t0 = static_cast<R>(t.p[0]), k = static_cast<R>(t.p[1]);
            e = R(0.5) * k * (th - t0) * (th - t0);
            dedt = k * (th - t0);
        } else {
            const FlaTable& T = fla[table];
            const R k = static_cast<R>(t.p[0]);
            const R tmin = static_cast<R>(T.tmin), tmax = static_cast<R>(T.tmax), h = static_cast<R>(T.h);
            R u = -static_cast<R>(T.emin), du = 0;
            if (th >= tmin && th <= tmax) {
                int lo = static_cast<int>(r_floor<R>(r_div<R>(th - tmin, h)));
                lo = lo < 0 ? 0 : (lo > 8 ? 8 : lo);
                const R xl = tmin + lo * h, xu = tmin + (lo + 1) * h;
                const R A = r_div<R>(xu - th, h), B = r_div<R>(th - xl, h);
                const R ylo = static_cast<R>(T.y[lo]), yhi = static_cast<R>(T.y[lo + 1]);
                const R dlo = static_cast<R>(T.y2[lo]), dhi = static_cast<R>(T.y2[lo + 1]);
                u += A * ylo + B * yhi + r_div<R>(((A * A * A - A) * dlo + (B * B * B - B) * dhi) * h * h, R(6));
                du += r_div<R>(yhi - ylo, h) + r_div<R>(((R(1) - R(3) * A * A) * dlo + (R(3) * B * B - R(1)) * dhi) * h, R(6));
            }
            if (tmin - th >= R(0)) { u += R(30) * (tmin - th) + static_cast<R>(T.y[0]); du -= R(30); }
            if (th - tmax >= R(0)) { u += R(30) * (th - tmax) + static_cast<R>(T.y[9]); du += R(30); }
            e = k * u;
            dedt = k * du;
        }
        if (!ENERGY && pn > R(0)) {
            const R ia = r_div<R>(R(1), dot3(a, a) * pn), ic = r_div<R>(R(1), dot3(c, c) * pn);
            const V3T<R> ap = cross3(a, p), pc = cross3(p, c);
            const V3T<R> gi = {ap.x * ia, ap.y * ia, ap.z * ia};
            const V3T<R> gk = {pc.x * ic, pc.y * ic, pc.z * ic};
            const FixV fi = fix_v<R>(-dedt * gi.x, -dedt * gi.y, -dedt * gi.z, bad);
            const FixV fk = fix_v<R>(-dedt * gk.x, -dedt * gk.y, -dedt * gk.z, bad);
            add_fix(out, t.idx[0], fi.x, fi.y, fi.z);
            add_fix(out, t.idx[2], fk.x, fk.y, fk.z);
            add_fix(out, t.idx[1], -(fi.x + fk.x), -(fi.y + fk.y), -(fi.z + fk.z));
        }
        break;
    }
    case 6: case 7: case 9: {  // Gaussian, flexible local and cosine dihedrals
        const V3T<R> b1 = bdisp<R>(px, py, pz, t.idx[0], t.idx[1], S, pbc);
        const V3T<R> b2 = bdisp<R>(px, py, pz, t.idx[1], t.idx[2], S, pbc);
        const V3T<R> b3 = bdisp<R>(px, py, pz, t.idx[2], t.idx[3], S, pbc);
        const V3T<R> n1 = cross3(b1, b2), n2 = cross3(b2, b3);
        const R b2n = r_sqrt<R>(dot3(b2, b2));
        const R phi = r_atan2<R>(b2n * dot3(b1, n2), dot3(n1, n2));
        R dedp;
        if (t.kind == 6) {
            const R k = static_cast<R>(t.p[0]), p0 = static_cast<R>(t.p[1]), s = static_cast<R>(t.p[2]);
            const R pi = static_cast<R>(3.1415926535897932385);
            const R dt = phi - p0;
            const R dtp = dt - r_floor<R>(r_div<R>(dt + pi, R(2) * pi)) * (R(2) * pi);
            const R is2 = r_div<R>(R(1), s * s);
            e = k * r_exp<R>(-R(0.5) * dtp * dtp * is2);
            dedp = -e * dtp * is2;
        } else if (t.kind == 9) {
            const R k = static_cast<R>(t.p[0]), t0 = static_cast<R>(t.p[1]), nn = static_cast<R>(t.p[2]);
            R sn, cs;
            r_sincos<R>(nn * (phi - t0), &sn, &cs);
            e = k * (R(1) - cs);
            dedp = k * nn * sn;
        } else {
            const FldTable& T = fld[table];
            const R k = static_cast<R>(t.p[0]);
            R s1, c1;
            r_sincos<R>(phi, &s1, &c1);
            const R s2 = R(2) * s1 * c1, c2 = c1 * c1 - s1 * s1;
            const R s3 = s1 * c2 + c1 * s2, c3 = c1 * c2 - s1 * s2;
            const R a1 = static_cast<R>(T.c[1]), b1c = static_cast<R>(T.c[2]);
            const R a2 = static_cast<R>(T.c[3]), b2c = static_cast<R>(T.c[4]);
            const R a3 = static_cast<R>(T.c[5]), b3c = static_cast<R>(T.c[6]);
            const R u = static_cast<R>(T.c[0] - T.c[7]) + a1 * c1 + b1c * s1 + a2 * c2 + b2c * s2 + a3 * c3 + b3c * s3;
            const R du = -a1 * s1 + b1c * c1 + R(2) * (-a2 * s2 + b2c * c2) + R(3) * (-a3 * s3 + b3c * c3);
            e = k * u;
            dedp = k * du;
        }
        if (!ENERGY) {
            const R n1s = dot3(n1, n1), n2s = dot3(n2, n2);
            if (n1s > R(0) && n2s > R(0) && b2n > R(0)) {
                const R fi = r_div<R>(-b2n, n1s), fl = r_div<R>(b2n, n2s);
                const V3T<R> gi = {n1.x * fi, n1.y * fi, n1.z * fi};
                const V3T<R> gl = {n2.x * fl, n2.y * fl, n2.z * fl};
                const R ib2 = r_div<R>(R(1), b2n * b2n);
                const R f1 = dot3(b1, b2) * ib2, f3 = dot3(b3, b2) * ib2;
                // middle beads: -(1 + f1) g_i + f3 g_l, and for bead k the
                // remainder f1 g_i - (1 + f3) g_l (the four sum to zero)
                const R a1 = -(R(1) + f1);
                const V3T<R> gj = {gi.x * a1 + gl.x * f3, gi.y * a1 + gl.y * f3, gi.z * a1 + gl.z * f3};
                const FixV wi = fix_v<R>(-dedp * gi.x, -dedp * gi.y, -dedp * gi.z, bad);
                const FixV wj = fix_v<R>(-dedp * gj.x, -dedp * gj.y, -dedp * gj.z, bad);
                const FixV wl = fix_v<R>(-dedp * gl.x, -dedp * gl.y, -dedp * gl.z, bad);
                add_fix(out, t.idx[0], wi.x, wi.y, wi.z);
                add_fix(out, t.idx[1], wj.x, wj.y, wj.z);
                add_fix(out, t.idx[2], -(wi.x + wj.x + wl.x), -(wi.y + wj.y + wl.y), -(wi.z + wj.z + wl.z));
                add_fix(out, t.idx[3], wl.x, wl.y, wl.z);
            }
        }
        break;
    }
    default:
        break;
    }
    if (ENERGY) *efix = to_fix_d(static_cast<double>(e), bad);   // the caller adds it to group `group` of `sys`
}

// Adds v to energy[key] for every lane with key >= 0: lanes with the same key
// are summed across the wavefront first, so there is one atomic per distinct
// key instead of one per lane (energies of a pass land on a few words).
// Every lane of the wavefront must call it.
__device__ inline void wave_add_energy(long long* energy, int key, long long v) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    unsigned long long todo = __ballot(key >= 0);
    while (todo) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const int lk = __shfl(key, leader, 64);
        const bool mine = key == lk;
        const unsigned long long peers = __ballot(mine) & todo;
        const long long sum = warp_sum(mine ? v : 0LL, 64);
        if (lane == leader)
            atomicAdd(reinterpret_cast<unsigned long long*>(&energy[lk]), static_cast<unsigned long long>(sum));
        todo &= ~peers;
    }
}

template <bool ENERGY>
__global__ void k_bonded(int nt, const BTerm* __restrict__ terms, const DevSys* __restrict__ sys,
                         const int* __restrict__ sys_of, const double* __restrict__ px,
                         const double* __restrict__ py, const double* __restrict__ pz,
                         const FlaTable* __restrict__ fla, const FldTable* __restrict__ fld,
                         ForceOut out, long long* __restrict__ energy, uint32_t groups,
                         int* __restrict__ err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    long long ef = 0;
    int key = -1;   // energy word of this term (ENERGY)
    if (t < nt) {
        const BTerm T = terms[t];
        if ((groups >> (T.meta & 31)) & 1u) {
            const int s = sys_of[T.idx[0]];
            const DevSys S = sys[s];
            int bad = 0;
            if (ENERGY) {
                bonded_term<true, double>(T, S, px, py, pz, fla, fld, out, &ef, s, bad);
                key = s * 32 + (T.meta & 31);
            } else {
                bonded_term<false, float>(T, S, px, py, pz, fla, fld, out, nullptr, s, bad);
            }
            if (bad) atomicOr(&err[s], kErrNumeric);
        }
    }
    if (ENERGY) wave_add_energy(energy, key, ef);
}

// ------------------------------------------------------------------- pair
// Reciprocals: the FP32 force pass uses the hardware v_rcp/v_rsq (1 ulp);
// the FP64 energy pass divides exactly.
// exp in the force pass: the hardware 2^x (v_exp_f32) of x*log2(e); its
// relative error stays near 1e-7 over the Debye-Hueckel range |x| <= 5.
template <typename R> __device__ inline R r_exp_fast(R v);
template <> __device__ inline float r_exp_fast<float>(float v) { return __builtin_amdgcn_exp2f(v * 1.44269504088896341f); }
template <> __device__ inline double r_exp_fast<double>(double v) { return exp(v); }

// Potential constants of a pair group in the arithmetic type of a pass, read
// once per block of rows rather than per pair.
template <typename R> struct PairK {
    R exv_eps, exv_corr, dh_K, dh_il, dh_corr, lja_rsw, lja_inv_sw;
    R u_sig2, u_eps;   // uniform kinds: sigma^2 and epsilon
    bool lja_switch;
    bool lj_shared;
};
template <typename R> __device__ inline PairK<R> pair_constants(const DevPairGroup& G) {
    PairK<R> k;
    k.exv_eps = static_cast<R>(G.exv_eps);
    k.exv_corr = static_cast<R>(G.exv_corr);
    k.dh_K = static_cast<R>(G.dh_K);
    k.dh_il = static_cast<R>(G.dh_inv_lambda);
    k.dh_corr = static_cast<R>(G.dh_corr);
    k.lja_rsw = static_cast<R>(G.lja_rsw);
    k.lja_inv_sw = static_cast<R>(G.lja_inv_sw);
    k.u_sig2 = static_cast<R>(G.u_sigma * G.u_sigma);
    k.u_eps = static_cast<R>(G.u_eps);
    k.lja_switch = G.lja_rsw > 0;
    k.lj_shared = G.lj_shared != 0;
    return k;
}

template <typename R> __device__ inline R r_rcp(R v);
template <> __device__ inline float r_rcp<float>(float v) { return __builtin_amdgcn_rcpf(v); }
template <> __device__ inline double r_rcp<double>(double v) { return 1.0 / v; }
template <typename R> __device__ inline R r_rsqrt(R v);
template <> __device__ inline float r_rsqrt<float>(float v) { return __builtin_amdgcn_rsqf(v); }
template <> __device__ inline double r_rsqrt<double>(double v) { return 1.0 / sqrt(v); }

// Energy terms and (dE/dr)/r of all kinds of a pair group for one pair.
// R = float (force pass, FP32 type rows tf) or double (energy pass, rows td).
template <typename R>
__device__ inline void pair_eval(int kinds, const PairK<R>& G, const float* __restrict__ tfi,
                                 const float* __restrict__ tfj, const double* __restrict__ tdi,
                                 const double* __restrict__ tdj, R r2, R* dedr_r, R e[4]) {
    auto P = [&](bool i_side, int k) -> R {
        if constexpr (sizeof(R) == 4) return i_side ? tfi[k] : tfj[k];
        else return i_side ? tdi[k] : tdj[k];
    };
    const R ir2 = r_rcp<R>(r2);
    R f = 0;
    R r = 0, ir = 0;
    if ((kinds & (kDh | kIsolf)) || ((kinds & (kLja | kUlja)) && G.lja_switch)) {
        ir = r_rsqrt<R>(r2);
        r = r2 * ir;
    }
    if (kinds & kExv) {
        const R s = P(true, kTyExv) + P(false, kTyExv);
        const R s2 = s * s;
        const R q2 = s2 * ir2;
        const R q6 = q2 * q2 * q2, q12 = q6 * q6;
        const R s6 = s2 * s2 * s2;
        const R eps = G.exv_eps;
        f += R(-12) * eps * q12 * ir2;
        e[0] = eps * (q12 - s6 * s6 * G.exv_corr);
    }
    if (kinds & kDh) {
        const R qq = P(true, kTyQ) * P(false, kTyQ);
        const R il = G.dh_il;
        const R ex = r_exp_fast<R>(-r * il);
        const R K = G.dh_K;
        f += -qq * K * ex * (ir + il) * ir2;
        e[1] = qq * K * (ex * ir - G.dh_corr);
    }
    const R c13 = static_cast<R>(1.2599210498948731648);  // 2^(1/3): (2^(1/6) sigma)^2 / sigma^2
    R s6 = 0, s12 = 0, sb_r = -1;
    if (kinds & kLjr) {
        const R eb = R(0.5) * (P(true, kTyLjrEps) + P(false, kTyLjrEps));
        const R sb = R(0.5) * (P(true, kTyLjrSig) + P(false, kTyLjrSig));
        const R s2 = sb * sb * ir2;
        s6 = s2 * s2 * s2;
        s12 = s6 * s6;
        sb_r = sb;
        if (c13 * sb * sb - r2 >= R(0)) {
            f += eb * (R(-48) * s12 + R(24) * s6) * ir2;
            e[2] = eb * (R(4) * (s12 - s6) + R(1));
        } else {
            e[2] = 0;
        }
    }
    if (kinds & kLja) {
        const R eb = R(0.5) * (P(true, kTyLjaEps) + P(false, kTyLjaEps));
        const R sb = R(0.5) * (P(true, kTyLjaSig) + P(false, kTyLjaSig));
        if (sb != sb_r) {  // Ashbaugh-Hatch pairs share sigma with the repulsive term
            const R s2 = sb * sb * ir2;
            s6 = s2 * s2 * s2;
            s12 = s6 * s6;
        }
        const R th2 = c13 * sb * sb;
        R u = 0, du = 0;  // du = (dU/dr)/r
        if (r2 - th2 >= R(0)) {
            u += R(4) * eb * (s12 - s6);
            du += eb * (R(-48) * s12 + R(24) * s6) * ir2;
        }
        if (th2 - r2 >= R(0)) u -= eb;
        if (G.lja_switch) {
            const R rsw = G.lja_rsw;
            if (r > rsw) {
                const R inv = G.lja_inv_sw;
                const R x = (r - rsw) * inv;
                const R x2 = x * x, x3 = x2 * x;
                const R S = R(1) - x3 * (R(10) - R(15) * x + R(6) * x2);
                const R dS = R(-30) * x2 * (R(1) - x) * (R(1) - x) * inv;
                du = du * S + u * dS * ir;
                u = u * S;
            }
        }
        f += du;
        e[3] = u;
    }
    // the kinds below each have a group of their own; their energy goes to e[0]
    if (kinds & kWca) {
        const R sb = R(0.5) * (P(true, kTySig) + P(false, kTySig));
        const R eb = r_sqrt<R>(P(true, kTyEps) * P(false, kTyEps));
        const R s2 = sb * sb * ir2, s6w = s2 * s2 * s2, s12w = s6w * s6w;
        if (c13 * sb * sb - r2 >= R(0)) {
            f += eb * (R(-48) * s12w + R(24) * s6w) * ir2;
            e[0] = eb * (R(4) * (s12w - s6w) + R(1));
        } else {
            e[0] = 0;
        }
    }
    if (kinds & kUwca) {
        const R s2 = G.u_sig2 * ir2, s6w = s2 * s2 * s2, s12w = s6w * s6w;
        f += G.u_eps * (R(-48) * s12w + R(24) * s6w) * ir2;
        e[0] = R(4) * G.u_eps * (s12w - s6w + R(0.25));
    }
    if (kinds & kUlja) {
        const R s2 = G.u_sig2 * ir2, s6u = s2 * s2 * s2, s12u = s6u * s6u;
        const R th2 = c13 * G.u_sig2;
        R u = 0, du = 0;
        if (r2 - th2 >= R(0)) {
            u += R(4) * G.u_eps * (s12u - s6u);
            du += G.u_eps * (R(-48) * s12u + R(24) * s6u) * ir2;
        }
        if (th2 - r2 >= R(0)) u -= G.u_eps;
        if (G.lja_switch && r > G.lja_rsw) {
            const R inv = G.lja_inv_sw;
            const R x = (r - G.lja_rsw) * inv;
            const R x2 = x * x, x3 = x2 * x;
            const R S = R(1) - x3 * (R(10) - R(15) * x + R(6) * x2);
            const R dS = R(-30) * x2 * (R(1) - x) * (R(1) - x) * inv;
            du = du * S + u * dS * ir;
            u = u * S;
        }
        f += du;
        e[0] = u;
    }
    if (kinds & kIsolf) {
        const R sb = R(0.5) * (P(true, kTySig) + P(false, kTySig));
        const R eb = r_sqrt<R>(P(true, kTyEps) * P(false, kTyEps));
        const R om = R(0.5) * (P(true, kTyOmega) + P(false, kTyOmega));
        const R b = static_cast<R>(1.1224620483093729814) * sb;   // 2^(1/6) sigma
        const R pi = static_cast<R>(3.1415926535897932385);
        R u = b - r >= R(0) ? -eb : R(0);
        if (r - b >= R(0) && b + om - r >= R(0)) {
            R sn, cs;
            r_sincos<R>(R(0.5) * pi * (r - b) / om, &sn, &cs);
            u -= eb * cs * cs;
            f += eb * R(2) * sn * cs * (R(0.5) * pi / om) * ir;
        }
        e[0] = u;
    }
    *dedr_r = f;
}

// The row's own type parameters, held in registers for the force pass.
// (A group of WCA or iSoLF reads its {sigma/2, sqrt(epsilon), omega/2} as
// exv, q and om.)
struct RowType {
    float exv, q, om, er, sr, ea, sa;
};
__device__ inline RowType row_type(const float* __restrict__ t) {
    return {t[kTyExv], t[kTyQ], t[kTyOmega], t[kTyLjrEps], t[kTyLjrSig], t[kTyLjaEps], t[kTyLjaSig]};
}

// (dE/dr)/r of the force pass.  Same terms as pair_eval<float>; when the
// repulsive and attractive LJ members of the group share sigma (Ashbaugh-
// Hatch) they are evaluated together: one set of sigma powers, one
// (24 s^6 - 48 s^12)/r^2 factor scaled by the repulsive epsilon inside r_min
// and by the attractive one outside, the switch on the attractive part.
// The FP32 type rows hold half of each LJ epsilon and sigma, so the
// Lorentz-Berthelot means are single adds (halving is exact, so the sums are
// the same bits as 0.5 * (a + b)).
// `kinds`, `shared` and `lsw` (the attractive switch is on) are compile-time
// constants in the specialized row loops, which then hold no branches; the
// switch polynomial is evaluated at x = max(0, (r - rsw)/w), where x = 0 gives
// S = 1 and dS = 0, the unswitched value exactly.
// (1/r is left in *ir_out when the kinds need it, else 0.)
__device__ __forceinline__ bool pair_force_has_ir(int kinds, bool lsw) {
    return (kinds & (kDh | kIsolf)) || ((kinds & (kLja | kUlja)) && lsw);
}
__device__ __forceinline__ float pair_force(int kinds, bool shared, bool lsw, const PairK<float>& G,
                                            const RowType& ti, const float* __restrict__ tj, float r2,
                                            float* ir_out = nullptr) {
    const float ir2 = __builtin_amdgcn_rcpf(r2);
    float f = 0.f, r = 0.f, ir = 0.f;
    if (pair_force_has_ir(kinds, lsw)) {
        ir = __builtin_amdgcn_rsqf(r2);
        r = r2 * ir;
    }
    if (ir_out) *ir_out = ir;
    if (kinds & kExv) {
        const float s = ti.exv + tj[kTyExv];
        const float q2 = s * s * ir2;
        const float q6 = q2 * q2 * q2, q12 = q6 * q6;
        f += -12.f * G.exv_eps * q12 * ir2;
    }
    if (kinds & kDh) {
        const float qq = ti.q * tj[kTyQ];
        const float ex = r_exp_fast<float>(-r * G.dh_il);
        f += -qq * G.dh_K * ex * (ir + G.dh_il) * ir2;
    }
    const float c13 = 1.2599210498948731648f;  // 2^(1/3): (2^(1/6) sigma)^2 / sigma^2
    if ((kinds & (kLjr | kLja)) == (kLjr | kLja) && shared) {
        const float4 lj = *reinterpret_cast<const float4*>(tj + kTyLjrEps);   // er, sr, ea, sa
        const float sb = ti.sr + lj.y;
        const float sb2 = sb * sb;
        const float s2 = sb2 * ir2;
        const float s6 = s2 * s2 * s2, s12 = s6 * s6;
        const float th2 = c13 * sb2;
        const float X = (24.f * s6 - 48.f * s12) * ir2;
        const float er = ti.er + lj.x, ea = ti.ea + lj.z;
        const bool inner = th2 - r2 >= 0.f, outer = r2 - th2 >= 0.f;
        f += inner ? er * X : 0.f;
        float u = outer ? 4.f * ea * (s12 - s6) : 0.f;
        float du = outer ? ea * X : 0.f;
        u = inner ? u - ea : u;
        if (lsw) {
            const float x = fmaxf((r - G.lja_rsw) * G.lja_inv_sw, 0.f);
            const float x2 = x * x, x3 = x2 * x;
            const float S = 1.f - x3 * (10.f - 15.f * x + 6.f * x2);
            const float dS = -30.f * x2 * (1.f - x) * (1.f - x) * G.lja_inv_sw;
            du = du * S + u * dS * ir;
        }
        f += du;
        return f;
    }
    float s6 = 0.f, s12 = 0.f, sb_r = -1.f;
    if (kinds & kLjr) {
        const float eb = ti.er + tj[kTyLjrEps];
        const float sb = ti.sr + tj[kTyLjrSig];
        const float s2 = sb * sb * ir2;
        s6 = s2 * s2 * s2;
        s12 = s6 * s6;
        sb_r = sb;
        if (c13 * sb * sb - r2 >= 0.f) f += eb * (-48.f * s12 + 24.f * s6) * ir2;
    }
    if (kinds & kLja) {
        const float eb = ti.ea + tj[kTyLjaEps];
        const float sb = ti.sa + tj[kTyLjaSig];
        if (sb != sb_r) {
            const float s2 = sb * sb * ir2;
            s6 = s2 * s2 * s2;
            s12 = s6 * s6;
        }
        const float th2 = c13 * sb * sb;
        float u = 0.f, du = 0.f;
        if (r2 - th2 >= 0.f) {
            u += 4.f * eb * (s12 - s6);
            du += eb * (-48.f * s12 + 24.f * s6) * ir2;
        }
        if (th2 - r2 >= 0.f) u -= eb;
        if (lsw) {
            const float x = fmaxf((r - G.lja_rsw) * G.lja_inv_sw, 0.f);
            const float x2 = x * x, x3 = x2 * x;
            const float S = 1.f - x3 * (10.f - 15.f * x + 6.f * x2);
            const float dS = -30.f * x2 * (1.f - x) * (1.f - x) * G.lja_inv_sw;
            du = du * S + u * dS * ir;
        }
        f += du;
    }
    // kinds with a group of their own (FP32 rows {sigma/2, sqrt(eps), omega/2})
    if (kinds & kWca) {
        const float sb = ti.exv + tj[kTySig];
        const float eb = ti.q * tj[kTyEps];
        const float sb2 = sb * sb;
        const float s2 = sb2 * ir2, s6w = s2 * s2 * s2, s12w = s6w * s6w;
        if (c13 * sb2 - r2 >= 0.f) f += eb * (-48.f * s12w + 24.f * s6w) * ir2;
    }
    if (kinds & kUwca) {
        const float s2 = G.u_sig2 * ir2, s6w = s2 * s2 * s2, s12w = s6w * s6w;
        f += G.u_eps * (-48.f * s12w + 24.f * s6w) * ir2;
    }
    if (kinds & kUlja) {
        const float s2 = G.u_sig2 * ir2, s6u = s2 * s2 * s2, s12u = s6u * s6u;
        const float th2 = c13 * G.u_sig2;
        float u = 0.f, du = 0.f;
        if (r2 - th2 >= 0.f) {
            u += 4.f * G.u_eps * (s12u - s6u);
            du += G.u_eps * (-48.f * s12u + 24.f * s6u) * ir2;
        }
        if (th2 - r2 >= 0.f) u -= G.u_eps;
        if (lsw) {
            const float x = fmaxf((r - G.lja_rsw) * G.lja_inv_sw, 0.f);
            const float x2 = x * x, x3 = x2 * x;
            const float S = 1.f - x3 * (10.f - 15.f * x + 6.f * x2);
            const float dS = -30.f * x2 * (1.f - x) * (1.f - x) * G.lja_inv_sw;
            du = du * S + u * dS * ir;
        }
        f += du;
    }
    if (kinds & kIsolf) {
        const float sb = ti.exv + tj[kTySig];
        const float eb = ti.q * tj[kTyEps];
        const float om = ti.om + tj[kTyOmega];
        const float b = 1.1224620483093729814f * sb;   // 2^(1/6) sigma
        if (r - b >= 0.f && b + om - r >= 0.f) {
            // dE/dr = eb (pi / (2 om)) sin(pi (r - b) / om); the hardware sine
            // takes revolutions, and (r - b) / (2 om) lies in [0, 1/2]
            const float iom = __builtin_amdgcn_rcpf(om);
            f += eb * (1.5707963267948966f * iom) * __builtin_amdgcn_sinf(0.5f * (r - b) * iom) * ir;
        }
    }
    return f;
}

// Pair contributions to fixed point.  fma(g * 2^32, d, 1.5 * 2^52) rounds the
// exact product g*d*2^32 to the nearest integer (ties to even) as long as
// |g d| < 2^19; that integer is the double's bit pattern minus the constant's.
// Lanes add the raw bit patterns (64-bit, wrapping) and remove n constants at
// the end.  Rounding to nearest even is odd-symmetric, so pair (j, i) adds
// exactly the negation of what pair (i, j) adds.
constexpr double kFixMagic = 6755399441055744.0;          // 1.5 * 2^52
constexpr unsigned long long kFixMagicBits = 0x4338000000000000ull;
constexpr float kPairLimit = 524288.0f;                   // 2^19 kJ/mol/nm per pair

// Bonded terms folded into the force pass of k_pair (block-table entries
// {-1, first term}).
struct BondedArgs {
    int nt;
    const BTerm* terms;
    const DevSys* sys;
    const int* sys_of;
    const double* px;
    const double* py;
    const double* pz;
    const FlaTable* fla;
    const FldTable* fld;
};

// Force-pass row loops, specialized by pair-group kind so the per-entry body
// has no branches and the four entries of a round interleave.
// kSpecGeneric covers any mix of the four round kinds (EXV, DH, LJ
// repulsive, LJ attractive), kSpecSolo the kinds that sit alone in a group.
constexpr int kSpecGeneric = 0, kSpecExvPeriodic = 1, kSpecExvOpen = 2, kSpecDh = 3, kSpecLjShared = 4,
              kSpecSolo = 5;

// Raw buffer over the pair records: 32-bit offsets instead of 64-bit
// addresses, and a read past the end (the -1 padding entries) returns zeros
// without touching memory (gfx9 buffer resource word 3 = 0x00020000, stride 0,
// range in bytes).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rec_rsrc(const PairRec* rec, int nrec) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<PairRec*>(rec), 0, nrec * static_cast<int>(sizeof(PairRec)),
                                             0x00020000);
}

__device__ __forceinline__ PairRec load_rec(__amdgpu_buffer_rsrc_t rs, int slot) {
    const unsigned off = static_cast<unsigned>(slot) * static_cast<unsigned>(sizeof(PairRec));   // -1: past the end
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(rs, static_cast<int>(off), 0, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b128(rs, static_cast<int>(off + 16u), 0, 0);
    PairRec r;
    r.x = __hiloint2double(static_cast<int>(a[1]), static_cast<int>(a[0]));
    r.y = __hiloint2double(static_cast<int>(a[3]), static_cast<int>(a[2]));
    r.z = __hiloint2double(static_cast<int>(b[1]), static_cast<int>(b[0]));
    r.type = static_cast<int>(b[2]);
    r.bead = static_cast<int>(b[3]);
    return r;
}

// Column entry e of a row of cnt (> 0) entries, -1 past its end: the load is
// clamped into the row, so it is issued unconditionally (no branch around it).
__device__ __forceinline__ int next_col(const int* __restrict__ crow, int e, int cnt) {
    const int v = crow[min(e, cnt - 1)];
    return e < cnt ? v : -1;
}

// FP32 box-wrapped position of slot j (prune input); slot -1 reads past the
// range and returns zeros.
__device__ __forceinline__ float4 load_rec32(__amdgpu_buffer_rsrc_t rs, int j) {
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(rs, static_cast<int>(static_cast<unsigned>(j) * 16u), 0, 0);
    return make_float4(__uint_as_float(a[0]), __uint_as_float(a[1]), __uint_as_float(a[2]), 0.f);
}

struct RowGeom {
    double bx, by, bz, ibx, iby, ibz, rc2, rin2;
    bool periodic;
};

struct RowAcc {
    unsigned long long ux = 0, uy = 0, uz = 0, nadd = 0;   // raw fma bit patterns
    int bad = 0;
};

// One list entry of a row: the displacement d = x_j - x_i (minimum image
// applied in place), and g * 2^32 (g = (dE/dr)/r, so that g d is the force on
// i) in FP64, or 0 when the entry is padding (valid false) or beyond the
// cutoff.  *bad is set when a pair inside the cutoff exceeded the per-pair
// limit.  The expression is symmetric in i and j, so the entry (j, i) gives
// the same g and the negated d.
template <int SPEC>
__device__ __forceinline__ double pair_entry(const RowGeom& g, int kinds, const PairK<float>& Kf, const RowType& rt,
                                             const float* __restrict__ tfj, bool valid, double& dx, double& dy,
                                             double& dz, int& bad) {
    // a group without a box has inverse box 0, for which the minimum
    // image is the identity
    if (SPEC == kSpecGeneric || SPEC == kSpecSolo ? g.periodic : SPEC != kSpecExvOpen) {
        dx = min_image(dx, g.bx, g.ibx);
        dy = min_image(dy, g.by, g.iby);
        dz = min_image(dz, g.bz, g.ibz);
    }
    const double r2 = dist2(dx, dy, dz);
    const bool in = valid && r2 < g.rc2;
    const float r2f = in ? static_cast<float>(r2) : 1.0f;
    float f, ir = 0.f;
    if (SPEC == kSpecExvPeriodic || SPEC == kSpecExvOpen) f = pair_force(kExv, false, false, Kf, rt, tfj, r2f);
    else if (SPEC == kSpecDh) f = pair_force(kDh, false, false, Kf, rt, tfj, r2f, &ir);
    else if (SPEC == kSpecLjShared) f = pair_force(kLjr | kLja, true, true, Kf, rt, tfj, r2f, &ir);
    else if (SPEC == kSpecSolo) f = pair_force(kinds & ~(kExv | kDh | kLjr | kLja), false, Kf.lja_switch, Kf, rt,
                                               tfj, r2f);
    else f = pair_force(kinds & (kExv | kDh | kLjr | kLja), Kf.lj_shared, Kf.lja_switch, Kf, rt, tfj, r2f);
    // |f d| <= |f| r: one range test covers the three components.  A
    // pair beyond the limit is flagged and clamped to it along its
    // direction (so relaxation still pushes the beads apart); NaN
    // contributes nothing.  Where the kinds already took 1/r the clamp
    // is formed every time and selected, with no branch in the loop.
    const bool ok = f * f * r2f < kPairLimit * kPairLimit;
    bad |= in && !ok;
    float gg;
    if (SPEC == kSpecDh || SPEC == kSpecLjShared) {
        const float cl = copysignf(0.999f * kPairLimit, f) * ir;
        gg = ok ? f : (f == f ? cl : 0.0f);
    } else {
        gg = ok ? f : (f == f ? copysignf(0.999f * kPairLimit, f) * __builtin_amdgcn_rsqf(r2f) : 0.0f);
    }
    gg = in ? gg : 0.0f;   // a masked entry adds exactly the bias constant
    // scaling by 2^32 is exact in FP32 (the result stays far below
    // FP32's largest value; a subnormal gg, flushed or not, rounds to
    // a zero contribution either way), so one FP32 multiply replaces
    // an FP64 one with the same bits
    return static_cast<double>(gg * 4294967296.0f);
}

// The pair kernel's spec (specialized row loop) of a group with these kinds.
__device__ __forceinline__ int pair_spec(int kinds, bool periodic, const PairK<float>& Kf) {
    if (kinds == kExv) return periodic ? kSpecExvPeriodic : kSpecExvOpen;
    if (kinds == kDh && periodic) return kSpecDh;
    if (kinds == (kLjr | kLja) && Kf.lj_shared && Kf.lja_switch && periodic) return kSpecLjShared;
    if (kinds & ~(kExv | kDh | kLjr | kLja)) return kSpecSolo;
    return kSpecGeneric;
}

// One member row for lanes `lane` of a team of TEAM: forces of its entries
// in fixed point (see kFixMagic).  U entries per lane per round (4, or 2:
// rounds of 2*TEAM entries waste less of a row's last round, with fewer
// gathers in flight; the autotuner picks).
template <int TEAM, int SPEC, int U>
__device__ __forceinline__ void force_row(const RowGeom& g, int kinds, const PairK<float>& Kf, const RowType& rt,
                                          const PairRec& ri, int beg, int cnt, int lane,
                                          const PairRec* __restrict__ rec, int nrec, const int* __restrict__ col,
                                          const float* __restrict__ lds_types, RowAcc& acc) {
    // A round covers 4*TEAM entries of the row; lane l takes entries l, l+T,
    // l+2T, l+3T of it, so the lanes of one gather read neighbouring entries,
    // which point at neighbouring (cell-ordered) records: a few cache lines
    // per gather instead of one per lane.  Entries past the row end read as
    // -1 (padding).  The round's 4 record gathers are all issued before any
    // of them is used (the scheduling barrier keeps the arithmetic below
    // them), and the next round's column indices are fetched while this
    // round computes.
    // (the next round's loads stay conditional here: clamped unconditional
    // loads, as in the prune, measured 4-6% slower in this loop)
    const int* crow = col + beg;
    const __amdgpu_buffer_rsrc_t rs = rec_rsrc(rec, nrec);
    int cn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) cn[u] = lane + u * TEAM < cnt ? crow[lane + u * TEAM] : -1;
    for (int base = 0; base < cnt; base += U * TEAM) {
        int cs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) cs[u] = cn[u];
        PairRec rjs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) rjs[u] = load_rec(rs, cs[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + U * TEAM + lane + u * TEAM;
            cn[u] = e < cnt ? crow[e] : -1;
        }
        __builtin_amdgcn_sched_barrier(0);
        acc.nadd += U;
        // the round's entries are evaluated without branches (an entry that is
        // padding or beyond the cutoff is computed at r2 = 1 and masked to
        // zero) so their dependency chains interleave
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const PairRec rj = rjs[u];
            double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
            const float* tfj = lds_types + (rj.type & 0xFFFF) * kTypeStride;
            const double gs = pair_entry<SPEC>(g, kinds, Kf, rt, tfj, cs[u] >= 0, dx, dy, dz, acc.bad);
            acc.ux += static_cast<unsigned long long>(__double_as_longlong(fma(gs, dx, kFixMagic)));
            acc.uy += static_cast<unsigned long long>(__double_as_longlong(fma(gs, dy, kFixMagic)));
            acc.uz += static_cast<unsigned long long>(__double_as_longlong(fma(gs, dz, kFixMagic)));
        }
    }
}

// Inner-list prune of one member row for lanes `lane` of a team of TEAM: its
// outer entries within cutoff + margin, written in list order to the same row
// slots of the inner list (col_w) and padded to a multiple of 4.  Lane l takes
// entries l, l+T, l+2T, l+3T of each round of 4T (as in the force pass).
// F32 (periodic groups): box-wrapped FP32 positions, one 16-byte gather per
// entry, kept while r^2 < rin2_wide, which exceeds (cutoff + margin)^2 by more
// than the FP32 coordinates' error, so the inner row still holds every entry
// an exact FP64 test would keep.  Otherwise the FP64 records and rin2.
template <int TEAM, bool F32>
__device__ __forceinline__ void prune_row(const DevPairGroup& G, int row, int lane, unsigned long long team_lanes,
                                          const PairRec* __restrict__ rec,
                                          const float4* __restrict__ rec32, const int* __restrict__ row_start,
                                          const int* __restrict__ row_count, const int* __restrict__ col,
                                          int* __restrict__ row_count_w, int* __restrict__ col_w, int nrec) {
    const int beg = row_start[row];
    const int cnt = row_count[row];
    double xi = 0, yi = 0, zi = 0;
    float4 qi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (F32) qi = rec32[row];
    else { xi = rec[row].x; yi = rec[row].y; zi = rec[row].z; }
    const float fbx = static_cast<float>(G.box[0]), fby = static_cast<float>(G.box[1]);
    const float fbz = static_cast<float>(G.box[2]);
    const float fibx = static_cast<float>(G.inv_box[0]), fiby = static_cast<float>(G.inv_box[1]);
    const float fibz = static_cast<float>(G.inv_box[2]);
    const float rin2f = static_cast<float>(G.rin2_wide);
    // every lane of a team runs the same rounds (the loop bound is the
    // row's), so the ballots below see the whole team
    int n = 0;   // entries the team has written to the inner row
    if (cnt == 0) {
        if (lane == 0) row_count_w[row] = 0;
        return;
    }
    const int* crow = col + beg;
    int* wrow = col_w + beg;
    const __amdgpu_buffer_rsrc_t rs32 = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float4*>(rec32), 0, F32 ? nrec * 16 : 0, 0x00020000);
    const __amdgpu_buffer_rsrc_t rs = rec_rsrc(rec, F32 ? 0 : nrec);
    if (F32) {
        // software-pipelined: round r's tests run while round r+1's gathers
        // and round r+2's column indices are in flight (the kernel has few
        // wavefronts per SIMD to hide a gather behind)
        int cs[4], cn[4];
        float4 qj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) cs[u] = next_col(crow, lane + u * TEAM, cnt);
#pragma unroll
        for (int u = 0; u < 4; ++u) cn[u] = next_col(crow, 4 * TEAM + lane + u * TEAM, cnt);
#pragma unroll
        for (int u = 0; u < 4; ++u) qj[u] = load_rec32(rs32, cs[u]);
        for (int base = 0; base < cnt; base += 4 * TEAM) {
            float4 qn[4];
            int cn2[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) qn[u] = load_rec32(rs32, cn[u]);   // -1 reads past the range: zeros
#pragma unroll
            for (int u = 0; u < 4; ++u) cn2[u] = next_col(crow, base + 8 * TEAM + lane + u * TEAM, cnt);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float dx = qj[u].x - qi.x, dy = qj[u].y - qi.y, dz = qj[u].z - qi.z;
                dx = fmaf(-fbx, rintf(dx * fibx), dx);
                dy = fmaf(-fby, rintf(dy * fiby), dy);
                dz = fmaf(-fbz, rintf(dz * fibz), dz);
                const bool in = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) < rin2f;
                const bool valid = cs[u] >= 0, keep = valid & in;
                const unsigned long long m =
                    __builtin_amdgcn_ballot_w64(valid) & __builtin_amdgcn_ballot_w64(in) & team_lanes;
                const int pos = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                          __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
                if (keep) wrow[n + pos] = cs[u];
                n += __popcll(m);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) { cs[u] = cn[u]; cn[u] = cn2[u]; qj[u] = qn[u]; }
        }
    }
    int cn[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) cn[u] = next_col(crow, lane + u * TEAM, cnt);
    for (int base = 0; base < (F32 ? 0 : cnt); base += 4 * TEAM) {
        const int cs[4] = {cn[0], cn[1], cn[2], cn[3]};
        double px[4], py[4], pz[4];
        float4 qj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // padding (-1) reads past the buffer's range: zeros, and keep is false
            if (F32) qj[u] = load_rec32(rs32, cs[u]);
            else {
                const PairRec r = load_rec(rs, cs[u]);
                px[u] = r.x; py[u] = r.y; pz[u] = r.z;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) cn[u] = next_col(crow, base + 4 * TEAM + lane + u * TEAM, cnt);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            bool in;
            if (F32) {
                float dx = qj[u].x - qi.x, dy = qj[u].y - qi.y, dz = qj[u].z - qi.z;
                dx = fmaf(-fbx, rintf(dx * fibx), dx);
                dy = fmaf(-fby, rintf(dy * fiby), dy);
                dz = fmaf(-fbz, rintf(dz * fibz), dz);
                in = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) < rin2f;
            } else {
                const double dx = min_image(px[u] - xi, G.box[0], G.inv_box[0]);
                const double dy = min_image(py[u] - yi, G.box[1], G.inv_box[1]);
                const double dz = min_image(pz[u] - zi, G.box[2], G.inv_box[2]);
                in = dist2(dx, dy, dz) < G.rin2;
            }
            // entry base + u*T + l precedes base + u*T + l' for l < l'; the
            // wave mask is the AND of the two tests' own masks
            const bool valid = cs[u] >= 0, keep = valid & in;
            const unsigned long long m =
                __builtin_amdgcn_ballot_w64(valid) & __builtin_amdgcn_ballot_w64(in) & team_lanes;
            const int pos = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
            if (keep) wrow[n + pos] = cs[u];
            n += __popcll(m);
        }
    }
    if (lane == 0) {
        int p = n;
        for (; p & 3; ++p) wrow[p] = -1;
        row_count_w[row] = p;
    }
}

// Inner-list prune: a team of TEAM lanes per member row, rows taken in order
// (row = global member index), independent of the force pass's block table
// so its team can differ.
template <int TEAM>
__global__ __launch_bounds__(kPairBlock) void k_prune(int nm, const DevPairGroup* __restrict__ groups,
                                                      const int* __restrict__ member_group,
                                                      const PairRec* __restrict__ rec,
                                                      const float4* __restrict__ rec32,
                                                      const int* __restrict__ row_start,
                                                      const int* __restrict__ row_count,
                                                      const int* __restrict__ col, int* __restrict__ row_count_w,
                                                      int* __restrict__ col_w) {
    const int wl = static_cast<int>(threadIdx.x & 63u);
    // the team's lanes within the wave
    const unsigned long long team_lanes = TEAM == 64 ? ~0ull : (((1ull << TEAM) - 1ull) << (wl / TEAM * TEAM));
    const int lane = threadIdx.x % TEAM;
    const int row = static_cast<int>((static_cast<long long>(blockIdx.x) * kPairBlock + threadIdx.x) / TEAM);
    if (row >= nm) return;   // whole teams leave together
    const DevPairGroup& G = groups[member_group[row]];
    if (G.periodic) prune_row<TEAM, true>(G, row, lane, team_lanes, rec, rec32, row_start, row_count, col,
                                          row_count_w, col_w, nm);
    else prune_row<TEAM, false>(G, row, lane, team_lanes, rec, rec32, row_start, row_count, col,
                                row_count_w, col_w, nm);
}

// At least 5 waves per SIMD (<= 96 VGPRs): the record gathers need the latency hiding.
// SLICED: walk the table in XCD slices (xcd_walk); otherwise a plain stride.
template <bool ENERGY, int TEAM, bool SLICED, int U>
__global__ __launch_bounds__(kPairBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_pair(
    const DevPairGroup* __restrict__ groups, const int2* __restrict__ blocks,
    const PairRec* __restrict__ rec, const int* __restrict__ row_start,
    const int* __restrict__ row_count, const int* __restrict__ col,
    const float* __restrict__ types_f, const double* __restrict__ types_d, ForceBuf out,
    long long* __restrict__ energy, uint32_t gmask, int* __restrict__ err, int nblocks, int nrec,
    unsigned long long* __restrict__ stamp, BondedArgs bd) {
    extern __shared__ __attribute__((aligned(16))) float lds_types[];
    int loaded = -1;
    // optional duration stamps: each workgroup stores its start and end on the
    // device's constant-rate wall clock (plain stores to its own two words)
    if (stamp && threadIdx.x == 0) stamp[2 * blockIdx.x] = static_cast<unsigned long long>(wall_clock64());
    // fixed grid: each workgroup walks the block table with stride gridDim.x,
    // or its XCD's slice of it
    const XcdWalk walk = SLICED ? xcd_walk(nblocks) : XcdWalk{static_cast<int>(blockIdx.x), nblocks,
                                                              static_cast<int>(gridDim.x)};
    for (int bi = walk.first; bi < walk.end; bi += walk.stride) {
    const int2 bt = blocks[bi];
    if (bt.x < 0) {
        // a chunk of bonded terms, one per thread (force pass only): their
        // gathers and atomics overlap the pair rows of the other workgroups
        if (!ENERGY) {
            const int t = bt.y + static_cast<int>(threadIdx.x);
            if (t < bd.nt) {
                const BTerm T = bd.terms[t];
                if ((gmask >> (T.meta & 31)) & 1u) {
                    const int s = bd.sys_of[T.idx[0]];
                    int bad = 0;
                    bonded_term<false, float>(T, bd.sys[s], bd.px, bd.py, bd.pz, bd.fla, bd.fld,
                                              ForceOut{out.fx, out.fy, out.fz}, nullptr, s, bad);
                    if (bad) atomicOr(&err[s], kErrNumeric);
                }
            }
        }
        continue;
    }
    const DevPairGroup& G = groups[bt.x];
    const RowGeom geo{G.box[0], G.box[1], G.box[2], G.inv_box[0], G.inv_box[1], G.inv_box[2], G.rc2, G.rin2,
                      G.periodic != 0};
    const PairK<float> Kf = pair_constants<float>(G);
    const PairK<double> Kd = pair_constants<double>(G);
    if (!ENERGY && bt.x != loaded) {  // bt.x is uniform over the workgroup
        __syncthreads();
        const float* src = types_f + static_cast<size_t>(G.type_begin) * kTypeStride;
        for (int k = threadIdx.x; k < G.n_types * kTypeStride; k += blockDim.x) lds_types[k] = src[k];
        __syncthreads();
        loaded = bt.x;
    }
    int kinds = 0;
#pragma unroll
    for (int k = 0; k < kPairKinds; ++k)
        if (((G.kinds >> k) & 1) && ((gmask >> G.egroup[k]) & 1u)) kinds |= 1 << k;
    const int lane = threadIdx.x % TEAM;
    const int row = bt.y + threadIdx.x / TEAM;
    const int row_end = G.member_begin + G.member_count;
    RowAcc acc;
    long long ae0 = 0, ae1 = 0, ae2 = 0, ae3 = 0;
    int bead_i = -1;
    if (row < row_end && kinds) {
        const PairRec ri = rec[row];
        bead_i = ri.bead;
        const int beg = row_start[row], cnt = row_count[row];
        const int ti = ri.type & 0xFFFF;
        if (!ENERGY) {
            const RowType rt = row_type(lds_types + ti * kTypeStride);
            // the group's kind selects the loop (uniform over the workgroup)
            switch (pair_spec(kinds, geo.periodic, Kf)) {
            case kSpecExvPeriodic:
                force_row<TEAM, kSpecExvPeriodic, U>(geo, kinds, Kf, rt, ri, beg, cnt, lane, rec, nrec, col,
                                                         lds_types, acc);
                break;
            case kSpecExvOpen:
                force_row<TEAM, kSpecExvOpen, U>(geo, kinds, Kf, rt, ri, beg, cnt, lane, rec, nrec, col,
                                                     lds_types, acc);
                break;
            case kSpecDh:
                force_row<TEAM, kSpecDh, U>(geo, kinds, Kf, rt, ri, beg, cnt, lane, rec, nrec, col,
                                                lds_types, acc);
                break;
            case kSpecLjShared:
                force_row<TEAM, kSpecLjShared, U>(geo, kinds, Kf, rt, ri, beg, cnt, lane, rec, nrec, col,
                                                      lds_types, acc);
                break;
            case kSpecSolo:
                force_row<TEAM, kSpecSolo, U>(geo, kinds, Kf, rt, ri, beg, cnt, lane, rec, nrec, col,
                                                  lds_types, acc);
                break;
            default:
                force_row<TEAM, kSpecGeneric, U>(geo, kinds, Kf, rt, ri, beg, cnt, lane, rec, nrec, col,
                                                     lds_types, acc);
                break;
            }
        } else {
            // energy pass: FP64 terms of each pair once (row < column)
            const double* tdi = types_d + static_cast<size_t>(G.type_begin + ti) * kTypeStride;
            const int4* col4 = reinterpret_cast<const int4*>(col + beg);
            for (int k = lane; 4 * k < cnt; k += TEAM) {
                const int4 c4 = col4[k];
                const int cs[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int s = cs[u];
                    if (s < 0 || row >= s) continue;
                    const PairRec rj = rec[s];
                    double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
                    if (geo.periodic) {
                        dx = min_image(dx, geo.bx, geo.ibx);
                        dy = min_image(dy, geo.by, geo.iby);
                        dz = min_image(dz, geo.bz, geo.ibz);
                    }
                    const double r2 = dist2(dx, dy, dz);
                    if (!(r2 < geo.rc2)) continue;
                    double f;
                    double e[4] = {0, 0, 0, 0};
                    const double* tdj = types_d + static_cast<size_t>(G.type_begin + (rj.type & 0xFFFF)) * kTypeStride;
                    pair_eval<double>(kinds, Kd, nullptr, nullptr, tdi, tdj, r2, &f, e);
                    ae0 += to_fix_d(e[0], acc.bad);
                    ae1 += to_fix_d(e[1], acc.bad);
                    ae2 += to_fix_d(e[2], acc.bad);
                    ae3 += to_fix_d(e[3], acc.bad);
                }
            }
        }
    }
    if (ENERGY) {
        // every row of a wavefront belongs to this block's group: one sum and
        // one atomic per kind for the whole wavefront
        ae0 = warp_sum(ae0, 64);
        ae1 = warp_sum(ae1, 64);
        ae2 = warp_sum(ae2, 64);
        ae3 = warp_sum(ae3, 64);
        if ((threadIdx.x & 63u) == 0) {
            const long long v[4] = {ae0, ae1, ae2, ae3};
            // kinds 4.. are alone in their groups and leave their energy in slot 0
#pragma unroll
            for (int k = 0; k < kPairKinds; ++k)
                if ((kinds >> k) & 1)
                    atomicAdd(reinterpret_cast<unsigned long long*>(&energy[G.sys * 32 + G.egroup[k]]),
                              static_cast<unsigned long long>(v[k < 4 ? k : 0]));
        }
    } else {
        const unsigned long long off = acc.nadd * kFixMagicBits;
        long long ax = static_cast<long long>(acc.ux - off);
        long long ay = static_cast<long long>(acc.uy - off);
        long long az = static_cast<long long>(acc.uz - off);
        ax = warp_sum(ax, TEAM);
        ay = warp_sum(ay, TEAM);
        az = warp_sum(az, TEAM);
        if (lane == 0 && bead_i >= 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&out.fx[bead_i]), static_cast<unsigned long long>(ax));
            atomicAdd(reinterpret_cast<unsigned long long*>(&out.fy[bead_i]), static_cast<unsigned long long>(ay));
            atomicAdd(reinterpret_cast<unsigned long long*>(&out.fz[bead_i]), static_cast<unsigned long long>(az));
        }
    }
    if (acc.bad) atomicOr(&err[G.sys], kErrNumeric);
    }  // block table
    if (stamp) {
        __syncthreads();
        if (threadIdx.x == 0) stamp[2 * blockIdx.x + 1] = static_cast<unsigned long long>(wall_clock64());
    }
}

// ------------------------------------------------------------ resident step
// Small systems (up to kResMaxBeads beads): one workgroup per system runs a
// whole chunk of steps on chip.  Positions and the fixed-point force words
// live in LDS, the forces are summed with LDS atomics, and the neighbor list
// (each pair once, i < j in member order, packed as
// {bead i, bead j, type row i, type row j} in 10 + 10 + 6 + 6 bits) is built
// inside the kernel by testing every member pair of the group, again whenever
// a bead has moved more than skin/2 since the last build.  The per-entry
// force (pair_entry), the bonded terms (bonded_term) and the per-bead step
// (baoab_bead_g with gauss3_fast noise) are the standard path's own
// functions, and a pair's contribution to j is the exact negation of its
// contribution to i, so every force word, and with it every step, has the
// same bits as the standard path.
// Everything a step reads sits in LDS where it fits (gfx950 gives a
// workgroup up to 160 KB): the system's bonded terms up to term_cap, its list
// up to list_cap entries (the rest stays in global memory), the groups'
// constants and type rows.  A step is two barriers: the integrator (thread b
// steps bead b), then the force pass, whose work -- bonded terms, list
// entries, and the next step's noise of every bead -- is dealt out as one
// sequence over all threads, so that no thread holds more than its share and
// the step's critical path is one item long when there are enough threads.
struct ResArgs {
    const ResSys* tab;
    const ResGroup* grp;
    const BTerm* terms;        // bead indices local to their system
    const float* rows;         // FP32 type rows, kTypeStride floats each
    const DevSys* sys;
    const DevPairGroup* groups;
    const int* member_bead;
    const int* member_type;    // type | category << 16
    const int* local_id;
    const double* inv_m;
    const double* bead_hs2;    // per bead: (resident skin / 2)^2
    double *px, *py, *pz, *vx, *vy, *vz;
    long long *fx, *fy, *fz;
    uint32_t* list;
    const FlaTable* fla;
    const FldTable* fld;
    const long long* gstep;
    int* err;
    int nsteps;
    int nbs;                   // LDS entries per bead array (>= every system's beads, multiple of 64)
    int term_cap;              // bonded terms staged in LDS per system
    int list_cap;              // list entries kept in LDS per system
    const int* excl_tab;       // the groups' exclusion tables (ResGroup::exo), by system
    int ex_cap;                // their LDS words per system
};

// A pair group's constants, staged in LDS for the whole chunk.
struct ResGroupK {
    RowGeom geo;
    PairK<float> Kf;
    long long l0;      // list region in global memory
    double rl2;             // (cutoff + resident skin)^2
    int kinds, spec, m0, nm, trow, gid, exo;
    int lds_off, lds_cnt;   // the group's entries kept in LDS: [lds_off, lds_off + lds_cnt)
    int cnt;                // entries of the last build
    float rl2w;
};

__host__ __device__ constexpr size_t res_align16(size_t v) { return (v + 15) & ~static_cast<size_t>(15); }

// LDS layout: x, y, z (doubles), f (3 nbs force words; the build's FP32
// positions lie over them), the next step's noise (3 nbs floats: gauss3_fast
// rounds its samples through FP32, so a float holds them exactly), mem and
// lid (nbs ints each), type rows, group constants, flags, terms, list.
__host__ __device__ constexpr size_t res_lds_fixed(int nbs) {
    return res_align16(static_cast<size_t>(nbs) * (6 * 8 + 3 * 4 + 2 * 4) +
                       kResMaxTypeRows * kTypeStride * sizeof(float) + kResMaxGroups * sizeof(ResGroupK) +
                       8 * sizeof(int));
}
__host__ __device__ constexpr size_t res_lds_bytes(int nbs, int term_cap, int list_cap, int ex_cap) {
    return res_lds_fixed(nbs) + res_align16(static_cast<size_t>(ex_cap) * sizeof(int)) +
           static_cast<size_t>(term_cap) * sizeof(BTerm) + static_cast<size_t>(list_cap) * sizeof(uint32_t);
}

struct ResLds {
    int nbs;
    double *x, *y, *z;
    long long* f;      // fx [0, nbs), fy [nbs, 2 nbs), fz [2 nbs, 3 nbs)
    float4* p32;       // the build's FP32 positions, over f (dead while the build runs)
    float* g;          // noise: g0 [0, nbs), g1 [nbs, 2 nbs), g2 [2 nbs, 3 nbs)
    int* mem;          // packed member info of the group being built
    int* lid;          // the beads' indices within their system (noise counters)
    float* rows;
    ResGroupK* gk;
    int* flag;         // [0] rebuild, [1] FP64 build, [2] entries kept in LDS so far
    int* ex;           // the system's exclusion tables
    BTerm* terms;
    uint32_t* list;
};

__device__ __forceinline__ ResLds res_lds(unsigned char* base, int nbs, int term_cap, int ex_cap) {
    ResLds L;
    L.nbs = nbs;
    L.x = reinterpret_cast<double*>(base);
    L.y = L.x + nbs;
    L.z = L.y + nbs;
    L.f = reinterpret_cast<long long*>(L.z + nbs);
    L.p32 = reinterpret_cast<float4*>(L.f);
    L.g = reinterpret_cast<float*>(L.f + 3 * nbs);
    L.mem = reinterpret_cast<int*>(L.g + 3 * nbs);
    L.lid = L.mem + nbs;
    L.rows = reinterpret_cast<float*>(L.lid + nbs);
    L.gk = reinterpret_cast<ResGroupK*>(L.rows + kResMaxTypeRows * kTypeStride);
    L.flag = reinterpret_cast<int*>(L.gk + kResMaxGroups);
    L.ex = reinterpret_cast<int*>(base + res_lds_fixed(nbs));
    L.terms = reinterpret_cast<BTerm*>(base + res_lds_fixed(nbs) + res_align16(static_cast<size_t>(ex_cap) * sizeof(int)));
    L.list = reinterpret_cast<uint32_t*>(L.terms + term_cap);
    return L;
}

// Lists of every group of the system at the positions in LDS.  All threads
// call it; it overwrites the force words (p32 lies over them) and ends with a
// barrier, after which they may be written again.  Entries go to LDS while
// the system's list_cap lasts, then to the group's global region.
// The member pairs (i, j > i) are dealt over all threads: rows r and
// nm - 1 - r together hold nm - 1 candidates, and each such row pair is cut
// into S segments so that about 2T segments cover the group; a thread tests
// its segments' candidates one after another against the row's exclusions (a
// mask bit for the next kResNearExcl members, a search of the short sorted
// list of farther ones, both in LDS).
template <int T>
__device__ void res_build(const ResArgs& a, const ResSys& R, const DevSys& S, const ResLds& L) {
    const int tid = threadIdx.x;
    if (R.ng == 0) return;   // uniform; nothing is written
    if (tid == 0) { L.flag[1] = 0; L.flag[2] = 0; }
    __syncthreads();
    if (tid < R.nb) {
        double x = L.x[tid], y = L.y[tid], z = L.z[tid];
        if (S.periodic) {
            x = wrap_coord(x, S.box[0], S.inv_box[0]);
            y = wrap_coord(y, S.box[1], S.inv_box[1]);
            z = wrap_coord(z, S.box[2], S.inv_box[2]);
        } else {
            x -= L.x[0];
            y -= L.y[0];
            z -= L.z[0];
        }
        L.p32[tid] = make_float4(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z), 0.f);
        if (!(fabs(x) <= kResWide && fabs(y) <= kResWide && fabs(z) <= kResWide)) L.flag[1] = 1;
    }
    for (int g = 0; g < R.ng; ++g) {
        ResGroupK& K = L.gk[g];
        const DevPairGroup& G = a.groups[K.gid];
        if (tid == 0) K.cnt = 0;
        if (tid < K.nm) {
            const int m = K.m0 + tid;
            const int mt = a.member_type[m];
            L.mem[tid] = (a.member_bead[m] - R.b0) | ((K.trow + (mt & 0xFFFF)) << 10) | ((mt >> 16) << 16);
        }
        __syncthreads();
        const bool wide = L.flag[1] != 0;
        const bool per = K.geo.periodic;
        const int nm = K.nm;
        const int lds_base = L.flag[2];
        const int lds_room = a.list_cap - lds_base;
        uint32_t* glist = a.list + K.l0;
        const float Lx = static_cast<float>(G.box[0]), Ly = static_cast<float>(G.box[1]);
        const float Lz = static_cast<float>(G.box[2]);
        const float iLx = static_cast<float>(G.inv_box[0]), iLy = static_cast<float>(G.inv_box[1]);
        const float iLz = static_cast<float>(G.inv_box[2]);
        const int nrp = (nm + 1) / 2;                       // row pairs
        const int nseg = max(1, (2 * T + nrp - 1) / max(nrp, 1));   // segments per row pair
        const int seglen = (nm - 1 + nseg - 1) / nseg;
        const int* exm = L.ex + K.exo;
        const int* far = exm + 2 * nm + 1;
        // consecutive threads take different row pairs: the entries one wave
        // appends together (consecutive list slots, read together by a wave of
        // the force pass) then hit different force words.  A wavefront's kept
        // candidates take their slots with one counter atomic.
        const int lane = tid & 63;
        auto append = [&](bool keep, uint32_t w) {
            const unsigned long long m = __ballot(keep);   // over the active lanes
            if (m == 0ull) return;                         // the same for every active lane
            const int leader = __ffsll(static_cast<long long>(m)) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&K.cnt, __popcll(m));
            base = __shfl(base, leader, 64);
            if (keep) {
                const int k = base + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                                         static_cast<uint32_t>(m >> 32),
                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
                if (k < lds_room) L.list[lds_base + k] = w;
                else glist[k] = w;
            }
        };
        auto excluded = [&](int i, int j, uint32_t near_mask, int f0, int f1) {
            const int d = j - i;
            if (d <= kResNearExcl) return ((near_mask >> (d - 1)) & 1u) != 0u;
            bool hit = false;
            for (int k = f0; k < f1; ++k) hit |= far[k] == j;
            return hit;
        };
        for (int u = tid; u < nrp * nseg; u += T) {
            const int r = u % nrp, sg = u / nrp;
            const int r2 = nm - 1 - r;
            const int split = nm - 1 - r;   // q < split: row r, candidate r + 1 + q; else row r2, q + 1
            const int len = r2 == r ? split : nm - 1;   // the middle row of an odd count stands alone
            const int q0 = sg * seglen, q1 = min(q0 + seglen, len);
            if (q0 >= q1) continue;
            // both rows' values: a segment crosses from row r to row r2 at most once
            const int ma = L.mem[r], mb = L.mem[r2];
            const int bia = ma & 1023, bib = mb & 1023;
            const int tia = (ma >> 10) & 63, tib = (mb >> 10) & 63;
            const uint32_t alla = G.allow[ma >> 16], allb = G.allow[mb >> 16];
            const uint32_t nma = static_cast<uint32_t>(exm[r]), nmb = static_cast<uint32_t>(exm[r2]);
            const int f0a = exm[nm + r], f1a = exm[nm + r + 1], f0b = exm[nm + r2], f1b = exm[nm + r2 + 1];
            if (!wide) {
                const float4 pa = L.p32[bia], pb = L.p32[bib];
                // four candidates a round, their loads issued before any test
                for (int q = q0; q < q1; q += 4) {
                    int mj[4];
                    float4 pj[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int qq = min(q + k, q1 - 1);   // past the end: the last again, not kept
                        mj[k] = L.mem[qq < split ? r + 1 + qq : qq + 1];
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) pj[k] = L.p32[mj[k] & 1023];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int qq = q + k;
                        const bool first = qq < split;
                        const float4 pi = first ? pa : pb;
                        float dx = pj[k].x - pi.x, dy = pj[k].y - pi.y, dz = pj[k].z - pi.z;
                        if (per) {
                            dx = fmaf(-Lx, rintf(dx * iLx), dx);
                            dy = fmaf(-Ly, rintf(dy * iLy), dy);
                            dz = fmaf(-Lz, rintf(dz * iLz), dz);
                        }
                        const uint32_t allow = first ? alla : allb;
                        bool keep = qq < q1 && fmaf(dz, dz, fmaf(dy, dy, dx * dx)) < K.rl2w &&
                                    ((allow >> (mj[k] >> 16)) & 1u);
                        const int i = first ? r : r2, j = first ? r + 1 + qq : qq + 1;
                        if (keep) keep = !excluded(i, j, first ? nma : nmb, first ? f0a : f0b, first ? f1a : f1b);
                        const uint32_t w = static_cast<uint32_t>(first ? bia : bib) |
                                           (static_cast<uint32_t>(mj[k] & 1023) << 10) |
                                           (static_cast<uint32_t>(first ? tia : tib) << 20) |
                                           (static_cast<uint32_t>((mj[k] >> 10) & 63) << 26);
                        append(keep, w);
                    }
                }
            } else {
                // a coordinate beyond kResWide: the FP64 test, one candidate at a time
                for (int q = q0; q < q1; ++q) {
                    const bool first = q < split;
                    const int i = first ? r : r2, j = first ? r + 1 + q : q + 1;
                    const int bi = first ? bia : bib;
                    const int mj = L.mem[j];
                    const int bj = mj & 1023;
                    double dx = L.x[bj] - L.x[bi], dy = L.y[bj] - L.y[bi], dz = L.z[bj] - L.z[bi];
                    if (per) {
                        dx = min_image(dx, G.box[0], G.inv_box[0]);
                        dy = min_image(dy, G.box[1], G.inv_box[1]);
                        dz = min_image(dz, G.box[2], G.inv_box[2]);
                    }
                    const uint32_t allow = first ? alla : allb;
                    bool keep = dist2(dx, dy, dz) < K.rl2 && ((allow >> (mj >> 16)) & 1u);
                    if (keep) keep = !excluded(i, j, first ? nma : nmb, first ? f0a : f0b, first ? f1a : f1b);
                    const uint32_t w = static_cast<uint32_t>(bi) | (static_cast<uint32_t>(bj) << 10) |
                                       (static_cast<uint32_t>(first ? tia : tib) << 20) |
                                       (static_cast<uint32_t>((mj >> 10) & 63) << 26);
                    append(keep, w);
                }
            }
        }
        __syncthreads();   // the count is final; the next group reuses mem
        if (tid == 0) {
            K.lds_off = lds_base;
            K.lds_cnt = min(K.cnt, max(lds_room, 0));
            L.flag[2] = lds_base + K.lds_cnt;
        }
        __syncthreads();   // p32 is no longer read; the LDS offset is published
    }
}

// The first item of the range [off, off + n) of the step's work sequence that
// thread tid takes (items u with u % T == tid), as an index into the range.
template <int T>
__device__ __forceinline__ int res_first(int off) {
    return static_cast<int>((static_cast<unsigned>(threadIdx.x) - static_cast<unsigned>(off)) % T);
}

template <int SPEC, int T>
__device__ __forceinline__ void res_pairs(const ResArgs& a, const ResGroupK& K, int off, const ResLds& L, int& bad) {
    // a thread with no entry of this group reads none of its constants
    const int e0 = res_first<T>(off);
    if (e0 >= K.cnt) return;
    const uint32_t* glist = a.list + K.l0;
    const uint32_t* llist = L.list + K.lds_off;
    const int cnt = K.cnt, lcnt = K.lds_cnt;
    const RowGeom geo = K.geo;
    const PairK<float> Kf = K.Kf;
    const int kinds = K.kinds;
    const int nbs = L.nbs;
    unsigned long long* F = reinterpret_cast<unsigned long long*>(L.f);
    auto entry = [&](uint32_t w) {
        const int bi = w & 1023, bj = (w >> 10) & 1023, ti = (w >> 20) & 63, tj = w >> 26;
        double dx = L.x[bj] - L.x[bi], dy = L.y[bj] - L.y[bi], dz = L.z[bj] - L.z[bi];
        const RowType rt = row_type(L.rows + ti * kTypeStride);
        const double gs = pair_entry<SPEC>(geo, kinds, Kf, rt, L.rows + tj * kTypeStride, true, dx, dy, dz, bad);
        const unsigned long long cx =
            static_cast<unsigned long long>(__double_as_longlong(fma(gs, dx, kFixMagic))) - kFixMagicBits;
        const unsigned long long cy =
            static_cast<unsigned long long>(__double_as_longlong(fma(gs, dy, kFixMagic))) - kFixMagicBits;
        const unsigned long long cz =
            static_cast<unsigned long long>(__double_as_longlong(fma(gs, dz, kFixMagic))) - kFixMagicBits;
        atomicAdd(F + bi, cx);
        atomicAdd(F + nbs + bi, cy);
        atomicAdd(F + 2 * nbs + bi, cz);
        atomicAdd(F + bj, 0ull - cx);
        atomicAdd(F + nbs + bj, 0ull - cy);
        atomicAdd(F + 2 * nbs + bj, 0ull - cz);
    };
    // entries kept in LDS, the rest read through a buffer resource over the
    // group's global region (a different load than LDS's, so neither becomes
    // a generic load, and the entry body is inlined once)
    const __amdgpu_buffer_rsrc_t grs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(glist), 0, cnt * static_cast<int>(sizeof(uint32_t)), 0x00020000);
    for (int e = e0; e < cnt; e += T) {
        uint32_t w;
        if (e < lcnt) w = llist[e];
        else w = static_cast<uint32_t>(__builtin_amdgcn_raw_buffer_load_b32(grs, e * 4, 0, 0));
        entry(w);
    }
}

// The force pass, and (noise != 0) every bead's noise for global step
// `nstep`: the sequence terms, list entries group by group, beads.
template <int T>
__device__ __forceinline__ void res_forces(const ResArgs& a, const ResSys& R, const DevSys& S, const ResLds& L,
                                           int noise, long long nstep, int& bad) {
    const ForceOut out{L.f, L.f + L.nbs, L.f + 2 * L.nbs};
    const int nl = min(R.nt, a.term_cap);
    // terms past term_cap through a buffer resource (as the list entries)
    static_assert(sizeof(BTerm) == 48, "a term is three 16-byte loads");
    const __amdgpu_buffer_rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<BTerm*>(a.terms + R.t0), 0, R.nt * static_cast<int>(sizeof(BTerm)), 0x00020000);
    // Each run of one bonded kind, and each pair group below, starts its
    // share of the work on a wavefront boundary (the work sequence is padded
    // to multiples of 64), so no wavefront runs two kinds' code one after the
    // other.  A slot finds its run by an unrolled walk over the run table,
    // which the system record holds in scalar registers (a loop indexing it
    // would load it from memory once per run and step).
    int off = 0;
#pragma unroll
    for (int q = 0; q < kResMaxSeg; ++q)
        if (q < R.nseg) off += (R.seg_n[q] + 63) & ~63;
    for (int v = threadIdx.x; v < off; v += T) {
        int t = -1, base = 0;
#pragma unroll
        for (int q = 0; q < kResMaxSeg; ++q) {
            if (q < R.nseg) {
                if (v >= base && v < base + R.seg_n[q]) t = R.seg_t[q] + (v - base);
                base += (R.seg_n[q] + 63) & ~63;
            }
        }
        if (t < 0) continue;   // padding of a run
        BTerm tm;
        if (t < nl) {
            tm = L.terms[t];
        } else {
            uint32_t w[12];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const auto u = __builtin_amdgcn_raw_buffer_load_b128(trs, t * 48 + 16 * k, 0, 0);
                w[4 * k] = u[0]; w[4 * k + 1] = u[1]; w[4 * k + 2] = u[2]; w[4 * k + 3] = u[3];
            }
            __builtin_memcpy(&tm, w, sizeof(tm));
        }
        bonded_term<false, float>(tm, S, L.x, L.y, L.z, a.fla, a.fld, out, nullptr, R.sys, bad);
    }
    for (int g = 0; g < R.ng; ++g) {
        const ResGroupK& K = L.gk[g];
        switch (K.spec) {
        case kSpecExvPeriodic: res_pairs<kSpecExvPeriodic, T>(a, K, off, L, bad); break;
        case kSpecExvOpen: res_pairs<kSpecExvOpen, T>(a, K, off, L, bad); break;
        case kSpecDh: res_pairs<kSpecDh, T>(a, K, off, L, bad); break;
        case kSpecLjShared: res_pairs<kSpecLjShared, T>(a, K, off, L, bad); break;
        case kSpecSolo: res_pairs<kSpecSolo, T>(a, K, off, L, bad); break;
        default: res_pairs<kSpecGeneric, T>(a, K, off, L, bad); break;
        }
        off += (K.cnt + 63) & ~63;
    }
    if (noise) {
        for (int b = res_first<T>(off); b < R.nb; b += T) {
            double g[3];
            gauss3_fast(S.seed, static_cast<uint64_t>(nstep - S.step_origin), static_cast<uint32_t>(L.lid[b]), g);
            L.g[b] = static_cast<float>(g[0]);
            L.g[L.nbs + b] = static_cast<float>(g[1]);
            L.g[2 * L.nbs + b] = static_cast<float>(g[2]);
        }
    }
}

// One workgroup of T threads per system; thread b owns bead b (nb <= T).
// On entry the force words hold F(x) of the current positions and the
// velocities are synchronized (no half kick owed); on exit the same holds
// nsteps later.  Every barrier is reached by every thread: the step count
// and the rebuild decision (read from LDS after a barrier) are uniform.
template <int T>
__global__ __launch_bounds__(T) void k_resident(ResArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char res_base[];
    const ResLds L = res_lds(res_base, a.nbs, a.term_cap, a.ex_cap);
    const ResSys R = a.tab[blockIdx.x];
    const DevSys S = a.sys[R.sys];
    const int tid = threadIdx.x;
    const int nbs = L.nbs;
    const bool own = tid < R.nb;
    const int gb = R.b0 + tid;
    double X[3] = {0, 0, 0}, V[3] = {0, 0, 0}, ref[3] = {0, 0, 0};
    double im = 0.0, hs2 = 0.0;
    int bad = 0;
    const long long step0 = a.gstep[0];
    if (own) {
        X[0] = a.px[gb]; X[1] = a.py[gb]; X[2] = a.pz[gb];
        V[0] = a.vx[gb]; V[1] = a.vy[gb]; V[2] = a.vz[gb];
        im = a.inv_m[gb];
        hs2 = a.bead_hs2[gb];
        L.x[tid] = X[0]; L.y[tid] = X[1]; L.z[tid] = X[2];
        const int lid = a.local_id[gb];
        L.lid[tid] = lid;
        double g[3];
        gauss3_fast(S.seed, static_cast<uint64_t>(step0 - S.step_origin), static_cast<uint32_t>(lid), g);
        L.g[tid] = static_cast<float>(g[0]);
        L.g[nbs + tid] = static_cast<float>(g[1]);
        L.g[2 * nbs + tid] = static_cast<float>(g[2]);
    }
    for (int k = tid; k < R.nrow * kTypeStride; k += T) L.rows[k] = a.rows[static_cast<size_t>(R.row0) * kTypeStride + k];
    {
        // the system's bonded terms, 12 words each, as far as term_cap holds
        const int nw = min(R.nt, a.term_cap) * static_cast<int>(sizeof(BTerm) / 4);
        const int* src = reinterpret_cast<const int*>(a.terms + R.t0);
        int* dst = reinterpret_cast<int*>(L.terms);
        for (int k = tid; k < nw; k += T) dst[k] = src[k];
    }
    for (int k = tid; k < R.exn; k += T) L.ex[k] = a.excl_tab[R.ex0 + k];
    if (tid < R.ng) {
        const ResGroup Q = a.grp[R.g0 + tid];
        const DevPairGroup& G = a.groups[Q.gid];
        ResGroupK& K = L.gk[tid];
        K.geo = RowGeom{G.box[0], G.box[1], G.box[2], G.inv_box[0], G.inv_box[1], G.inv_box[2], G.rc2, G.rin2,
                        G.periodic != 0};
        K.Kf = pair_constants<float>(G);
        K.l0 = Q.l0;
        K.kinds = G.kinds;
        K.spec = pair_spec(G.kinds, G.periodic != 0, K.Kf);
        K.m0 = Q.m0;
        K.nm = Q.nm;
        K.trow = Q.trow;
        K.gid = Q.gid;
        K.exo = Q.exo;
        K.rl2 = Q.rl2;
        K.lds_off = 0;
        K.lds_cnt = 0;
        K.cnt = 0;
        K.rl2w = Q.rl2w;
    }
    if (tid == 0) L.flag[0] = 0;
    __syncthreads();
    res_build<T>(a, R, S, L);
    ref[0] = X[0]; ref[1] = X[1]; ref[2] = X[2];
    if (own) {
        L.f[tid] = a.fx[gb];
        L.f[nbs + tid] = a.fy[gb];
        L.f[2 * nbs + tid] = a.fz[gb];
    }
    __syncthreads();
    for (int s = 0; s < a.nsteps; ++s) {
        if (own) {
            const long long F0 = L.f[tid], F1 = L.f[nbs + tid], F2 = L.f[2 * nbs + tid];
            L.f[tid] = 0;   // read: the force pass after the barrier refills them
            L.f[nbs + tid] = 0;
            L.f[2 * nbs + tid] = 0;
            const double g[3] = {L.g[tid], L.g[nbs + tid], L.g[2 * nbs + tid]};
            double xn[3], vn[3];
            if (!baoab_bead_g(S, im, F0, F1, F2, X, V, s > 0 ? 1 : 0, g, xn, vn)) bad = 1;
            X[0] = xn[0]; X[1] = xn[1]; X[2] = xn[2];
            V[0] = vn[0]; V[1] = vn[1]; V[2] = vn[2];
            L.x[tid] = X[0]; L.y[tid] = X[1]; L.z[tid] = X[2];
            const double dx = X[0] - ref[0], dy = X[1] - ref[1], dz = X[2] - ref[2];
            if (!(dist2(dx, dy, dz) <= hs2)) L.flag[0] = 1;
        }
        __syncthreads();
        if (L.flag[0] != 0) {   // uniform: every thread reads it after the barrier
            res_build<T>(a, R, S, L);
            ref[0] = X[0]; ref[1] = X[1]; ref[2] = X[2];
            if (tid == 0) L.flag[0] = 0;   // read by all before the build's barriers
            if (own) {   // the build's FP32 positions lay over the force words
                L.f[tid] = 0;
                L.f[nbs + tid] = 0;
                L.f[2 * nbs + tid] = 0;
            }
            __syncthreads();
        }
        // the noise slots were read by the integrator before the barrier
        res_forces<T>(a, R, S, L, s + 1 < a.nsteps ? 1 : 0, step0 + s + 1, bad);
        __syncthreads();
    }
    if (own) {
        const long long F0 = L.f[tid], F1 = L.f[nbs + tid], F2 = L.f[2 * nbs + tid];
        if (im != 0.0) {   // the closing half kick, as k_half_kick
            V[0] = half_kick(V[0], S.half_dt, F0 * kFixInv, im);
            V[1] = half_kick(V[1], S.half_dt, F1 * kFixInv, im);
            V[2] = half_kick(V[2], S.half_dt, F2 * kFixInv, im);
        }
        a.px[gb] = X[0]; a.py[gb] = X[1]; a.pz[gb] = X[2];
        a.vx[gb] = V[0]; a.vy[gb] = V[1]; a.vz[gb] = V[2];
        a.fx[gb] = F0; a.fy[gb] = F1; a.fz[gb] = F2;
    }
    if (bad) atomicOr(&a.err[R.sys], kErrNumeric);
}

// Both step slots advance by n after a resident chunk.
__global__ void k_gstep_advance(long long* gstep, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const long long s = gstep[0] + n;
        gstep[0] = s;
        gstep[1] = s;
    }
}

// Launch q of a chunk left nblk {start, end} pairs at stamp + 2*nblk*q; one
// workgroup per launch reduces them to {earliest start, latest end} in out.
__global__ __launch_bounds__(256) void k_reduce_stamps(const unsigned long long* __restrict__ stamp, int nblk,
                                                       unsigned long long* __restrict__ out) {
    __shared__ unsigned long long lo[4], hi[4];
    const unsigned long long* s = stamp + 2 * static_cast<size_t>(nblk) * blockIdx.x;
    unsigned long long a = ~0ull, b = 0ull;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        a = min(a, s[2 * i]);
        b = max(b, s[2 * i + 1]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        a = min(a, static_cast<unsigned long long>(__shfl_xor(a, o)));
        b = max(b, static_cast<unsigned long long>(__shfl_xor(b, o)));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { lo[w] = a; hi[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long A = lo[0], B = hi[0];
        for (int k = 1; k < static_cast<int>(blockDim.x >> 6); ++k) { A = min(A, lo[k]); B = max(B, hi[k]); }
        out[2 * blockIdx.x] = A;
        out[2 * blockIdx.x + 1] = B;
    }
}

// ------------------------------------------------------------- energies
__global__ void k_kinetic(int nb, const int* __restrict__ sys_of, const double* __restrict__ mass,
                          const double* __restrict__ vx, const double* __restrict__ vy,
                          const double* __restrict__ vz, long long* __restrict__ ke,
                          int* __restrict__ err) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double e = 0.5 * mass[b] * (vx[b] * vx[b] + vy[b] * vy[b] + vz[b] * vz[b]);
    int bad = 0;
    const long long q = to_fix_d(e, bad);
    atomicAdd(reinterpret_cast<unsigned long long*>(&ke[sys_of[b]]), static_cast<unsigned long long>(q));
    if (bad) atomicOr(&err[sys_of[b]], kErrNumeric);
}

// Lists the interacting pairs (r < cutoff) of one pair group as bead pairs
// i<j, using the same distance test as k_pair.
__global__ void k_neighbor_dump(int group, const DevPairGroup* __restrict__ groups,
                                const PairRec* __restrict__ rec, const int* __restrict__ row_start,
                                const int* __restrict__ row_count, const int* __restrict__ col,
                                int* __restrict__ count, int2* __restrict__ out, long long cap) {
    const DevPairGroup& G = groups[group];
    const int row = G.member_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= G.member_begin + G.member_count) return;
    const PairRec ri = rec[row];
    const int beg = row_start[row], cnt = row_count[row];
    for (int k = 0; k < cnt; ++k) {
        if (col[beg + k] < 0) continue;
        const PairRec rj = rec[col[beg + k]];
        if (!(ri.bead < rj.bead)) continue;
        double dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
        if (G.periodic) {
            dx = min_image(dx, G.box[0], G.inv_box[0]);
            dy = min_image(dy, G.box[1], G.inv_box[1]);
            dz = min_image(dz, G.box[2], G.inv_box[2]);
        }
        if (!(dist2(dx, dy, dz) < G.rc2)) continue;
        const int slot = atomicAdd(count, 1);
        if (slot < cap) out[slot] = make_int2(ri.bead, rj.bead);
    }
}

// Maxwell-Boltzmann velocities for the beads of one system (Philox stream 1).
__global__ void k_velocity_noise(int nb, int sys, const int* __restrict__ sys_of,
                                 const int* __restrict__ local_id, const double* __restrict__ inv_m,
                                 double kT, uint64_t seed, double* __restrict__ vx,
                                 double* __restrict__ vy, double* __restrict__ vz) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || sys_of[b] != sys) return;
    double g[3];
    gauss3(seed, 0, static_cast<uint32_t>(local_id[b]), 1u, g);
    const double s = sqrt(kT * inv_m[b]);
    vx[b] = s * g[0];
    vy[b] = s * g[1];
    vz[b] = s * g[2];
}

// Capped steepest-descent move: x += min(alpha, dmax/|F|) F for mobile beads;
// clears the force for the next evaluation.

// ------------------------------------------------------------ launchers
static inline int nblk(long long n, int b) { return static_cast<int>((n + b - 1) / b); }

void launch_zero_forces(const DevPtrs& d, const ForceBuf& f, hipStream_t st) {
    if (d.nb == 0) return;
    k_fill_i64<<<nblk(d.nb, 256), 256, 0, st>>>(f.fx, 0, d.nb);
    k_fill_i64<<<nblk(d.nb, 256), 256, 0, st>>>(f.fy, 0, d.nb);
    k_fill_i64<<<nblk(d.nb, 256), 256, 0, st>>>(f.fz, 0, d.nb);
}

// Small workgroups: a step touches each bead once, so the launch is one
// latency chain per wave, and more workgroups reach more CUs.
constexpr int kIntegrateBlock = 64;

void launch_integrate(const DevPtrs& d, hipStream_t st, int pin, int pout, int pending, int pack, int check,
                      int prune) {
    PackArgs pk{pack && d.nm > 0, check, d.maxm, d.bead_slot, d.bead_hs2, d.refx, d.refy, d.refz, d.rec,
                d.rec32, d.prx, d.pry, d.prz, d.prune_hm2, d.nm > 0 && d.col_in ? prune : kPruneNone};
    k_integrate<<<nblk(d.nb, kIntegrateBlock), kIntegrateBlock, 0, st>>>(d.nb, d.sys, d.sys_of, d.local_id, d.inv_m, d.px, d.py,
                                                 d.pz, d.vx, d.vy, d.vz, d.f.fx, d.f.fy, d.f.fz, d.gstep,
                                                 pin, pout, pending, d.err, pk, d.sys0, d.nsys == 1 ? 1 : 0);
}

void launch_snapshot(const DevPtrs& d, hipStream_t st, double* snap) {
    const int n3 = 3 * d.nb;
    k_snapshot<<<nblk(std::max(n3, 1), 256), 256, 0, st>>>(n3, d.px, d.vx, d.f.fx, d.gstep, snap, d.flags);
}

void launch_status_out(const DevPtrs& d, hipStream_t st, int* host) {
    k_status_out<<<1, 256, 0, st>>>(d.nsys, d.flags, d.err, host);
}

void launch_half_kick(const DevPtrs& d, hipStream_t st, int pcur) {
    k_half_kick<<<nblk(d.nb, 256), 256, 0, st>>>(d.nb, d.sys, d.sys_of, d.inv_m, d.vx, d.vy, d.vz, d.f.fx,
                                                 d.f.fy, d.f.fz, d.gstep, pcur);
}

void launch_pack(const DevPtrs& d, hipStream_t st) {
    if (d.nm == 0) return;
    k_pack<<<nblk(d.nm, 256), 256, 0, st>>>(d.nm, d.groups, d.member_group, d.member_bead, d.member_type,
                                            d.member_slot, d.px, d.py, d.pz, d.refx, d.refy, d.refz, d.rec,
                                            d.rec32, d.err);
}

static void launch_scan(const DevPtrs& d, hipStream_t st, const int* in, int* out, int n, int* total) {
    const int nb = nblk(n, kScanItems);
    k_scan_blocks<<<nb, kScanThreads, 0, st>>>(in, out, n, d.scan_partial);
    k_scan_partials<<<1, kScanThreads, 0, st>>>(d.scan_partial, nb, total);
    k_scan_add<<<nb, kScanThreads, 0, st>>>(out, n, d.scan_partial);
}

// Exclusive scan of the cell counts in one workgroup (grids up to
// kSmallCells cells): each thread takes a run of consecutive cells, the block
// scans the run sums, then every thread writes its cells' starts.
// start[ncell] receives the total.
constexpr int kSmallCells = 16384;
constexpr int kSmallScanThreads = 1024;

__global__ __launch_bounds__(kSmallScanThreads) void k_scan_cells_small(int ncell, const int* __restrict__ count,
                                                                          int* __restrict__ start) {
    __shared__ int wb[kSmallScanThreads / 64];
    const int per = (ncell + kSmallScanThreads - 1) / kSmallScanThreads;
    const int c0 = min(ncell, static_cast<int>(threadIdx.x) * per), c1 = min(ncell, c0 + per);
    int sb = 0;
    for (int c = c0; c < c1; ++c) sb += count[c];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int ib = sb;   // inclusive scan within the wave
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(ib, o, 64);
        if (lane >= o) ib += t;
    }
    if (lane == 63) wb[w] = ib;
    __syncthreads();
    if (threadIdx.x == 0) {
        int rb = 0;
        for (int k = 0; k < kSmallScanThreads / 64; ++k) {
            const int t = wb[k];
            wb[k] = rb;
            rb += t;
        }
        start[ncell] = rb;
    }
    __syncthreads();
    int rb = ib - sb + wb[w];
    for (int c = c0; c < c1; ++c) {
        start[c] = rb;
        rb += count[c];
    }
}

void launch_rebuild(const DevPtrs& d, hipStream_t st) {
    if (d.nm == 0) return;
    if (d.open_grids > 0)
        k_open_grid<<<d.ngroups, kOpenGridThreads, 0, st>>>(d.groups, d.member_bead, d.px, d.py, d.pz);
    // cell_count is zero here: the previous build's k_place cleared it
    k_cell_assign<<<nblk(d.nm, 256), 256, 0, st>>>(d.nm, d.groups, d.member_group, d.member_bead, d.px, d.py,
                                                   d.pz, d.member_cell, d.member_rank, d.cell_count,
                                                   d.tile_nitems);
    if (d.ncell <= kSmallCells)
        k_scan_cells_small<<<1, kSmallScanThreads, 0, st>>>(d.ncell, d.cell_count, d.cell_start);
    else
        launch_scan(d, st, d.cell_count, d.cell_start, d.ncell, d.cell_start + d.ncell);
    k_place<<<nblk(std::max(d.nm, 1), 256), 256, 0, st>>>(d.nm, d.groups, d.member_group, d.member_bead,
                                                          d.member_type, d.member_cell, d.member_rank, d.cell_start,
                                                          d.px, d.py, d.pz, d.member_slot, d.slot_member,
                                                          d.member_kpos, d.bead_slot, d.rec, d.rec32, d.refx, d.refy,
                                                          d.refz, d.cell_count, d.ncell);
    // work items appended with atomics: the order interleaves the groups, which
    // measured faster than cell order
    k_tile_items<<<nblk(d.ncell, 256), 256, 0, st>>>(d.ncell, d.cell_start, d.tile_items, d.tile_nitems,
                                                     d.tile_chunk);
    k_nl_tiles<<<d.tile_grid, 256, 0, st>>>(d.tile_items, d.tile_nitems, d.groups, d.cell_group, d.cell_start,
                                            d.rec, d.slot_member, d.excl_start, d.excl_count, d.excl,
                                            d.row_start, d.row_count, d.group_need, d.col, d.flags,
                                            d.tile_chunk);
}

void launch_bonded(const DevPtrs& d, hipStream_t st, bool energy, const ForceBuf& f, uint32_t groups) {
    if (d.nt == 0) return;
    ForceOut o{f.fx, f.fy, f.fz};
    if (energy)
        k_bonded<true><<<nblk(d.nt, 256), 256, 0, st>>>(d.nt, d.terms, d.sys, d.sys_of, d.px, d.py, d.pz,
                                                        d.fla, d.fld, o, d.energy, groups, d.err);
    else
        k_bonded<false><<<nblk(d.nt, 256), 256, 0, st>>>(d.nt, d.terms, d.sys, d.sys_of, d.px, d.py, d.pz,
                                                         d.fla, d.fld, o, d.energy, groups, d.err);
}

template <bool E>
static void pair_dispatch(const DevPtrs& d, hipStream_t st, const ForceBuf& f, uint32_t groups,
                          const int* count, const int* col, unsigned long long* stamp) {
    const size_t lds = E ? 0 : static_cast<size_t>(d.max_types) * kTypeStride * sizeof(float);
    const BondedArgs bd{d.nt, d.terms, d.sys, d.sys_of, d.px, d.py, d.pz, d.fla, d.fld};
    const dim3 grid(std::min(d.n_pair_blocks, d.pair_grid)), block(kPairBlock);
#define OCM_PAIR(T, U)                                                                                    \
    if (d.xcd_sliced)                                                                                     \
        k_pair<E, T, true, U><<<grid, block, lds, st>>>(d.groups, d.pair_blocks, d.rec, d.row_start, count, \
                                                        col, d.types_f, d.types_d, f, d.energy, groups,   \
                                                        d.err, d.n_pair_blocks, d.nm, stamp, bd);         \
    else                                                                                                  \
        k_pair<E, T, false, U><<<grid, block, lds, st>>>(d.groups, d.pair_blocks, d.rec, d.row_start, count, \
                                                         col, d.types_f, d.types_d, f, d.energy, groups,  \
                                                         d.err, d.n_pair_blocks, d.nm, stamp, bd)
    // the energy pass has its own loop and always runs the 4-entry instance
    if (!E && d.pair_u == 2) {
        switch (d.team) {
        case 8: OCM_PAIR(8, 2); break;
        case 32: OCM_PAIR(32, 2); break;
        default: OCM_PAIR(16, 2); break;
        }
    } else {
        switch (d.team) {
        case 4: OCM_PAIR(4, 4); break;
        case 8: OCM_PAIR(8, 4); break;
        case 32: OCM_PAIR(32, 4); break;
        case 64: OCM_PAIR(64, 4); break;
        default: OCM_PAIR(16, 4); break;
        }
    }
#undef OCM_PAIR
}

void launch_prune(const DevPtrs& d, hipStream_t st) {
    if (d.nm == 0 || !d.col_in) return;
    const dim3 block(kPairBlock);
#define OCM_PRUNE(T)                                                                                           \
    k_prune<T><<<dim3(static_cast<unsigned>((static_cast<long long>(d.nm) * T + kPairBlock - 1) / kPairBlock)), \
                 block, 0, st>>>(d.nm, d.groups, d.member_group, d.rec, d.rec32, d.row_start, d.row_count,       \
                                 d.col, d.row_count_in, d.col_in)
    switch (d.prune_team) {
    case 4: OCM_PRUNE(4); break;
    case 8: OCM_PRUNE(8); break;
    case 32: OCM_PRUNE(32); break;
    case 64: OCM_PRUNE(64); break;
    default: OCM_PRUNE(16); break;
    }
#undef OCM_PRUNE
}

void launch_pair(const DevPtrs& d, hipStream_t st, bool energy, const ForceBuf& f, uint32_t groups,
                 int list, unsigned long long* stamp) {
    if (d.n_pair_blocks == 0) return;
    // without an inner list, or for a subset of groups or the energy pass, the
    // outer list is read and nothing is written
    if (!d.col_in || energy || (list == kListPrune && groups != 0xFFFFFFFFu)) list = kListOuter;
    if (energy) {
        pair_dispatch<true>(d, st, f, groups, d.row_count, d.col, stamp);
        return;
    }
    if (list == kListPrune) launch_prune(d, st);   // then the force pass reads the fresh inner list
    if (list == kListOuter) pair_dispatch<false>(d, st, f, groups, d.row_count, d.col, stamp);
    else pair_dispatch<false>(d, st, f, groups, d.row_count_in, d.col_in, stamp);
}

int pair_grid_size(const DevPtrs& d) { return std::min(d.n_pair_blocks, d.pair_grid); }

size_t resident_lds_bytes(int nbs, int term_cap, int list_cap, int ex_cap) {
    return res_lds_bytes(nbs, term_cap, list_cap, ex_cap);
}

hipError_t resident_prepare(int threads, size_t lds) {
    const void* f = threads == 256 ? reinterpret_cast<const void*>(&k_resident<256>)
                  : threads == 512 ? reinterpret_cast<const void*>(&k_resident<512>)
                                   : reinterpret_cast<const void*>(&k_resident<1024>);
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
}

void launch_resident(const DevPtrs& d, hipStream_t st, int nsteps) {
    if (d.res_n == 0 || nsteps <= 0) return;
    const ResArgs a{d.res_tab, d.res_grp, d.res_terms, d.res_rows, d.sys, d.groups, d.member_bead, d.member_type,
                    d.local_id, d.inv_m, d.res_hs2, d.px, d.py, d.pz,
                    d.vx, d.vy, d.vz, d.f.fx, d.f.fy, d.f.fz, d.res_list, d.fla, d.fld, d.gstep, d.err, nsteps,
                    d.res_nbs, d.res_term_cap, d.res_list_cap, d.res_excl, d.res_ex_cap};
    const size_t lds = res_lds_bytes(d.res_nbs, d.res_term_cap, d.res_list_cap, d.res_ex_cap);
    switch (d.res_threads) {
    case 256: k_resident<256><<<d.res_n, 256, lds, st>>>(a); break;
    case 512: k_resident<512><<<d.res_n, 512, lds, st>>>(a); break;
    default: k_resident<1024><<<d.res_n, 1024, lds, st>>>(a); break;
    }
    k_gstep_advance<<<1, 64, 0, st>>>(d.gstep, nsteps);
}

void launch_reduce_stamps(hipStream_t st, const unsigned long long* stamp, int nlaunch, int nblk,
                          unsigned long long* out) {
    if (nlaunch > 0) k_reduce_stamps<<<nlaunch, 256, 0, st>>>(stamp, nblk, out);
}

void launch_zero_energy(const DevPtrs& d, hipStream_t st) {
    k_fill_i64<<<nblk(d.nsys * 32, 256), 256, 0, st>>>(d.energy, 0, d.nsys * 32);
    k_fill_i64<<<nblk(d.nsys, 256), 256, 0, st>>>(d.ke, 0, d.nsys);
}

void launch_kinetic(const DevPtrs& d, hipStream_t st) {
    k_kinetic<<<nblk(d.nb, 256), 256, 0, st>>>(d.nb, d.sys_of, d.mass, d.vx, d.vy, d.vz, d.ke, d.err);
}

void launch_neighbor_dump(const DevPtrs& d, hipStream_t st, int group, long long cap, int members,
                          int* count, int2* out) {
    k_fill_i32<<<1, 64, 0, st>>>(count, 0, 1);
    if (members == 0) return;
    k_neighbor_dump<<<nblk(members, 128), 128, 0, st>>>(group, d.groups, d.rec, d.row_start, d.row_count,
                                                        d.col, count, out, cap);
}

void launch_fill_i32(hipStream_t st, int* p, int v, int n) {
    if (n > 0) k_fill_i32<<<nblk(n, 256), 256, 0, st>>>(p, v, n);
}

// ------------------------------------------- pair dump without allocation
// Fixed scratch of the no-allocation pair dump: the count word in the first
// 256 bytes, then up to kNoAllocMaxPairs pairs.  Every GPU the library runs
// on holds its own copy until the process ends.  In host code `scratch` names
// a host-side stand-in, not a device address: kernels receive the address
// hipGetSymbolAddress returns on the device, never the array's name.
static __device__ __align__(256) int scratch[kNoAllocHeaderInts + 2 * kNoAllocMaxPairs];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, int** count, int2** pairs) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // once per device, with that device current
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<int*>(p);
    }
    *count = g_scratch_ptr[device];
    *pairs = reinterpret_cast<int2*>(g_scratch_ptr[device] + kNoAllocHeaderInts);
    return hipSuccess;
}

void launch_velocity_noise(const DevPtrs& d, hipStream_t st, int sys, double kT, uint64_t seed) {
    k_velocity_noise<<<nblk(d.nb, 256), 256, 0, st>>>(d.nb, sys, d.sys_of, d.local_id, d.inv_m, kT, seed,
                                                      d.vx, d.vy, d.vz);
}


}  // namespace ocm
