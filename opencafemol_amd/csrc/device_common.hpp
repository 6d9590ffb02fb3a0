// device_common.hpp -- data layout shared by the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocm {

// Fixed-point force/energy accumulation: value * 2^32 stored in int64, so every
// sum over contributions is exact and independent of summation order.
constexpr double kFixScale = 4294967296.0;          // 2^32
constexpr double kFixInv = 1.0 / 4294967296.0;
constexpr float kFixLimitF = 1073741824.0f;          // |contribution| must stay < 2^30

// Per-system error bits (DevSys::err).
constexpr int kErrNumeric = 1;   // NaN or fixed-point overflow in a force/energy term
constexpr int kErrStale = 2;     // a bead moved more than skin/2 since the list build
constexpr int kErrPrune = 4;     // a bead moved more than margin/2 since the inner list was pruned
// Batch-wide flag word bits.
constexpr int kFlagListOverflow = 1;

// Pair-group kind bits (bit k: nonbonded kind OCM_EXCLUDED_VOLUME + k).
constexpr int kExv = 1, kDh = 2, kLjr = 4, kLja = 8;
constexpr int kWca = 16, kUwca = 32, kUlja = 64, kIsolf = 128;
constexpr int kPairKinds = 8;

// Bonded term record (all kinds in one array, sorted by kind).
struct BTerm {
    int idx[4];      // bead indices; unused slots -1
    int kind;        // ocm_force_kind
    int meta;        // bits 0-4 energy group, bit 5 periodic, bits 8.. table id
    double p[3];     // per-term parameters, OpenMM units
};

// Flexible-local-angle table: knot values and natural-spline second derivatives.
struct FlaTable {
    double y[10];
    double y2[10];
    double tmin, tmax, h, emin;
};
// Flexible-local-dihedral table: c, a1, b1, a2, b2, a3, b3, Emin.
struct FldTable {
    double c[8];
};

struct DevSys {
    double box[3];
    double inv_box[3];
    int periodic;
    int bead_begin;
    int bead_count;
    int pad0;
    double half_dt;
    double c1;            // exp(-gamma dt)
    double noise_kT;      // kT (1 - c1^2)
    uint64_t seed;
    int64_t step_origin;  // global step at which this system's step count was 0
};

// Pair-record of one member bead in slot (cell-sorted) order.
struct alignas(16) PairRec {
    double x, y, z;
    int type;   // bits 0-15: index into the group's type table; bits 16-20: category
    int bead;   // global bead index
};

struct DevPairGroup {
    int sys;
    int kinds;
    int periodic;
    int member_begin;
    int member_count;
    int type_begin;
    int n_types;
    int cell_begin;
    int ncell[3];
    int reach[3];          // 1 when the stencil spans neighbour cells along an axis
    int egroup[kPairKinds];   // energy group of the member force of each kind
    int n_cat;
    int stride;            // list entries reserved per row (multiple of 4)
    int lj_shared;         // LJ repulsive and attractive members share sigma for every type
    int open_grid;         // open boundary with a cell grid over the members' bounding box,
                           // laid out again by k_open_grid at every build
    int grid_cap;          // cells reserved for that grid
    int pad1;
    double rc2;            // cutoff^2 (nm^2)
    double rl2;            // (cutoff + skin)^2
    double half_skin2;     // (skin/2)^2
    double rin2;           // (cutoff + margin)^2: radius of the pruned inner list
    double rin2_wide;      // periodic groups: the FP32 prune's threshold, (cutoff + margin + 2 delta)^2
                           // with delta a bound on its per-axis coordinate error
    double rc;
    double box[3], inv_box[3];
    double cell_scale[3];  // cells per nm (ncell / box when periodic)
    double cell_lo[3];     // open grid: the lower corner of cell 0
    // potential constants (FP64; the force pass rounds them to FP32)
    double exv_eps, exv_corr;
    double dh_K, dh_inv_lambda, dh_corr;
    double ljr_rsw, lja_rsw, exv_rsw, dh_rsw;   // switching distances (<= 0: none)
    double lja_inv_sw;     // 1 / (cutoff - lja_rsw)
    double u_sigma, u_eps; // the uniform WCA / LJ-attractive kinds' sigma and epsilon
    uint32_t allow[32];    // category allow masks
};

// Resident stepping (small systems, one workgroup per system, whole chunks
// on chip): per system, its bonded terms (bead indices local to the system)
// and its pair groups, whose FP32 type rows are staged together in LDS.
constexpr int kResMaxBeads = 1024;      // beads per system (local indices fit 10 bits)
constexpr int kResMaxGroups = 8;        // pair groups per system
constexpr int kResMaxTypeRows = 64;     // staged type rows per system (6 bits)
// The in-kernel build tests pairs on FP32 coordinates: periodic systems'
// wrapped into boxes of at most kResWide nm, open systems' relative to bead 0
// (a build that finds one beyond kResWide tests that build's pairs in FP64).
constexpr double kResWide = 1000.0;
constexpr int kResMaxSeg = 9;   // bonded kinds 1-9: one segment each
struct ResSys {
    int sys;          // DevSys index
    int nb;           // beads
    int b0;           // first global bead
    int t0, nt;       // bonded terms [t0, t0 + nt) of the resident term array
    int g0, ng;       // resident pair groups [g0, g0 + ng)
    int row0, nrow;   // staged type rows [row0, row0 + nrow) of the resident row array
    int ex0, exn;     // the groups' exclusion tables: words [ex0, ex0 + exn) of the resident exclusion array
    // the terms in runs of one kind: run q is terms [seg_t[q], seg_t[q] + seg_n[q])
    // of the system's own (kind-ordered) list
    int nseg;
    int seg_t[kResMaxSeg], seg_n[kResMaxSeg];
};
// A group's exclusions in member order (members follow bead order), for the
// in-kernel build, which tests member pairs (i, j > i): per member a mask
// whose bit k says member i + 1 + k is excluded (k < 32), then nm + 1 starts
// and the partners further away (member indices, sorted) -- chains exclude
// their near neighbours, so most exclusions are mask bits.
constexpr int kResNearExcl = 32;
struct ResGroup {
    int gid;          // pair group (DevPairGroup index)
    int m0, nm;       // members [m0, m0 + nm)
    int trow;         // the group's first type row within its system's staged rows
    long long l0;     // list region: nm (nm - 1) / 2 entries from l0
    double rl2;       // (cutoff + skin)^2 with the resident skin
    float rl2w;       // (cutoff + skin + FP32 margin)^2: the in-kernel build's threshold
    int exo;          // its exclusion table within the system's: mask [exo, exo + nm),
                      // starts [exo + nm, exo + 2 nm + 1), far partners after them
};

// Per-type parameters: {exv_radius, dh_charge, 0, 0, ljr_eps, ljr_sigma, lja_eps, lja_sigma};
// the four LJ values sit 16-byte aligned for one vector read.  A group of
// WCA or iSoLF (which holds no other kind) uses {sigma, epsilon, omega}.
constexpr int kTypeStride = 8;
constexpr int kTyExv = 0, kTyQ = 1, kTyLjrEps = 4, kTyLjrSig = 5, kTyLjaEps = 6, kTyLjaSig = 7;
constexpr int kTySig = 0, kTyEps = 1, kTyOmega = 2;

// ------------------------------------------------------------------ Philox
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c[3] ^ k1;
        c[1] = static_cast<uint32_t>(p1);
        c[3] = static_cast<uint32_t>(p0);
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Three N(0,1) samples for (seed, step, bead, stream) via Box-Muller in FP64.
__host__ __device__ inline void gauss3(uint64_t seed, uint64_t step, uint32_t bead,
                                       uint32_t stream, double g[3]) {
    uint32_t c[4] = {static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32), bead, stream};
    philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    const double s = 2.3283064365386963e-10;  // 2^-32
    const double u0 = (c[0] + 0.5) * s, u1 = (c[1] + 0.5) * s;
    const double u2 = (c[2] + 0.5) * s, u3 = (c[3] + 0.5) * s;
    const double twopi = 6.283185307179586476925286766559;
    const double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
    g[0] = r0 * cos(twopi * u1);
    g[1] = r0 * sin(twopi * u1);
    g[2] = r1 * cos(twopi * u3);
}

}  // namespace ocm
