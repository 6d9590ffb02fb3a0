// kernels.hpp -- device buffer table and launch entry points of the step stages.
#pragma once

#include "device_common.hpp"

#include <hip/hip_runtime.h>

namespace ocm {

struct ForceBuf {
    long long* fx;
    long long* fy;
    long long* fz;
};

// Every device array of a packed batch.  Sizes: nb beads, nsys systems,
// nt bonded terms, nm pair-group members (= slots), ncell cells.
struct DevPtrs {
    int nb = 0, nsys = 0, nt = 0, nm = 0, ncell = 0, ngroups = 0;
    int open_grids = 0;            // pair groups with an open-boundary cell grid
    DevSys sys0{};                 // host copy of system 0's record (kernel argument when nsys == 1)
    DevSys* sys = nullptr;
    int* sys_of = nullptr;
    int* local_id = nullptr;
    double* inv_m = nullptr;
    double* mass = nullptr;
    double *px = nullptr, *py = nullptr, *pz = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    ForceBuf f{};          // step forces
    ForceBuf f2{};         // scratch forces for group-restricted getState
    long long* gstep = nullptr;   // two slots, see k_integrate
    int* err = nullptr;           // nsys error words
    int* flags = nullptr;         // batch-wide flag word
    int* group_need = nullptr;    // per pair group: longest row seen that did not fit
    BTerm* terms = nullptr;
    FlaTable* fla = nullptr;
    FldTable* fld = nullptr;
    DevPairGroup* groups = nullptr;
    int *member_group = nullptr, *member_bead = nullptr, *member_type = nullptr, *member_cat = nullptr;
    int *member_slot = nullptr, *member_cell = nullptr, *member_rank = nullptr;
    // bead -> its pair-group members, `maxm` slots per bead: the current record
    // slot of each (-1 past the bead's last membership) and the smallest
    // (skin/2)^2 of its groups
    int maxm = 1;
    int* bead_slot = nullptr;
    double* bead_hs2 = nullptr;
    int* member_kpos = nullptr;   // member -> its position in the bead tables
    int *excl_start = nullptr, *excl_count = nullptr, *excl = nullptr;
    double *refx = nullptr, *refy = nullptr, *refz = nullptr;   // positions at the last build, by bead
    PairRec* rec = nullptr;
    float4* rec32 = nullptr;       // periodic groups: box-wrapped FP32 positions by slot (the prune's input)
    int* slot_member = nullptr;
    int *cell_count = nullptr, *cell_start = nullptr, *cell_group = nullptr;
    int *row_count = nullptr, *row_start = nullptr, *col = nullptr;
    // inner list: the outer rows pruned to cutoff + margin, same row layout
    int *row_count_in = nullptr, *col_in = nullptr;
    double *prx = nullptr, *pry = nullptr, *prz = nullptr;   // positions at the last prune, by bead
    double prune_hm2 = 0.0;        // (margin/2)^2, slightly reduced

    long long list_capacity = 0;
    int2* tile_items = nullptr;    // (cell, chunk) work items of the list build
    int* tile_nitems = nullptr;
    int tile_grid = 2048;          // workgroups of the list-build kernels
    int tile_chunk = 16;           // member rows per list-build work item (16 or 64)
    int pair_grid = 2048;          // workgroups of the pair kernel
    int* scan_partial = nullptr;
    int* scan_total = nullptr;     // [0] cells total, [1] list total
    float* types_f = nullptr;
    double* types_d = nullptr;
    int max_types = 0;
    int2* pair_blocks = nullptr;   // (group, first slot) per pair-kernel block
    int n_pair_blocks = 0;
    int xcd_sliced = 0;            // pair_blocks holds 8 contiguous XCD slices of equal work mix (batches of many systems)
    int team = 16;
    int pair_u = 4;                // force pass: entries per lane per round (4 or 2)
    int prune_team = 32;           // lanes per row of the inner-list prune
    // resident stepping (every system small): one workgroup of res_threads
    // threads per system, res_n systems
    int res_n = 0, res_threads = 0;
    int res_term_cap = 0, res_list_cap = 0;   // bonded terms and list entries held in LDS per system
    int res_nbs = 0;                          // LDS entries per bead array (multiple of 64)
    int res_ex_cap = 0;                       // exclusion-table words per system (staged in LDS)
    int* res_excl = nullptr;
    double* res_hs2 = nullptr;                // per bead: (resident skin / 2)^2
    ResSys* res_tab = nullptr;
    ResGroup* res_grp = nullptr;
    BTerm* res_terms = nullptr;
    float* res_rows = nullptr;
    uint32_t* res_list = nullptr;
    long long* energy = nullptr;   // nsys * 32 fixed point
    long long* ke = nullptr;       // nsys fixed point
    int* dump_count = nullptr;
};

constexpr int kPairBlock = 128;
// Inner-list margin beyond the cutoff (nm), capped by the group's skin.
constexpr double kPruneMargin = 0.1;

void launch_zero_forces(const DevPtrs& d, const ForceBuf& f, hipStream_t st);
// pack != 0: also refresh the pair records (and check list validity when check != 0).
// prune: kPruneNone, kPruneWrite (this step's pair pass prunes: record the
// positions) or kPruneCheck (flag beads that moved more than margin/2 since).
constexpr int kPruneNone = 0, kPruneWrite = 1, kPruneCheck = 2;
void launch_integrate(const DevPtrs& d, hipStream_t st, int pin, int pout, int pending, int pack, int check,
                      int prune = kPruneNone);
void launch_half_kick(const DevPtrs& d, hipStream_t st, int pcur);
// Chunk bracket: snap receives positions, velocities, force words (3*nb each)
// and the two step slots, and the flag word is cleared; host (pinned, 1 + nsys
// words) receives the flag word and the error words, which are then cleared.
void launch_snapshot(const DevPtrs& d, hipStream_t st, double* snap);
void launch_status_out(const DevPtrs& d, hipStream_t st, int* host);
void launch_pack(const DevPtrs& d, hipStream_t st);
void launch_rebuild(const DevPtrs& d, hipStream_t st);
void launch_bonded(const DevPtrs& d, hipStream_t st, bool energy, const ForceBuf& f, uint32_t groups);
// stamp (optional): 2 * pair_grid_size words; workgroup w stores its start
// and end on the device wall clock at stamp[2w], stamp[2w + 1].
// list: kListOuter reads the outer list; kListPrune reads it and writes the
// inner list (full group mask only; otherwise it reads the outer list);
// kListInner reads the inner list.
constexpr int kListOuter = 0, kListPrune = 1, kListInner = 2;
void launch_pair(const DevPtrs& d, hipStream_t st, bool energy, const ForceBuf& f, uint32_t groups,
                 int list = kListOuter, unsigned long long* stamp = nullptr);
// Rewrites the inner list from the outer one at the current positions.
void launch_prune(const DevPtrs& d, hipStream_t st);
int pair_grid_size(const DevPtrs& d);
// Resident stepping: nsteps steps of every system, each in one workgroup, and
// both step slots advanced by nsteps.  Forces at the current positions must be
// in d.f and the velocities synchronized; both hold again afterwards.
void launch_resident(const DevPtrs& d, hipStream_t st, int nsteps);
// LDS of one resident workgroup (bead arrays of nbs entries), and the
// attribute that allows it (gfx950 gives a workgroup up to 160 KB).
size_t resident_lds_bytes(int nbs, int term_cap, int list_cap, int ex_cap);
hipError_t resident_prepare(int threads, size_t lds);
constexpr size_t kResLdsMax = 160 * 1024;
void launch_reduce_stamps(hipStream_t st, const unsigned long long* stamp, int nlaunch, int nblk,
                          unsigned long long* out);
void launch_zero_energy(const DevPtrs& d, hipStream_t st);
void launch_kinetic(const DevPtrs& d, hipStream_t st);
// Interacting pairs (r < cutoff, i < j as global bead ids) of one pair group
// into out (up to cap pairs); *count receives how many there are.
void launch_neighbor_dump(const DevPtrs& d, hipStream_t st, int group, long long cap, int members,
                          int* count, int2* out);
void launch_fill_i32(hipStream_t st, int* p, int v, int n);
// Fixed device scratch of the no-allocation pair dump (a count word, then
// kNoAllocMaxPairs int2 pairs), one copy per GPU for the life of the process.
// The first call on a device, with that device current, looks its address up;
// later calls return the kept address.  Calls that overlap on one GPU share it.
constexpr int kNoAllocMaxPairs = 1 << 21;
constexpr int kNoAllocHeaderInts = 64;
hipError_t noalloc_scratch(int device, int** count, int2** pairs);
void launch_velocity_noise(const DevPtrs& d, hipStream_t st, int sys, double kT, uint64_t seed);

}  // namespace ocm
