// am_solve.hpp - argument blocks and launchers of the solve-side kernels.
#pragma once

#include "am_internal.hpp"

namespace am {

// Inputs of the record set-up (b, D, mu in sweep order) and of the unpack.
struct SetupArgs {
    BodiesView B;
    ContactsView C;
    const int* n_contacts;
    const int* order;    // sweep position -> contact index
    double* rec;         // kRecDoubles per sweep position
    int* meta;           // int4 per sweep position: solve body a, solve body b, clamp, 0
    const int* pos_a;    // first solve body of the group of each sweep position
    int* sstate;         // contact state per sweep position
    double dt;
    double feedback;
    double compliance;
    double rest_val;
    double fric_val;
    int rest_ovr;
    int fric_ovr;
    int post;
    int in_collection;
    const double* lam_alt;  // unpack: odd-sweep lambda of the dataflow schedule, stride 4 (null: coloured)
    const int* iters_used;  // unpack: sweeps used (with lam_alt)
    const int* scene_used = nullptr;  // unpack with lam_alt: sweeps used per scene (null: iters_used for all)
};

// Inputs of the persistent coloured PGS kernel.
struct PgsArgs {
    const int* colour_start;  // kMaxColours + 2 entries over sorted group positions
    const int* group_desc;    // int4 per sorted group: first, end sweep position, solve bodies A, B
    const int* meta;
    double* rec;
    int* sstate;
    double* dv;
    const double* minv;
    const double* jinv;
    const int* bflags;
    int iterations;
    int check_tolerance;      // main solve: early exit and changes
    int warm_start;           // apply confidentWarmStart first
    double tolerance;
    double omega;
    double sliding;
    double compliance;
    unsigned long long* iter_max;  // per sweep, bit pattern of max |dl|
    int* iters_used;
    unsigned* barrier;             // grid-barrier counter (zeroed before launch)
    int* error;                    // set if a barrier spin limit is hit
    // Independent scenes that stop one by one (PGS.java:190 per scene): with
    // per_scene set, a group whose scene has converged skips the remaining
    // sweeps.  Arrays are indexed by scene s; scene_max[s * iterations + t]
    // is sweep t+1's max |dl| over the scene's groups, scene_done[s *
    // (iterations + 1) + p] the scene's groups that finished pass p.
    int per_scene = 0;
    int n_scenes = 1;
    const int* g_scene = nullptr;              // scene of each sorted group
    unsigned long long* scene_max = nullptr;
    // per scene and sweep (dataflow schedule): groups done in the high 32
    // bits, those whose max |dl| was not under the tolerance in the low 32
    unsigned long long* scene_done = nullptr;
    const unsigned* scene_groups = nullptr;    // groups per scene
    int* scene_used = nullptr;                 // sweeps each scene used (written after the solve)
};

// Inputs of the dataflow PGS kernel (the same sweeps without grid barriers).
// A group's updated deltaV of each dynamic body goes to its own slot
// (kSlotDoubles doubles per incidence 2 q + side, one set per pass parity);
// the body's next update polls that slot until it is no longer empty, takes
// the value and empties the slot again.
constexpr int kSlotDoubles = 8;
constexpr unsigned long long kSlotEmpty = 0xFFF5A5A5A5A5A5A5ull;  // a signalling NaN: no arithmetic result
constexpr int kPassWords = 8;  // the pass words spread the waves' adds over this many addresses

struct FlowArgs {
    PgsArgs P;
    const int* flow;       // int4 per sorted group: rank_a, deg_a, rank_b, deg_b
    const int* pred;       // int2 per sorted group: incidence of the previous update of A, of B
    double* slot;          // 2 parities x 2 n_groups incidences x kSlotDoubles (all empty at launch)
    int n_groups;
    double* dv_alt;        // deltaV of odd passes (a copy of dv at launch)
    double* lam_alt;       // lambda of odd sweeps, 4 per sweep position (3 used)
    // per pass, kPassWords words (zeroed before launch): waves that have
    // finished it in the high 32 bits, those whose max |dl| was not under the
    // tolerance in the low; wave w adds to word w % kPassWords
    unsigned long long* pass_done;
    unsigned n_waves;
};

void fill_int(hipStream_t s, int* p, int n, int v);
void fill_double(hipStream_t s, double* p, int n, double v);
void fill_u64(hipStream_t s, void* p, size_t n, unsigned long long v);
void launch_groups(hipStream_t s, int cap, const int* n_contacts, const int* cb1, const int* cb2, const int* slot,
                   const int* parent, int* flags, int* scanned, int* group_start, int* group_a, int* group_b,
                   int* d_n_groups, void* scan_temp);
void launch_incidence(hipStream_t s, bool fill, int n_groups, const int* ga, const int* gb, const int* bflags,
                      int* inc_count, const int* inc_start, int* inc_cursor, int* inc_items);
void launch_colour_round(hipStream_t s, int n_groups, const int* ga, const int* gb, const int* bflags,
                         const int* inc_start, const int* inc_count, const int* inc_items, const int* col_in,
                         int* col_out, int* remaining, const int* layer, int* blocker);
void launch_colour_keys(hipStream_t s, int n_groups, const int* colour, const int* group_start, uint64_t* keys);
void launch_colour_layout(hipStream_t s, int n_groups, const uint64_t* sorted, const int* group_start,
                          const int* group_a, const int* group_b, int* colour_start, int* sizes, int* pos_start,
                          int* order, int* group_desc, int* pos_a, int* spos, int* d_total, void* scan_temp);
void launch_setup(hipStream_t s, const SetupArgs& S, int cap);
// scene of every sorted group (from its first contact's body 1) and the group count per scene
void launch_group_scene(hipStream_t s, const int* group_desc, int n_groups, const int* order, const int* cb1,
                        const am_body_desc* desc, int n_scenes, int* g_scene, unsigned* scene_groups);
// sweeps used per scene (first sweep under the tolerance, else all)
void launch_scene_finish(hipStream_t s, const PgsArgs& P);
void launch_unpack(hipStream_t s, const SetupArgs& S, int cap);
int pgs_max_blocks(int block, bool* ok);
hipError_t launch_pgs(hipStream_t s, const PgsArgs& P, int grid, int block, bool plain);
int pgs_flow_max_blocks(bool* ok);
hipError_t launch_pgs_flow(hipStream_t s, const FlowArgs& F, int grid, bool plain, int n_dv, const BodiesView& B);
void launch_flow_ranks(hipStream_t s, int n_ent, int n_groups, const int* bflags, const int* inc_start,
                       const int* inc_count, const int* inc_items, const int* spos, const int* group_desc,
                       int* flow, int* pred);
void launch_map_insert(hipStream_t s, const OldContactsView& O, int* table);
void launch_warm_start(hipStream_t s, BodiesView B, ContactsView C, OldContactsView O, const int* n_groups, int cap,
                       const int* group_start, int post, int* counters);
void launch_redo_warm(hipStream_t s, ContactsView C, const int* n, int cap, int zero);
void launch_clear(hipStream_t s, BodiesView B);
void launch_gravity(hipStream_t s, BodiesView B, const int* targets, int n, double gx, double gy, double gz);
void launch_advance(hipStream_t s, BodiesView B, const int* targets, int n, double dt, bool positions);
// the listed entities' poses advanced by dt with deltaV as the velocities
void launch_advance_post(hipStream_t s, BodiesView B, const int* targets, int n, double dt);
// every entity's pose advanced by dt into xa (3 per entity) and Ra (9 per entity)
void launch_advanced_poses(hipStream_t s, BodiesView B, double dt, double* xa, double* Ra);
void launch_springs(hipStream_t s, BodiesView B, const am_spring* springs, int n, double ks, double ds);
// a force held for the next external-force pass (am_impulse_add)
struct ImpulseRec {
    int body;
    int pad;
    double p[3];
    double f[3];
};
void launch_kinematic(hipStream_t s, BodiesView B, const am_kinematic* k, int n, double t);
void launch_impulses(hipStream_t s, BodiesView B, const ImpulseRec* I, int n);
void launch_viscous(hipStream_t s, BodiesView B, const int* targets, int n, double a1, double a2);
void launch_update_jinv(hipStream_t s, BodiesView B);
void launch_metric(hipStream_t s, BodiesView B, const int* pairs, int n, double* out);

// am_detect.hip
void launch_buckets(hipStream_t s, const int* cell, int n, int* count, const int* start, int* cursor, int* items,
                    bool scatter);
void launch_contacts_set(hipStream_t s, BodiesView B, int n, const int* b1, const int* b2, const double* pw,
                         const double* nw, const double* viol, const int* info, const int* bv1, const int* bv2,
                         ContactsView C, int* overflow);

}  // namespace am
