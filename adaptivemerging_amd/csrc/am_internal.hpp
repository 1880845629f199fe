// am_internal.hpp - device data views and launch entry points shared by the
// translation units of libam.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/am.h"
#include "am_math.hpp"

namespace am {

// Body arrays, one slot per entity (physical bodies, then collections).
struct BodiesView {
    int n;              // entities
    int n_phys;         // physical bodies (have geometry)
    const am_body_desc* desc;
    double* x;          // 3 per entity
    double* R;          // 9
    double* v;          // 3
    double* w;          // 3
    double* f;          // 3
    double* t;          // 3
    double* jinv;       // 9, world-aligned
    double* dv;         // 6 (deltaV: v then w)
    const double* minv; // 1
    const int* flags;   // am_body_flags
    const int* parent;  // collection entity or -1
    const int* slot;    // position of the body's top-level entity in the broad-phase list
    const int* sub;     // index inside its collection (0 otherwise)
    const am_body_desc* sub_desc;  // sub-bodies of composite bodies
    const double* sub_b2c;         // their transforms into the composite's frame: R (9) then t (3)
};

struct TreeView {
    const am_sphere_node* nodes;
    double* cw;         // world centre per node, 3 per node
};

// Contact arrays (structure of arrays), index = position in the contact list.
struct ContactsView {
    int cap;
    int* b1;
    int* b2;
    int* bv1;
    int* bv2;
    int* info;
    double* pB1;
    double* nB1;
    double* t1B1;
    double* t2B1;
    double* viol;
    double* prev_viol;
    double* J;          // 36 per contact
    double* lam;
    double* lamw;
    double* b;
    double* D;
    int* state;
    int* is_new;
};

// Velocity of a world point moving with body b (RigidBody.getSpatialVelocity).
AMD_INLINE d3 spatial_velocity(const BodiesView& B, int b, d3 pW) {
    d3 r = sub(pW, ld3(&B.x[b * 3]));
    return add(cross(ld3(&B.w[b * 3]), r), ld3(&B.v[b * 3]));
}
// Largest speed of body a's bounding-box corners
// (MotionMetricProcessor.getLargestVelocityNorm(body), :62-73).
AMD_INLINE double metric_one(const BodiesView& B, int a) {
    m33 R = ldm(&B.R[a * 9]);
    d3 x = ld3(&B.x[a * 3]);
    double largest = 0;
    for (int q = 0; q < 8; q++) {
        d3 pW = xform(R, x, ld3(&B.desc[a].bbox[q * 3]));
        largest = jmax(length(spatial_velocity(B, a, pW)), largest);
    }
    return largest;
}
// Largest relative speed over both bodies' bounding-box corners (:39-57).
AMD_INLINE double metric_two(const BodiesView& B, int a, int b) {
    double largest = 0;
    for (int s = 0; s < 2; s++) {
        int body = s == 0 ? a : b;
        m33 R = ldm(&B.R[body * 9]);
        d3 x = ld3(&B.x[body * 3]);
        for (int q = 0; q < 8; q++) {
            d3 pW = xform(R, x, ld3(&B.desc[body].bbox[q * 3]));
            d3 v1 = spatial_velocity(B, a, pW);
            d3 v2 = spatial_velocity(B, b, pW);
            largest = jmax(length(sub(v1, v2)), largest);
        }
    }
    return largest;
}
// The same measure at the position level (:75-116): each corner of one body
// goes into the other body's frame, out again through the other body's pose
// advanced by dt, back through the first body's advanced pose, and its
// displacement over dt is the speed.  xa/Ra hold every entity's pose after
// RigidBody.advancePositions(dt) on a copy.
AMD_INLINE double metric_two_pos(const BodiesView& B, const double* xa, const double* Ra, int a, int b, double dt) {
    double largest = 0;
    for (int s = 0; s < 2; s++) {
        const int self = s == 0 ? a : b, other = s == 0 ? b : a;
        const m33 R1 = ldm(&B.R[self * 9]), R2 = ldm(&B.R[other * 9]);
        const m33 R1a = ldm(&Ra[self * 9]), R2a = ldm(&Ra[other * 9]);
        const d3 x1 = ld3(&B.x[self * 3]), x2 = ld3(&B.x[other * 3]);
        const d3 x1a = ld3(&xa[self * 3]), x2a = ld3(&xa[other * 3]);
        for (int q = 0; q < 8; q++) {
            const d3 p = ld3(&B.desc[self].bbox[q * 3]);
            const d3 pB = inv_xform(R2, x2, xform(R1, x1, p));
            const d3 back = inv_xform(R1a, x1a, xform(R2a, x2a, pB));
            const double s_dt = 1. / dt;
            d3 v = sub(p, back);
            v = d3{v.x * s_dt, v.y * s_dt, v.z * s_dt};
            largest = jmax(length(v), largest);
        }
    }
    return largest;
}

// Previous step's contacts needed by the warm start.
struct OldContactsView {
    int n;
    const int* b1;
    const int* b2;
    const int* bv1;
    const int* bv2;
    const int* info;
    const double* pB1;
    double* lam;        // zeroed when taken ("take and zero")
    const double* viol;
    const double* prev_viol;
    const int* table;   // open-addressing table of old indices, -1 empty
    int table_mask;
};

// Pair keys: slot_lo:20 | slot_hi:20 | sub_lo:12 | sub_hi:12.
constexpr int kSlotBits = 20;
constexpr int kSubBits = 12;
constexpr int kMaxSlots = 1 << kSlotBits;
// broad phase: bodies whose bounding radius exceeds kLargeRatio times the
// median stay out of the grid (at most kMaxLarge of them; beyond that the grid
// takes every body and its cells fit the largest)
constexpr double kLargeRatio = 4.0;
constexpr int kMaxLarge = 256;
constexpr int kMaxSub = 1 << kSubBits;
constexpr int kTreeStack = 128;
constexpr int kMaxColours = 64;     // colours 0..63 run in parallel; 64 is the sequential tail

// Solve record: the per-contact data the PGS sweep reads, packed in sweep order.
// Per row (n, t1, t2), 15 doubles: ja.w, jb.v, jb.w (ja.v = -jb.v exactly),
// then jinv_a ja.w and jinv_b jb.w (constant during a solve, so computed once
// in setup with the same operations the update would do).  Then b[3], D[3],
// mu, lambda[3] (read and written in place), the solve bodies a | b << 32 as
// raw bits, the clamp flag, then D + C [3] and the refined reciprocals of
// D + C [3] (the dataflow sweep's division), one pad: 64 doubles (512 B,
// 16-byte aligned).
constexpr int kRecDoubles = 64;
constexpr int kRecRow = 15;
constexpr int kRecB = 45;
constexpr int kRecD = 48;
constexpr int kRecMu = 51;
constexpr int kRecLambda = 52;
constexpr int kRecAB = 55;
constexpr int kRecClamp = 56;
constexpr int kRecDen = 57;
constexpr int kRecY = 60;

// ---- launchers (am_detect.hip) ----
struct PairArgs {
    BodiesView B;
    TreeView T;
    const int* grid_order;   // gridded index -> physical body
    int n_grid;
    const double* sphere;    // 4 per gridded body: centre, radius
    const int* cell;         // 4 per gridded body: cx, cy, cz, bucket
    const int* bucket_start;
    const int* bucket_count;
    const int* bucket_items;
    int mask;
    const int* planes;       // grouped by scene, each group in body order
    const int* plane_start;  // per scene: its planes are planes[plane_start[s] .. plane_start[s + 1])
    const int* top_pinned;   // per slot
    const int* top_sleeping; // per slot
    const int* large;        // gridded indices of the bodies too large for the grid (not in its buckets)
    int n_large;
};
void launch_tree_world_centres(hipStream_t s, BodiesView B, TreeView T, const int* node_body, const int* node_sub,
                               int n_nodes);
void launch_bounds(hipStream_t s, BodiesView B, TreeView T, const int* order, int n, double* sphere,
                   double cell, int* cell_key, int hash_mask, const int* is_large);
int pair_tasks(int n_grid);  // threads of the pair passes: one per (grid body, task)
void launch_pairs_impl(hipStream_t s, bool write, const PairArgs& A, int* counts, const int* offsets, uint64_t* keys,
                       int cap, int* overflow);
void launch_narrow(hipStream_t s, bool write, BodiesView B, TreeView T, const uint64_t* keys,
                   const int* n_pairs, int pair_cap, const int* slot_entity, const int* member_first,
                   const int* members, int* counts, const int* offsets, ContactsView C, int* overflow);

// ---- launchers (am_scan.hip) ----
size_t scan_temp_bytes(int n);
// Exclusive scan of n ints; writes the total to *d_total (device).
void exclusive_scan(hipStream_t s, const int* in, int* out, int n, int* d_total, void* temp);
size_t sort_temp_bytes(int n);
void sort_u64(hipStream_t s, uint64_t* keys_in, uint64_t* keys_out, int n, void* temp, size_t temp_bytes,
              int begin_bit, int end_bit);

}  // namespace am
