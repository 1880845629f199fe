// am_merge.hpp - records and launchers of adaptive merging: body-pair
// histories (BodyPairContact), merge/unmerge conditions and the per-step
// physics of collections (RigidCollection).
#pragma once

#include "am_internal.hpp"

namespace am {

// Longest history a body pair keeps (stepAccumMerging / stepAccumUnmerging
// are bounded by this; the reference default is 3).
constexpr int kMaxAccum = 8;

// One external body pair of the current step (BodyPairContact.java:15-49),
// built from the contact list after the main solve.
struct BpcRec {
    int b1, b2;       // physical bodies, in the order of the pair's first contact
    int first, end;   // the pair's contacts in the current list
    int n_active;     // contacts with |lambda0| > 1e-14 (clearBodyPairContacts)
    int not_on_edge;  // active contacts whose state is not ONEDGE
    int breath;       // an active contact moved more than thresholdBreath
    int hist_n;       // motionMetricHist size
    int state_n;      // contactStateHist size
    int states;       // contactStateHist, 2 bits per entry, oldest in the low bits
    int flag;         // merge condition without the same-collection test
    int pad;
    double hist[kMaxAccum];  // motionMetricHist, oldest first
};

// One internal body pair (a BodyPairContact with inCollection set).
struct IbpcRec {
    int b1, b2;
    int first, count;  // its contacts in the internal contact arrays
    int hist_n;        // motionMetricHist size (unmerging keeps a count only)
    int flag;          // unmerge condition
    int coll;          // owning collection entity
    int pad;
};

// Longest motion-metric history of a body for sleeping (the reference's
// stepAccum slider goes up to 200); one more slot holds the value appended
// before the oldest is dropped.
constexpr int kMaxSleepAccum = 200;
constexpr int kSleepHistStride = kMaxSleepAccum + 1;

struct MergeParams {
    int step_accum_merging, step_accum_unmerging;
    int merge_pinned, stable_contact, let_it_breathe;
    int unmerge_friction, unmerge_normal;
    double threshold_merge, threshold_unmerge, threshold_breath;
};

void launch_push_down(hipStream_t s, BodiesView B, const int* members, int n, int integrated_only);
void launch_member_transforms(hipStream_t s, BodiesView B, const double* b2c, const int* members, int n);
void launch_zero_dv(hipStream_t s, BodiesView B, const int* targets, int n);
void launch_clear_targets(hipStream_t s, BodiesView B, const int* targets, int n);
void launch_bpc_build(hipStream_t s, ContactsView C, const int* n_contacts, const int* group_start, int n_groups,
                      unsigned char* active, int* keep, int* keep_scan, int* d_n_keep, void* scan_temp,
                      const BpcRec* prev, int n_prev, const int* prev_hash, int prev_mask, BpcRec* out,
                      double threshold_breath);
void launch_bpc_hash(hipStream_t s, const BpcRec* bpc, int n, int* hash, int mask);
// Poses advanced by dt for the position-level metric; xa == nullptr selects
// the velocity-level metric.
struct PoseLevel {
    const double* xa;
    const double* Ra;
    double dt;
};
void launch_bpc_accumulate(hipStream_t s, BodiesView B, BpcRec* bpc, int n, MergeParams P, PoseLevel L);
void launch_ibpc_accumulate(hipStream_t s, BodiesView B, IbpcRec* ib, int n, ContactsView IC, MergeParams P,
                            PoseLevel L);
// mode 0 copy, 1 fresh copy (lambda_warm, b, D zeroed, not new), 2 lambda_warm
// = lambda; 3 scatter: dst[idx[k]] = src[src_off + k]
void launch_gather_contacts(hipStream_t s, ContactsView dst, int dst_off, ContactsView src, const int* idx,
                            int src_off, int n, int mode);
void launch_wake_count(hipStream_t s, BodiesView B, const int* b1, const int* b2, int n, int* count);
void launch_sleep_external(hipStream_t s, BodiesView B, const int* b1, const int* b2, const double* lam, int n,
                           int* ext);
void launch_sleep_decide(hipStream_t s, BodiesView B, int* flags, const int* slot_entity, int n_slots, const int* ext,
                         double* hist, int* hist_n, int accum, double threshold, int* counts);

}  // namespace am
