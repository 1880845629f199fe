// am_merge.hip - adaptive merging.
//
// Device side: the per-step physics of collections (velocity push-down,
// member transforms), the external body-pair table with its merge histories,
// the unmerge histories of internal pairs, and contact gathers.
// Host side: the graph edits of Merging.merge / Merging.unmerge and the mass
// properties of RigidCollection, done when a merge or unmerge event happens
// (rare), on host mirrors downloaded for that event.
//
// The reference iterates HashSets here (bodyPairContacts, bpcsToUnmerge,
// subbodies); this file uses fixed orders instead: body pairs in contact-list
// order, internal pairs in creation order, connected components in
// depth-first order from the first member.
#include <cmath>
#include <functional>
#include <new>
#include <thread>
#include <unordered_map>

#include "am_ctx.hpp"

namespace am {

// ---------------------------------------------------------------------------
// Collection physics on the device.

// RigidCollection.applyVelocitiesTo (RigidCollection.java:925-937) for every
// listed member: v = v_c + w_c x (x - x_c), w = w_c.  integrated_only limits it
// to members of collections that advanceVelocities visits (not pinned).
__global__ void k_push_down(BodiesView B, const int* members, int n, int integrated_only) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int b = members[k];
    int c = B.parent[b];
    if (integrated_only && (B.flags[c] & (AM_BODY_PINNED | AM_BODY_SLEEPING))) return;
    d3 r = sub(ld3(&B.x[b * 3]), ld3(&B.x[c * 3]));
    d3 wxr = cross(ld3(&B.w[c * 3]), r);
    st3(&B.v[b * 3], add(ld3(&B.v[c * 3]), wxr));
    st3(&B.w[b * 3], ld3(&B.w[c * 3]));
}
void launch_push_down(hipStream_t s, BodiesView B, const int* members, int n, int integrated_only) {
    if (n > 0)
        hipLaunchKernelGGL(k_push_down, dim3((n + 255) / 256), dim3(256), 0, s, B, members, n, integrated_only);
}

// RigidCollection.updateBodiesPositionAndTransformations (:898-909):
// B2W = C2W * B2C, then jinv = R jinv0 R^T, for members of collections that
// advancePositions visits.
__global__ void k_member_transforms(BodiesView B, const double* b2c, const int* members, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int b = members[k];
    int c = B.parent[b];
    if (B.flags[c] & (AM_BODY_PINNED | AM_BODY_SLEEPING)) return;
    m33 Rc = ldm(&B.R[c * 9]);
    m33 Rbc = ldm(&b2c[(size_t)b * 12]);
    d3 t = add(mul(Rc, ld3(&b2c[(size_t)b * 12 + 9])), ld3(&B.x[c * 3]));
    m33 R = mmul(Rc, Rbc);
    st3(&B.x[b * 3], t);
    stm(&B.R[b * 9], R);
    stm(&B.jinv[b * 9], rm0rt(R, ldm(B.desc[b].jinv0)));
}
void launch_member_transforms(hipStream_t s, BodiesView B, const double* b2c, const int* members, int n) {
    if (n > 0) hipLaunchKernelGGL(k_member_transforms, dim3((n + 127) / 128), dim3(128), 0, s, B, b2c, members, n);
}

__global__ void k_zero_dv(BodiesView B, const int* targets, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = targets[k];
#pragma unroll
    for (int a = 0; a < 6; a++) B.dv[i * 6 + a] = 0.0;
}
void launch_zero_dv(hipStream_t s, BodiesView B, const int* targets, int n) {
    if (n > 0) hipLaunchKernelGGL(k_zero_dv, dim3((n + 255) / 256), dim3(256), 0, s, B, targets, n);
}

// RigidBody.clear for the listed entities only.
__global__ void k_clear_targets(BodiesView B, const int* targets, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = targets[k];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        B.f[i * 3 + a] = 0.0;
        B.t[i * 3 + a] = 0.0;
    }
#pragma unroll
    for (int a = 0; a < 6; a++) B.dv[i * 6 + a] = 0.0;
}
void launch_clear_targets(hipStream_t s, BodiesView B, const int* targets, int n) {
    if (n > 0) hipLaunchKernelGGL(k_clear_targets, dim3((n + 255) / 256), dim3(256), 0, s, B, targets, n);
}

// ---------------------------------------------------------------------------
// External body pairs.

AMD_INLINE uint64_t pair_key(int a, int b) {
    uint32_t lo = (uint32_t)min(a, b), hi = (uint32_t)max(a, b);
    return ((uint64_t)lo << 32) | hi;
}
AMD_INLINE uint32_t pair_hash(uint64_t z) {
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    z ^= z >> 33;
    return (uint32_t)z;
}

// clearBodyPairContacts (CollisionProcessor.java:213-226), pass 1: a
// contact is active when |lambda0| > 1e-14; a pair survives with one active
// contact or more.
// Entity records an event changed on the host, one per listed entity:
// x, v, w (3 each), R, jinv (9 each), minv, then the descriptor (60 doubles).
constexpr int kEntRecDoubles = 3 + 3 + 3 + 9 + 9 + 1 + (int)(sizeof(am_body_desc) / 8);
static_assert(sizeof(am_body_desc) % 8 == 0, "descriptor is copied as doubles");

__global__ void k_scatter_entities(double* x, double* v, double* w, double* R, double* jinv, double* minv,
                                   am_body_desc* desc, const int* list, const double* rec, int n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * kEntRecDoubles) return;
    const int k = (int)(i / kEntRecDoubles), q = (int)(i % kEntRecDoubles), e = list[k];
    const double val = rec[i];
    if (q < 3) x[(size_t)e * 3 + q] = val;
    else if (q < 6) v[(size_t)e * 3 + q - 3] = val;
    else if (q < 9) w[(size_t)e * 3 + q - 6] = val;
    else if (q < 18) R[(size_t)e * 9 + q - 9] = val;
    else if (q < 27) jinv[(size_t)e * 9 + q - 18] = val;
    else if (q < 28) minv[e] = val;
    else reinterpret_cast<double*>(desc + e)[q - 28] = val;
}

// Body-to-collection transforms (12 doubles) of the listed bodies.
__global__ void k_scatter_b2c(double* b2c, const int* list, const double* rec, int n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * 12) return;
    b2c[(size_t)list[i / 12] * 12 + i % 12] = rec[i];
}

__global__ void k_bpc_keep(ContactsView C, const int* n_groups, const int* group_start, unsigned char* active,
                           int* keep, int cap_groups) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= cap_groups) return;
    int ng = *n_groups;
    if (g >= ng) {
        keep[g] = 0;
        return;
    }
    int any = 0;
    for (int k = group_start[g]; k < group_start[g + 1]; k++) {
        unsigned char a = fabs(C.lam[k * 3]) > 1e-14 ? 1 : 0;
        active[k] = a;
        any |= a;
    }
    keep[g] = any;
}

// Pass 2: one record per surviving pair, with the histories carried over from
// the previous step's record of the same pair (BodyPairContact.checkExists).
__global__ void k_bpc_write(ContactsView C, const int* n_groups, const int* group_start, const unsigned char* active,
                            const int* keep, const int* keep_scan, const BpcRec* prev, const int* prev_hash,
                            int prev_mask, BpcRec* out, double threshold_breath) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= *n_groups || !keep[g]) return;
    BpcRec r;
    int first = group_start[g], end = group_start[g + 1];
    r.b1 = C.b1[first];
    r.b2 = C.b2[first];
    r.first = first;
    r.end = end;
    r.n_active = 0;
    r.not_on_edge = 0;
    r.breath = 0;
    for (int k = first; k < end; k++) {
        if (!active[k]) continue;
        r.n_active++;
        if (C.state[k] != AM_STATE_ONEDGE) r.not_on_edge++;
        if (fabs(C.prev_viol[k] - C.viol[k]) > threshold_breath) r.breath = 1;
    }
    r.hist_n = 0;
    r.state_n = 0;
    r.states = 0;
    r.flag = 0;
    r.pad = 0;
    for (int q = 0; q < kMaxAccum; q++) r.hist[q] = 0.0;
    if (prev_mask > 0) {
        uint64_t key = pair_key(r.b1, r.b2);
        uint32_t h = pair_hash(key) & (uint32_t)prev_mask;
        for (int probe = 0; probe <= prev_mask; probe++) {
            int idx = prev_hash[h];
            if (idx < 0) break;
            const BpcRec& p = prev[idx];
            if (pair_key(p.b1, p.b2) == key) {
                r.hist_n = p.hist_n;
                r.state_n = p.state_n;
                r.states = p.states;
                for (int q = 0; q < kMaxAccum; q++) r.hist[q] = p.hist[q];
                break;
            }
            h = (h + 1) & (uint32_t)prev_mask;
        }
    }
    out[keep_scan[g]] = r;
}

void launch_bpc_build(hipStream_t s, ContactsView C, const int* n_contacts, const int* group_start, int n_groups_cap,
                      unsigned char* active, int* keep, int* keep_scan, int* d_n_keep, void* scan_temp,
                      const BpcRec* prev, int n_prev, const int* prev_hash, int prev_mask, BpcRec* out,
                      double threshold_breath) {
    (void)n_contacts;
    (void)n_prev;
    if (n_groups_cap <= 0) return;
    // the group count sits in the slot after the scan total written by launch_groups
    const int* n_groups = d_n_keep + 1;
    dim3 g((n_groups_cap + 255) / 256), b(256);
    hipLaunchKernelGGL(k_bpc_keep, g, b, 0, s, C, n_groups, group_start, active, keep, n_groups_cap);
    exclusive_scan(s, keep, keep_scan, n_groups_cap, d_n_keep, scan_temp);
    hipLaunchKernelGGL(k_bpc_write, g, b, 0, s, C, n_groups, group_start, active, keep, keep_scan, prev, prev_hash,
                       prev_mask, out, threshold_breath);
}

// Open-addressing index over a pair table (pairs with b1 < 0 are dead).
__global__ void k_bpc_hash(const BpcRec* bpc, int n, int* hash, int mask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || bpc[i].b1 < 0) return;
    uint32_t h = pair_hash(pair_key(bpc[i].b1, bpc[i].b2)) & (uint32_t)mask;
    for (int probe = 0; probe <= mask; probe++) {
        if (atomicCAS(&hash[h], -1, i) == -1) return;
        h = (h + 1) & (uint32_t)mask;
    }
}
void launch_bpc_hash(hipStream_t s, const BpcRec* bpc, int n, int* hash, int mask) {
    if (n > 0) hipLaunchKernelGGL(k_bpc_hash, dim3((n + 255) / 256), dim3(256), 0, s, bpc, n, hash, mask);
}

AMD_INLINE int top_of(const BodiesView& B, int b) { return B.parent[b] >= 0 ? B.parent[b] : b; }
AMD_INLINE bool pinned(const BodiesView& B, int e) { return (B.flags[e] & AM_BODY_PINNED) != 0; }
// BodyPairContact.java:88-95 and :130-137: a pinned body leaves the other's
// own motion; otherwise the relative motion at the velocity or position level.
AMD_INLINE double pair_metric(const BodiesView& B, int a, int b, const PoseLevel& L) {
    if (pinned(B, a)) return metric_one(B, b);
    if (pinned(B, b)) return metric_one(B, a);
    return L.xa ? metric_two_pos(B, L.xa, L.Ra, a, b, L.dt) : metric_two(B, a, b);
}

// accumulateForMerging (BodyPairContact.java:83-121) followed by the parts of
// checkMergeCondition (:151-177) that do not depend on the merges done
// earlier in the same call: breathing, pinned bodies, the motion-metric
// history (:187-198) and stable contacts (:212-224).
__global__ void k_bpc_accumulate(BodiesView B, BpcRec* bpc, int n, MergeParams P, PoseLevel L) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    BpcRec& r = bpc[i];
    const int accum = P.step_accum_merging;
    double metric = pair_metric(B, top_of(B, r.b1), top_of(B, r.b2), L);
    // motionMetricHist.add, then drop the oldest beyond stepAccumMerging
    if (r.hist_n < accum) {
        r.hist[r.hist_n++] = metric;
    } else if (accum > 0) {
        for (int q = 0; q + 1 < accum; q++) r.hist[q] = r.hist[q + 1];
        r.hist[accum - 1] = metric;
    }
    int state;
    if (r.not_on_edge == 0) state = AM_STATE_ONEDGE;
    else if (r.not_on_edge >= 2) state = AM_STATE_CLEAR;
    else if (r.n_active == 1) state = AM_STATE_CLEAR;
    else state = AM_STATE_ONEDGE;
    if (r.state_n < accum) {
        r.states |= state << (2 * r.state_n);
        r.state_n++;
    } else if (accum > 0) {
        r.states = (r.states >> 2) | (state << (2 * (accum - 1)));
    }
    int flag = 1;
    if (P.let_it_breathe && r.breath) flag = 0;
    if (!P.merge_pinned && (pinned(B, r.b1) || pinned(B, r.b2))) flag = 0;
    // a body on a prescribed motion never merges (its motion would freeze into a pinned collection)
    if ((B.flags[r.b1] | B.flags[r.b2]) & AM_BODY_KINEMATIC) flag = 0;
    if (r.hist_n != accum) {
        flag = 0;
    } else {
        for (int q = 0; q < accum; q++)
            if (r.hist[q] > P.threshold_merge) flag = 0;
    }
    if (P.stable_contact) {
        if (r.state_n != accum) {
            flag = 0;
        } else {
            for (int q = 0; q < accum; q++)
                if (((r.states >> (2 * q)) & 3) == AM_STATE_ONEDGE) flag = 0;
        }
    }
    r.flag = flag;
}
void launch_bpc_accumulate(hipStream_t s, BodiesView B, BpcRec* bpc, int n, MergeParams P, PoseLevel L) {
    if (n > 0) hipLaunchKernelGGL(k_bpc_accumulate, dim3((n + 127) / 128), dim3(128), 0, s, B, bpc, n, P, L);
}

// accumulateForUnmerging (BodyPairContact.java:127-145) on the members'
// velocities after the single sweep, then the unmerge test of
// Merging.unmerge: checkContactsState (:255-266) or
// checkMotionMetricForUnmerging (:200-205).
__global__ void k_ibpc_accumulate(BodiesView B, IbpcRec* ib, int n, ContactsView IC, MergeParams P, PoseLevel L) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    IbpcRec& r = ib[i];
    if (B.flags[r.coll] & AM_BODY_SLEEPING) return;
    double metric = pair_metric(B, r.b1, r.b2, L);
    if (metric > P.threshold_unmerge) r.hist_n = min(r.hist_n + 1, 1 << 30);
    else r.hist_n = 0;
    int flag = 0;
    for (int k = r.first; k < r.first + r.count; k++) {
        int st = IC.state[k];
        if (st == AM_STATE_BROKEN && P.unmerge_normal) flag = 1;
        if (st == AM_STATE_ONEDGE && P.unmerge_friction) flag = 1;
    }
    if (!flag && r.hist_n >= P.step_accum_unmerging && r.count < 3) flag = 1;
    r.flag = flag;
}
void launch_ibpc_accumulate(hipStream_t s, BodiesView B, IbpcRec* ib, int n, ContactsView IC, MergeParams P,
                            PoseLevel L) {
    if (n > 0) hipLaunchKernelGGL(k_ibpc_accumulate, dim3((n + 127) / 128), dim3(128), 0, s, B, ib, n, IC, P, L);
}

// ---------------------------------------------------------------------------
// Contact gathers: dst[dst_off + k] = src[idx ? idx[k] : src_off + k].
// mode 1 is the copy of Contact(Contact) (Contact.java:120-150) made when a
// pair becomes internal: lambda_warm, b and D start at zero and the contact
// is no longer new.  mode 2 re-exposes internal contacts (Merging.java:346-353):
// lambda_warm = lambda.
__global__ void k_gather_contacts(ContactsView dst, int dst_off, ContactsView src, const int* idx, int src_off, int n,
                                  int mode) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    size_t s = (size_t)(mode == 3 ? src_off + k : (idx ? idx[k] : src_off + k));
    size_t d = (size_t)(mode == 3 ? idx[k] : dst_off + k);
    dst.b1[d] = src.b1[s];
    dst.b2[d] = src.b2[s];
    dst.bv1[d] = src.bv1[s];
    dst.bv2[d] = src.bv2[s];
    dst.info[d] = src.info[s];
    dst.state[d] = src.state[s];
    dst.is_new[d] = mode == 1 ? 0 : src.is_new[s];
    dst.viol[d] = src.viol[s];
    dst.prev_viol[d] = src.prev_viol[s];
    for (int a = 0; a < 3; a++) {
        dst.pB1[d * 3 + a] = src.pB1[s * 3 + a];
        dst.nB1[d * 3 + a] = src.nB1[s * 3 + a];
        dst.t1B1[d * 3 + a] = src.t1B1[s * 3 + a];
        dst.t2B1[d * 3 + a] = src.t2B1[s * 3 + a];
        double lam = src.lam[s * 3 + a];
        dst.lam[d * 3 + a] = lam;
        dst.lamw[d * 3 + a] = mode == 1 ? 0.0 : (mode == 2 ? lam : src.lamw[s * 3 + a]);
        dst.b[d * 3 + a] = mode == 1 ? 0.0 : src.b[s * 3 + a];
        dst.D[d * 3 + a] = mode == 1 ? 0.0 : src.D[s * 3 + a];
    }
    for (int a = 0; a < 36; a++) dst.J[d * 36 + a] = src.J[s * 36 + a];
}
// ---------------------------------------------------------------------------
// Sleeping (Sleeping.java).

// Sleeping.wake's test (:112-121) per contact of this step: both bodies not
// pinned and one of their top-level bodies asleep.
__global__ void k_wake_count(BodiesView B, const int* b1, const int* b2, int n, int* count) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int a = b1[k], b = b2[k];
    if ((B.flags[a] | B.flags[b]) & AM_BODY_PINNED) return;
    if ((B.flags[top_of(B, a)] | B.flags[top_of(B, b)]) & AM_BODY_SLEEPING) atomicAdd(count, 1);
}
void launch_wake_count(hipStream_t s, BodiesView B, const int* b1, const int* b2, int n, int* count) {
    if (n > 0) hipLaunchKernelGGL(k_wake_count, dim3((n + 255) / 256), dim3(256), 0, s, B, b1, b2, n, count);
}

// Sleeping.sleep's external-contact test (:64-72): the body pairs left by
// clearBodyPairContacts (a contact with |lambda0| > 1e-14) that are not inside
// a collection and join two bodies that are not pinned mark both top-level
// bodies.
__global__ void k_sleep_external(BodiesView B, const int* b1, const int* b2, const double* lam, int n, int* ext) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (!(fabs(lam[(size_t)k * 3]) > 1e-14)) return;
    int a = b1[k], b = b2[k];
    if ((B.flags[a] | B.flags[b]) & AM_BODY_PINNED) return;
    if (B.parent[a] >= 0 && B.parent[a] == B.parent[b]) return;
    ext[top_of(B, a)] = 1;
    ext[top_of(B, b)] = 1;
}
void launch_sleep_external(hipStream_t s, BodiesView B, const int* b1, const int* b2, const double* lam, int n,
                           int* ext) {
    if (n > 0) hipLaunchKernelGGL(k_sleep_external, dim3((n + 255) / 256), dim3(256), 0, s, B, b1, b2, lam, n, ext);
}

// Sleeping.sleep (:49-98) and accumulate (:155-160) per top-level body:
// append the one-body motion metric, drop the oldest value once the history
// is longer than stepAccum, then sleep if the history is full, every value is
// at most the threshold and none exceeds its predecessor by more than 5e-5.
// counts[0] += bodies that fell asleep, counts[1] += bodies asleep afterwards.
__global__ void k_sleep_decide(BodiesView B, int* flags, const int* slot_entity, int n_slots, const int* ext,
                               double* hist, int* hist_n, int accum, double threshold, int* counts) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    int e = slot_entity[s];
    if (flags[e] & AM_BODY_SLEEPING) {
        atomicAdd(&counts[1], 1);
        return;
    }
    // a body on a prescribed motion never sleeps
    if (ext[e] || (flags[e] & AM_BODY_KINEMATIC)) return;
    double* h = hist + (size_t)e * kSleepHistStride;
    int n = hist_n[e];
    h[n++] = metric_one(B, e);
    if (n > accum) {
        for (int i = 1; i < n; i++) h[i - 1] = h[i];
        n--;
    }
    hist_n[e] = n;
    bool sleep = n >= accum;
    double prev = 1.7976931348623157e308;
    for (int i = 0; sleep && i < n; i++) {
        if (h[i] > prev + 5e-5 || h[i] > threshold) sleep = false;
        prev = h[i];
    }
    if (sleep) {
        flags[e] |= AM_BODY_SLEEPING;
        atomicAdd(&counts[0], 1);
        atomicAdd(&counts[1], 1);
    }
}
void launch_sleep_decide(hipStream_t s, BodiesView B, int* flags, const int* slot_entity, int n_slots, const int* ext,
                         double* hist, int* hist_n, int accum, double threshold, int* counts) {
    if (n_slots > 0)
        hipLaunchKernelGGL(k_sleep_decide, dim3((n_slots + 127) / 128), dim3(128), 0, s, B, flags, slot_entity,
                           n_slots, ext, hist, hist_n, accum, threshold, counts);
}

void launch_gather_contacts(hipStream_t s, ContactsView dst, int dst_off, ContactsView src, const int* idx,
                            int src_off, int n, int mode) {
    if (n > 0)
        hipLaunchKernelGGL(k_gather_contacts, dim3((n + 127) / 128), dim3(128), 0, s, dst, dst_off, src, idx, src_off,
                           n, mode);
}

}  // namespace am

// ===========================================================================
// Host side.
namespace am_rt {

using namespace am;

namespace {

// ---- javax.vecmath operations in the reference's operation order ----------
struct V3 {
    double x, y, z;
};
struct M3 {
    double m[9];
};
V3 v3(const double* p) { return V3{p[0], p[1], p[2]}; }
void put(double* p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
M3 m3(const double* p) {
    M3 a;
    for (int i = 0; i < 9; i++) a.m[i] = p[i];
    return a;
}
M3 zero3() {
    M3 a;
    for (int i = 0; i < 9; i++) a.m[i] = 0.0;
    return a;
}
M3 ident3() {
    M3 a = zero3();
    a.m[0] = a.m[4] = a.m[8] = 1.0;
    return a;
}
// Matrix3d.transform: m00*x + m01*y + m02*z
V3 mulv(const M3& A, V3 v) {
    return V3{A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z, A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
              A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z};
}
// Matrix3d.mul(A, B)
M3 mmul(const M3& A, const M3& B) {
    M3 C;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C.m[i * 3 + j] = A.m[i * 3 + 0] * B.m[0 * 3 + j] + A.m[i * 3 + 1] * B.m[1 * 3 + j] +
                             A.m[i * 3 + 2] * B.m[2 * 3 + j];
    return C;
}
// Matrix3d.mulTransposeRight(A, B) = A B^T
M3 mmul_bt(const M3& A, const M3& B) {
    M3 C;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C.m[i * 3 + j] = A.m[i * 3 + 0] * B.m[j * 3 + 0] + A.m[i * 3 + 1] * B.m[j * 3 + 1] +
                             A.m[i * 3 + 2] * B.m[j * 3 + 2];
    return C;
}
// Matrix3d.mulTransposeLeft(A, B) = A^T B
M3 mmul_at(const M3& A, const M3& B) {
    M3 C;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C.m[i * 3 + j] = A.m[0 * 3 + i] * B.m[0 * 3 + j] + A.m[1 * 3 + i] * B.m[1 * 3 + j] +
                             A.m[2 * 3 + i] * B.m[2 * 3 + j];
    return C;
}
M3 transpose(const M3& A) {
    M3 C;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C.m[i * 3 + j] = A.m[j * 3 + i];
    return C;
}
// RigidTransform3D.computeRM0RT / computeRTMR (RigidTransform3D.java:225-242)
M3 rm0rt(const M3& R, const M3& M) { return mmul_bt(mmul(R, M), R); }
M3 rtmr(const M3& R, const M3& M) { return mmul(mmul_at(R, M), R); }
// 3x3 inverse by cofactors (Matrix3d.invert; vecmath's own elimination order
// is not reproduced).
M3 invert3(const M3& A) {
    const double* a = A.m;
    double c00 = a[4] * a[8] - a[5] * a[7];
    double c01 = a[5] * a[6] - a[3] * a[8];
    double c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double s = 1.0 / det;
    M3 r;
    r.m[0] = c00 * s;
    r.m[1] = (a[2] * a[7] - a[1] * a[8]) * s;
    r.m[2] = (a[1] * a[5] - a[2] * a[4]) * s;
    r.m[3] = c01 * s;
    r.m[4] = (a[0] * a[8] - a[2] * a[6]) * s;
    r.m[5] = (a[2] * a[3] - a[0] * a[5]) * s;
    r.m[6] = c02 * s;
    r.m[7] = (a[1] * a[6] - a[0] * a[7]) * s;
    r.m[8] = (a[0] * a[4] - a[1] * a[3]) * s;
    return r;
}
void madd(M3& A, const M3& B) {
    for (int i = 0; i < 9; i++) A.m[i] = A.m[i] + B.m[i];
}
void msub(M3& A, const M3& B) {
    for (int i = 0; i < 9; i++) A.m[i] = A.m[i] - B.m[i];
}
// Java Math.max / Math.min
double jmax_h(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0) return std::signbit(a) ? b : a;
    return a >= b ? a : b;
}
double jmin_h(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0) return std::signbit(a) ? a : b;
    return a <= b ? a : b;
}

// RigidCollection.getOp (:808-830): parallel-axis term of body at distance
// (x - com), scaled by the body's mass.
M3 get_op(V3 bx, double mass, V3 com) {
    double x = bx.x - com.x, y = bx.y - com.y, z = bx.z - com.z;
    double x2 = x * x, y2 = y * y, z2 = z * z;
    M3 op;
    op.m[0] = y2 + z2;
    op.m[1] = -x * y;
    op.m[2] = -x * z;
    op.m[3] = -y * x;
    op.m[4] = x2 + z2;
    op.m[5] = -y * z;
    op.m[6] = -z * x;
    op.m[7] = -z * y;
    op.m[8] = x2 + y2;
    for (int i = 0; i < 9; i++) op.m[i] = op.m[i] * mass;
    return op;
}

// Runs body(k) for every k in [0, n) on up to kEventThreads host threads
// (at most max_threads when that is positive), in contiguous ranges of about
// equal total weight (serially when the work is small).  The jobs of an
// event write disjoint entities, so the result is the serial one.
constexpr int kEventThreads = 16;
constexpr long kWorkPerThread = 4096;  // weight units (bodies) worth a thread
template <class W, class F>
void parallel_jobs(int max_threads, int n, W weight, F body) {
    long total = 0;
    for (int k = 0; k < n; k++) total += weight(k);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int cap = max_threads > 0 ? std::min(max_threads, kEventThreads) : kEventThreads;
    const int nt = (int)std::min<long>(std::min<long>(hw, cap), std::max(total / kWorkPerThread, 1L));
    if (nt <= 1 || n <= 1) {
        for (int k = 0; k < n; k++) body(k);
        return;
    }
    std::vector<int> start(nt + 1, n);
    start[0] = 0;
    long acc = 0;
    for (int k = 0, t = 1; k < n && t < nt; k++) {
        acc += weight(k);
        while (t < nt && acc * nt >= total * t) start[t++] = k + 1;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (int t = 1; t < nt; t++)
        pool.emplace_back([&, t] {
            for (int k = start[t]; k < start[t + 1]; k++) body(k);
        });
    for (int k = start[0]; k < start[1]; k++) body(k);
    for (auto& th : pool) th.join();
}

// Items listed per body in item order (compressed rows): for each item i,
// first(i) and second(i) name its bodies (-1 for none).
struct Adjacency {
    std::vector<int> start, item;
    template <class F1, class F2>
    void build(int n_bodies, int n_items, F1 first, F2 second) {
        start.assign(n_bodies + 1, 0);
        for (int i = 0; i < n_items; i++) {
            const int a = first(i), b = second(i);
            if (a >= 0) start[a + 1]++;
            if (b >= 0) start[b + 1]++;
        }
        for (int k = 0; k < n_bodies; k++) start[k + 1] += start[k];
        item.resize(start[n_bodies]);
        std::vector<int> at(start.begin(), start.end() - 1);
        for (int i = 0; i < n_items; i++) {
            const int a = first(i), b = second(i);
            if (a >= 0) item[at[a]++] = i;
            if (b >= 0) item[at[b]++] = i;
        }
    }
};

// Host mirror of one entity for the duration of a merge or unmerge event.
struct Ent {
    V3 x, v, w;
    M3 R, I, I0, jinv, jinv0;  // massAngular, massAngular0, jinv, jinv0
    double mass, minv;
    bool pinned, plane;
    bool sleeping;  // RigidBody.sleeping (a collection's own flag)
    double bb[24];
};

struct Event {
    am_ctx* c;
    PinnedVec<double> x, R, v, w, jinv;    // device state of every entity
    std::vector<char> touched;             // entity state changed on the host
    std::vector<char> dv_zero;             // deltaV to clear (unmergeBody)
    std::vector<std::vector<int>> members; // per collection entity: RigidCollection.bodies
    // Top-level list edits.  A removal (ArrayList.remove, the others keep
    // their order) marks the entity's position; the list is compacted once,
    // when the event is committed.
    std::vector<int> top_pos;    // entity -> position in c->slot_entity, -1 if not listed
    std::vector<char> top_gone;  // position removed
    // Collections whose member transforms are recomputed at commit from the
    // collection's stored pose (members keep their world poses during an
    // event, so the last pose a collection stores gives the same transforms
    // as recomputing them after each change).
    std::vector<char> b2c_dirty;
    std::vector<int> b2c_list;
    std::vector<char> b2c_changed;  // body transforms to upload
    int n_begin = 0;                // entities when the event began

    void top_remove(int e) {
        if (e < (int)top_pos.size() && top_pos[e] >= 0) {
            top_gone[top_pos[e]] = 1;
            top_pos[e] = -1;
        }
    }
    void top_append(int e) {
        if (e >= (int)top_pos.size()) top_pos.resize(e + 1, -1);
        top_pos[e] = (int)c->slot_entity.size();
        c->slot_entity.push_back(e);
        top_gone.push_back(0);
    }
    void mark_b2c(int e) {
        if (e >= (int)b2c_dirty.size()) b2c_dirty.resize(e + 1, 0);
        if (!b2c_dirty[e]) {
            b2c_dirty[e] = 1;
            b2c_list.push_back(e);
        }
    }

    Ent load(int e) const {
        const am_body_desc& d = c->desc[e];
        Ent E;
        E.x = v3(&x[e * 3]);
        E.v = v3(&v[e * 3]);
        E.w = v3(&w[e * 3]);
        E.R = m3(&R[e * 9]);
        E.jinv = m3(&jinv[e * 9]);
        E.jinv0 = m3(d.jinv0);
        E.I0 = m3(d.inertia0);
        E.pinned = (c->flags[e] & AM_BODY_PINNED) != 0;
        E.sleeping = (c->flags[e] & AM_BODY_SLEEPING) != 0;
        E.plane = d.geom == AM_GEOM_PLANE;
        // massAngular as RigidBody.updateRotationalInertiaFromTransformation left it
        E.I = E.pinned ? zero3() : rm0rt(E.R, E.I0);
        E.mass = E.pinned ? 0.0 : d.mass;
        E.minv = E.pinned ? 0.0 : d.minv;
        for (int i = 0; i < 24; i++) E.bb[i] = d.bbox[i];
        return E;
    }
    // Writes a collection's state and mass properties back to the mirrors.
    void store_collection(int e, const Ent& E) {
        put(&x[e * 3], E.x);
        put(&v[e * 3], E.v);
        put(&w[e * 3], E.w);
        for (int i = 0; i < 9; i++) {
            R[e * 9 + i] = E.R.m[i];
            jinv[e * 9 + i] = E.jinv.m[i];
        }
        am_body_desc& d = c->desc[e];
        std::memset(&d, 0, sizeof d);
        d.geom = AM_GEOM_COLLECTION;
        d.flags = (E.pinned ? AM_BODY_PINNED : 0) | (E.sleeping ? AM_BODY_SLEEPING : 0);
        // members share one scene (bodies of different scenes never touch)
        d.scene = members[e].empty() ? 0 : c->desc[members[e][0]].scene;
        d.mass = E.mass;
        d.minv = E.minv;
        for (int i = 0; i < 9; i++) {
            d.jinv0[i] = E.jinv0.m[i];
            d.inertia0[i] = E.I0.m[i];
        }
        for (int i = 0; i < 24; i++) d.bbox[i] = E.bb[i];
        d.friction = 0.8;
        c->flags[e] = d.flags;
        touched[e] = 1;
    }
};

// RigidCollection.updateBB(body) (:612-650), bounding box temporarily in world
// coordinates; skipped for planes.
void update_bb_with(Ent& C, const Ent& B) {
    if (B.plane) return;
    // Double.MAX_VALUE, not infinity, as the reference initialises them
    double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    const Ent* two[2] = {&C, &B};
    for (int s = 0; s < 2; s++) {
        const Ent& b = *two[s];
        for (int q = 0; q < 8; q++) {
            V3 p = mulv(b.R, v3(&b.bb[q * 3]));
            p = V3{p.x + b.x.x, p.y + b.x.y, p.z + b.x.z};
            mn[0] = jmin_h(mn[0], p.x);
            mn[1] = jmin_h(mn[1], p.y);
            mn[2] = jmin_h(mn[2], p.z);
            mx[0] = jmax_h(mx[0], p.x);
            mx[1] = jmax_h(mx[1], p.y);
            mx[2] = jmax_h(mx[2], p.z);
        }
    }
    const double pts[8][3] = {{mx[0], mx[1], mx[2]}, {mx[0], mn[1], mn[2]}, {mn[0], mx[1], mn[2]},
                              {mn[0], mn[1], mx[2]}, {mn[0], mn[1], mn[2]}, {mn[0], mx[1], mx[2]},
                              {mx[0], mn[1], mx[2]}, {mx[0], mx[1], mn[2]}};
    for (int q = 0; q < 8; q++)
        for (int a = 0; a < 3; a++) C.bb[q * 3 + a] = pts[q][a];
}

// RigidCollection.addBodyInternalMethod (:477-514).
void add_body_internal(Ent& C, const Ent& B) {
    C.R = ident3();  // updateTheta: the collection frame is world-aligned
    if (C.pinned) {
        C.v = V3{0, 0, 0};
        C.w = V3{0, 0, 0};
        C.I = zero3();
        C.I0 = zero3();
        C.jinv = zero3();
        C.jinv0 = zero3();
        C.mass = 0;
        C.minv = 0;
        return;
    }
    V3 com = C.x;
    com = V3{com.x * C.mass, com.y * C.mass, com.z * C.mass};
    com = V3{B.mass * B.x.x + com.x, B.mass * B.x.y + com.y, B.mass * B.x.z + com.z};
    double inv = 1. / (B.mass + C.mass);
    com = V3{com.x * inv, com.y * inv, com.z * inv};
    // updateVelocitiesFrom (:526-534)
    V3 v = V3{C.v.x * C.mass, C.v.y * C.mass, C.v.z * C.mass};
    v = V3{B.mass * B.v.x + v.x, B.mass * B.v.y + v.y, B.mass * B.v.z + v.z};
    C.v = V3{v.x * inv, v.y * inv, v.z * inv};
    V3 w = V3{C.w.x * C.mass, C.w.y * C.mass, C.w.z * C.mass};
    w = V3{B.mass * B.w.x + w.x, B.mass * B.w.y + w.y, B.mass * B.w.z + w.z};
    C.w = V3{w.x * inv, w.y * inv, w.z * inv};
    // updateInertia (:791-803)
    M3 M0 = zero3();
    madd(M0, C.I);
    madd(M0, get_op(C.x, C.mass, com));
    madd(M0, B.I);
    madd(M0, get_op(B.x, B.mass, com));
    C.I = M0;
    update_bb_with(C, B);
    C.mass = C.mass + B.mass;
    C.minv = inv;
    C.x = com;
    // bounding box back in collection coordinates, R^T (p - x)
    for (int q = 0; q < 8; q++) {
        double px = C.bb[q * 3] - C.x.x, py = C.bb[q * 3 + 1] - C.x.y, pz = C.bb[q * 3 + 2] - C.x.z;
        const double* r = C.R.m;
        C.bb[q * 3 + 0] = r[0] * px + r[3] * py + r[6] * pz;
        C.bb[q * 3 + 1] = r[1] * px + r[4] * py + r[7] * pz;
        C.bb[q * 3 + 2] = r[2] * px + r[5] * py + r[8] * pz;
    }
}

// updateInertiaRestAndInvert (:674-680) and updateRotationalInertiaFromTransformation.
void finish_inertia(Ent& C) {
    if (C.pinned) return;
    C.jinv = invert3(C.I);
    C.I0 = rtmr(C.R, C.I);
    C.jinv0 = rtmr(C.R, C.jinv);
    C.I = rm0rt(C.R, C.I0);
    C.jinv = rm0rt(C.R, C.jinv0);
}


// RigidTransform3D.multAinvB (:201-215): body-to-collection transform of
// member b, t = R_c^T (x_b - x_c), R = R_c^T R_b.
void body_to_collection(V3 xc, const M3& Rc, V3 xb, const M3& Rb, double* out12) {
    V3 t = V3{xb.x - xc.x, xb.y - xc.y, xb.z - xc.z};
    M3 Rt = transpose(Rc);
    t = mulv(Rt, t);
    M3 R = mmul(Rt, Rb);
    for (int i = 0; i < 9; i++) out12[i] = R.m[i];
    put(out12 + 9, t);
}

// RigidCollection.updateBodiesTransformations (:589-592).
void update_bodies_transformations(Event& E, int ce, const Ent& C) {
    am_ctx* c = E.c;
    for (int b : E.members[ce]) {
        body_to_collection(C.x, C.R, v3(&E.x[b * 3]), m3(&E.R[b * 9]), &c->b2c[(size_t)b * 12]);
        E.b2c_changed[b] = 1;
    }
}

// The deferred updateBodiesTransformations of the marked collections, from
// the pose each one stored last.
void flush_transformations(Event& E) {
    am_ctx* c = E.c;
    const std::vector<int>& L = E.b2c_list;
    auto weight = [&](int k) { return (long)E.members[L[k]].size(); };
    parallel_jobs(c->params.host_threads, (int)L.size(), weight, [&](int k) {
        const int ce = L[k];
        E.b2c_dirty[ce] = 0;
        const V3 xc = v3(&E.x[ce * 3]);
        const M3 Rc = m3(&E.R[ce * 9]);
        for (int b : E.members[ce]) {
            body_to_collection(xc, Rc, v3(&E.x[b * 3]), m3(&E.R[b * 9]), &c->b2c[(size_t)b * 12]);
            E.b2c_changed[b] = 1;
        }
    });
    E.b2c_list.clear();
}

int alloc_entity(Event& E) {
    am_ctx* c = E.c;
    if (c->free_entities.empty()) {
        int old = c->n_ent, add = std::max(16, c->n_phys / 4);
        int n = old + add;
        am_body_desc blank;
        std::memset(&blank, 0, sizeof blank);
        blank.geom = AM_GEOM_COLLECTION;
        blank.flags = AM_BODY_PINNED;
        c->desc.resize(n, blank);
        c->flags.resize(n, AM_BODY_PINNED);
        c->parent.resize(n, -1);
        c->slot.resize(n, -1);
        c->sub.resize(n, 0);
        c->b2c.resize((size_t)n * 12, 0.0);
        E.x.resize((size_t)n * 3, 0.0);
        E.v.resize((size_t)n * 3, 0.0);
        E.w.resize((size_t)n * 3, 0.0);
        E.R.resize((size_t)n * 9, 0.0);
        E.jinv.resize((size_t)n * 9, 0.0);
        for (int e = old; e < n; e++) E.R[(size_t)e * 9] = E.R[(size_t)e * 9 + 4] = E.R[(size_t)e * 9 + 8] = 1.0;
        E.touched.resize(n, 0);
        E.dv_zero.resize(n, 0);
        E.members.resize(n);
        E.top_pos.resize(n, -1);
        E.b2c_dirty.resize(n, 0);
        E.b2c_changed.resize(n, 0);
        for (int e = n - 1; e >= old; e--) c->free_entities.push_back(e);
        c->n_ent = n;
    }
    int e = c->free_entities.back();
    c->free_entities.pop_back();
    E.members[e].clear();
    c->sleep_fresh.push_back(e);  // a new RigidCollection has an empty metric history
    return e;
}

void free_entity(Event& E, int e) {
    am_ctx* c = E.c;
    E.members[e].clear();
    std::memset(&c->desc[e], 0, sizeof(am_body_desc));
    c->desc[e].geom = AM_GEOM_COLLECTION;
    c->desc[e].flags = AM_BODY_PINNED;
    c->flags[e] = AM_BODY_PINNED;
    c->parent[e] = -1;
    E.touched[e] = 1;
    c->free_entities.push_back(e);
}

// RigidCollection.updateCollectionState (:459-470): the collection is pinned
// if the body is and asleep if the body was; the body's own flag is cleared.
void collection_state(am_ctx* c, Ent& C, int b, const Ent& B) {
    C.pinned = C.pinned || B.pinned;
    C.sleeping = C.sleeping || B.sleeping;
    c->flags[b] &= ~AM_BODY_SLEEPING;
    c->desc[b].flags &= ~AM_BODY_SLEEPING;
}

// RigidCollection.addBody (:136-147).
void add_body(Event& E, int ce, Ent& C, int b) {
    am_ctx* c = E.c;
    c->parent[b] = ce;
    E.members[ce].push_back(b);
    Ent B = E.load(b);
    collection_state(c, C, b, B);
    add_body_internal(C, B);
    finish_inertia(C);
    E.mark_b2c(ce);  // updateBodiesTransformations
}

// RigidCollection.addBodies (:153-166).
void add_bodies(Event& E, int ce, Ent& C, const std::vector<int>& list) {
    am_ctx* c = E.c;
    for (int b : list) {
        c->parent[b] = ce;
        E.members[ce].push_back(b);
        Ent B = E.load(b);
        collection_state(c, C, b, B);
        add_body_internal(C, B);
    }
    finish_inertia(C);
    E.mark_b2c(ce);
}

// RigidCollection.addCollection (:172-186): the other collection's members
// join, its state is averaged in as one body.
void add_collection(Event& E, int ce, Ent& C, int other) {
    am_ctx* c = E.c;
    for (int b : E.members[other]) {
        c->parent[b] = ce;
        E.members[ce].push_back(b);
    }
    Ent O = E.load(other);
    collection_state(c, C, other, O);
    add_body_internal(C, O);
    finish_inertia(C);
    E.mark_b2c(ce);
}

// new RigidCollection(body1, body2) (:55-76): starts as a copy of the first
// body (RigidBody.set) unless that one is a plane.
int new_collection(Event& E, int b1, int b2, Ent& C) {
    am_ctx* c = E.c;
    int ce = alloc_entity(E);
    if (c->desc[b1].geom == AM_GEOM_PLANE) std::swap(b1, b2);
    C = E.load(b1);
    C.plane = false;
    c->parent[b1] = ce;
    E.members[ce].push_back(b1);
    add_body(E, ce, C, b2);
    return ce;
}

// RigidCollection.applyVelocitiesTo + unmergeBody (:925-951).
void unmerge_body(Event& E, const Ent& C, int b) {
    am_ctx* c = E.c;
    V3 xb = v3(&E.x[b * 3]);
    V3 r = V3{xb.x - C.x.x, xb.y - C.x.y, xb.z - C.x.z};
    V3 wxr = V3{C.w.y * r.z - C.w.z * r.y, C.w.z * r.x - C.w.x * r.z, C.w.x * r.y - C.w.y * r.x};
    put(&E.v[b * 3], V3{C.v.x + wxr.x, C.v.y + wxr.y, C.v.z + wxr.z});
    put(&E.w[b * 3], C.w);
    E.touched[b] = 1;
    E.dv_zero[b] = 1;
    c->parent[b] = -1;
}

// RigidCollection.updateBB() (:561-606): box of the members in collection
// coordinates.
void update_bb_members(Event& E, int ce, Ent& C) {
    am_ctx* c = E.c;
    double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    for (int b : E.members[ce]) {
        const am_body_desc& d = c->desc[b];
        if (d.geom == AM_GEOM_PLANE) continue;
        M3 Rbc = m3(&c->b2c[(size_t)b * 12]);
        V3 tbc = v3(&c->b2c[(size_t)b * 12 + 9]);
        for (int q = 0; q < 8; q++) {
            V3 p = mulv(Rbc, v3(&d.bbox[q * 3]));
            p = V3{p.x + tbc.x, p.y + tbc.y, p.z + tbc.z};
            mn[0] = jmin_h(mn[0], p.x);
            mn[1] = jmin_h(mn[1], p.y);
            mn[2] = jmin_h(mn[2], p.z);
            mx[0] = jmax_h(mx[0], p.x);
            mx[1] = jmax_h(mx[1], p.y);
            mx[2] = jmax_h(mx[2], p.z);
        }
    }
    const double pts[8][3] = {{mx[0], mx[1], mx[2]}, {mx[0], mn[1], mn[2]}, {mn[0], mx[1], mn[2]},
                              {mn[0], mn[1], mx[2]}, {mn[0], mn[1], mn[2]}, {mn[0], mx[1], mx[2]},
                              {mx[0], mn[1], mx[2]}, {mx[0], mx[1], mn[2]}};
    for (int q = 0; q < 8; q++)
        for (int a = 0; a < 3; a++) C.bb[q * 3 + a] = pts[q][a];
}

// RigidCollection.removeBodies (:686-753).
void remove_bodies(Event& E, int ce, Ent& C, const std::vector<int>& gone) {
    am_ctx* c = E.c;
    auto& mem = E.members[ce];
    std::vector<int> sorted_gone = gone;
    std::sort(sorted_gone.begin(), sorted_gone.end());
    std::vector<int> keep;
    for (int b : mem)
        if (!std::binary_search(sorted_gone.begin(), sorted_gone.end(), b)) keep.push_back(b);
    mem = keep;
    bool was_pinned = C.pinned;
    C.pinned = false;
    for (int b : mem) C.pinned = C.pinned || ((c->flags[b] & AM_BODY_PINNED) != 0);
    C.R = ident3();
    if (C.pinned) {
        C.v = V3{0, 0, 0};
        C.w = V3{0, 0, 0};
        C.I = zero3();
        C.I0 = zero3();
        C.jinv = zero3();
        C.jinv0 = zero3();
        C.mass = 0;
        C.minv = 0;
    } else if (was_pinned) {
        C.x = V3{0, 0, 0};
        C.mass = 0;
        for (int b : mem) {
            Ent B = E.load(b);
            C.x = V3{B.mass * B.x.x + C.x.x, B.mass * B.x.y + C.x.y, B.mass * B.x.z + C.x.z};
            C.mass = C.mass + B.mass;
        }
        C.minv = 1. / C.mass;
        C.x = V3{C.x.x * C.minv, C.x.y * C.minv, C.x.z * C.minv};
        // computeInertia (:766-774)
        M3 I = zero3();
        for (int b : mem) {
            Ent B = E.load(b);
            madd(I, B.I);
            madd(I, get_op(B.x, B.mass, C.x));
        }
        C.I = I;
        C.jinv = invert3(C.I);
        C.I0 = rtmr(C.R, C.I);
        C.jinv0 = rtmr(C.R, C.jinv);
    } else {
        for (int b : gone) {
            Ent B = E.load(b);
            V3 com = C.x;
            com = V3{com.x * C.mass, com.y * C.mass, com.z * C.mass};
            C.x = V3{-B.mass * B.x.x + com.x, -B.mass * B.x.y + com.y, -B.mass * B.x.z + com.z};
            double s = 1. / C.mass;
            com = V3{com.x * s, com.y * s, com.z * s};
            C.mass = C.mass - B.mass;
            C.minv = 1. / C.mass;
            C.x = V3{C.x.x * C.minv, C.x.y * C.minv, C.x.z * C.minv};
            // updateInertiaReverse (:779-789)
            msub(C.I, B.I);
            msub(C.I, get_op(C.x, C.mass, com));
            msub(C.I, get_op(B.x, B.mass, com));
        }
        C.jinv = invert3(C.I);
        C.I0 = rtmr(C.R, C.I);
        C.jinv0 = rtmr(C.R, C.jinv);
    }
    if (!C.pinned) {
        C.I = rm0rt(C.R, C.I0);
        C.jinv = rm0rt(C.R, C.jinv0);
    }
    update_bodies_transformations(E, ce, C);
    update_bb_members(E, ce, C);
}

int begin_event(am_ctx* c, Event& E) {
    E.c = c;
    size_t n = (size_t)c->n_ent;
    E.x.swap(c->ev_x);
    E.v.swap(c->ev_v);
    E.w.swap(c->ev_w);
    E.R.swap(c->ev_R);
    E.jinv.swap(c->ev_jinv);
    E.members.swap(c->ev_members);
    E.x.resize(n * 3);
    E.v.resize(n * 3);
    E.w.resize(n * 3);
    E.R.resize(n * 9);
    E.jinv.resize(n * 9);
    int r;
    if ((r = pull_d(c, E.x.data(), c->d_x.p, n * 3))) return r;
    if ((r = pull_d(c, E.v.data(), c->d_v.p, n * 3))) return r;
    if ((r = pull_d(c, E.w.data(), c->d_w.p, n * 3))) return r;
    if ((r = pull_d(c, E.R.data(), c->d_R.p, n * 9))) return r;
    if ((r = pull_d(c, E.jinv.data(), c->d_jinv.p, n * 9))) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    E.touched.assign(n, 0);
    E.dv_zero.assign(n, 0);
    for (auto& m : E.members) m.clear();
    E.members.resize(n);
    E.top_pos.assign(n, -1);
    E.top_gone.assign(c->slot_entity.size(), 0);
    for (int q = 0; q < (int)c->slot_entity.size(); q++) E.top_pos[c->slot_entity[q]] = q;
    E.b2c_dirty.assign(n, 0);
    E.b2c_list.clear();
    E.b2c_changed.assign(n, 0);
    E.n_begin = (int)n;
    for (int e : c->slot_entity)
        if (is_collection(c, e))
            for (int q = 0; q < c->member_count[e]; q++) E.members[e].push_back(c->members[c->member_first[e] + q]);
    return AM_OK;
}

// Flattens the structure, uploads what the event changed.
int commit_event(am_ctx* c, Event& E) {
    hipStream_t s = c->stream;
    const int n = c->n_ent;
    flush_transformations(E);
    {
        auto& t = c->slot_entity;
        int k = 0;
        for (int q = 0; q < (int)t.size(); q++)
            if (!E.top_gone[q]) t[k++] = t[q];
        t.resize(k);
    }
    const int old_n = (int)(c->d_flags.n);
    // Only the entities the event changed are written back while the entity
    // count stays; a grown count reallocates and uploads every array.
    const bool sparse = n == E.n_begin && c->d_x.n >= (size_t)n * 3 && c->d_desc.n >= (size_t)n &&
                        c->d_minv.n >= (size_t)n && c->d_b2c.n >= (size_t)n * 12;
    if (sparse) {
        std::vector<int> list;
        for (int e = 0; e < n; e++)
            if (E.touched[e]) list.push_back(e);
        if (!list.empty()) {
            std::vector<double> rec(list.size() * kEntRecDoubles);
            for (size_t k = 0; k < list.size(); k++) {
                const int e = list[k];
                double* o = &rec[k * kEntRecDoubles];
                for (int q = 0; q < 3; q++) {
                    o[q] = E.x[(size_t)e * 3 + q];
                    o[3 + q] = E.v[(size_t)e * 3 + q];
                    o[6 + q] = E.w[(size_t)e * 3 + q];
                }
                for (int q = 0; q < 9; q++) {
                    o[9 + q] = E.R[(size_t)e * 9 + q];
                    o[18 + q] = E.jinv[(size_t)e * 9 + q];
                }
                o[27] = (c->flags[e] & AM_BODY_PINNED) ? 0.0 : c->desc[e].minv;
                std::memcpy(o + 28, &c->desc[e], sizeof(am_body_desc));
            }
            AM_HIP(c, upload(c, c->ev_list, list));
            AM_HIP(c, upload(c, c->ev_rec, rec));
            const long m = (long)list.size() * kEntRecDoubles;
            hipLaunchKernelGGL(k_scatter_entities, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, c->d_x.p,
                               c->d_v.p, c->d_w.p, c->d_R.p, c->d_jinv.p, c->d_minv.p, c->d_desc.p, c->ev_list.p,
                               c->ev_rec.p, (int)list.size());
            AM_HIP(c, hipGetLastError());
            AM_HIP(c, hipStreamSynchronize(s));  // the staging buffers are reused below
        }
        std::vector<int> blist;
        for (int e = 0; e < n; e++)
            if (E.b2c_changed[e]) blist.push_back(e);
        if (!blist.empty()) {
            std::vector<double> rec(blist.size() * 12);
            for (size_t k = 0; k < blist.size(); k++)
                std::memcpy(&rec[k * 12], &c->b2c[(size_t)blist[k] * 12], 12 * sizeof(double));
            AM_HIP(c, upload(c, c->ev_list, blist));
            AM_HIP(c, upload(c, c->ev_rec, rec));
            const long m = (long)blist.size() * 12;
            hipLaunchKernelGGL(k_scatter_b2c, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, c->d_b2c.p,
                               c->ev_list.p, c->ev_rec.p, (int)blist.size());
            AM_HIP(c, hipGetLastError());
            AM_HIP(c, hipStreamSynchronize(s));
        }
    } else {
        // per-entity device arrays sized by n_ent
        if (c->d_x.n < (size_t)n * 3) {
            int keep = old_n;
            AM_HIP(c, grow_keep(s, c->d_f, n, 3, keep));
            AM_HIP(c, grow_keep(s, c->d_t, n, 3, keep));
            AM_HIP(c, grow_keep(s, c->d_dv, n, 6, keep));
            AM_HIP(c, c->d_x.ensure((size_t)n * 3));
            AM_HIP(c, c->d_v.ensure((size_t)n * 3));
            AM_HIP(c, c->d_w.ensure((size_t)n * 3));
            AM_HIP(c, c->d_R.ensure((size_t)n * 9));
            AM_HIP(c, c->d_jinv.ensure((size_t)n * 9));
            AM_HIP(c, c->d_minv.ensure(n));
            // clear the new tail of the accumulators
            fill_double(s, c->d_f.p + (size_t)keep * 3, (n - keep) * 3, 0.0);
            fill_double(s, c->d_t.p + (size_t)keep * 3, (n - keep) * 3, 0.0);
            fill_double(s, c->d_dv.p + (size_t)keep * 6, (n - keep) * 6, 0.0);
        }
        AM_HIP(c, hipMemcpyAsync(c->d_x.p, E.x.data(), (size_t)n * 24, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->d_v.p, E.v.data(), (size_t)n * 24, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->d_w.p, E.w.data(), (size_t)n * 24, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->d_R.p, E.R.data(), (size_t)n * 72, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->d_jinv.p, E.jinv.data(), (size_t)n * 72, hipMemcpyHostToDevice, s));
        std::vector<double> minv(n);
        for (int i = 0; i < n; i++) minv[i] = (c->flags[i] & AM_BODY_PINNED) ? 0.0 : c->desc[i].minv;
        AM_HIP(c, upload(c, c->d_minv, minv));
        AM_HIP(c, upload(c, c->d_desc, c->desc));
        AM_HIP(c, upload(c, c->d_b2c, c->b2c));
    }
    std::vector<int> dz;
    for (int i = 0; i < n; i++)
        if (E.dv_zero[i]) dz.push_back(i);
    if (!dz.empty()) {
        DBuf<int> d_dz;
        AM_HIP(c, upload(c, d_dz, dz));
        launch_zero_dv(s, bodies_view(c), d_dz.p, (int)dz.size());
        AM_HIP(c, hipStreamSynchronize(s));
    }
    // structure: top list order, members in RigidCollection.bodies order
    c->member_first.assign(n, 0);
    c->member_count.assign(n, 0);
    c->members.clear();
    c->slot.assign(n, -1);
    c->sub.assign(n, 0);
    for (int s_ = 0; s_ < (int)c->slot_entity.size(); s_++) {
        int e = c->slot_entity[s_];
        c->slot[e] = s_;
        c->parent[e] = -1;
        if (is_collection(c, e)) {
            c->member_first[e] = (int)c->members.size();
            c->member_count[e] = (int)E.members[e].size();
            int q = 0;
            for (int b : E.members[e]) {
                c->members.push_back(b);
                c->parent[b] = e;
                c->slot[b] = s_;
                c->sub[b] = q++;
            }
        }
    }
    E.x.swap(c->ev_x);
    E.v.swap(c->ev_v);
    E.w.swap(c->ev_w);
    E.R.swap(c->ev_R);
    E.jinv.swap(c->ev_jinv);
    E.members.swap(c->ev_members);
    return rebuild_structure(c, false);
}

}  // namespace

MergeParams merge_params(const am_params& p) {
    MergeParams P;
    P.step_accum_merging = p.step_accum_merging;
    P.step_accum_unmerging = p.step_accum_unmerging;
    P.merge_pinned = p.enable_merge_pinned;
    P.stable_contact = p.enable_merge_stable_contact;
    P.let_it_breathe = p.enable_merge_let_it_breathe;
    P.unmerge_friction = p.enable_unmerge_friction;
    P.unmerge_normal = p.enable_unmerge_normal;
    P.threshold_merge = p.threshold_merge;
    P.threshold_unmerge = p.threshold_unmerge;
    P.threshold_breath = p.threshold_breath;
    return P;
}

bool has_collections(const am_ctx* c) { return !c->member_list.empty(); }

hipError_t ContactStore::ensure(int want, int keep, hipStream_t s) {
    if (want <= cap && b1.p) return hipSuccess;
    int nc = std::max(want + want / 4, 256);
    size_t k = (size_t)std::min(keep, cap);
    hipError_t e;
#define AM_GROW(buf, stride) \
    if ((e = grow_keep(s, buf, nc, stride, k)) != hipSuccess) return e;
    AM_GROW(b1, 1) AM_GROW(b2, 1) AM_GROW(bv1, 1) AM_GROW(bv2, 1) AM_GROW(info, 1) AM_GROW(state, 1)
    AM_GROW(is_new, 1) AM_GROW(pB1, 3) AM_GROW(nB1, 3) AM_GROW(t1B1, 3) AM_GROW(t2B1, 3) AM_GROW(viol, 1)
    AM_GROW(prev, 1) AM_GROW(J, 36) AM_GROW(lam, 3) AM_GROW(lamw, 3) AM_GROW(b, 3) AM_GROW(D, 3)
#undef AM_GROW
    cap = nc;
    return hipSuccess;
}

void ContactStore::swap(ContactStore& o) {
#define AM_SW(buf)                \
    std::swap(buf.p, o.buf.p);    \
    std::swap(buf.n, o.buf.n);
    AM_SW(b1) AM_SW(b2) AM_SW(bv1) AM_SW(bv2) AM_SW(info) AM_SW(state) AM_SW(is_new) AM_SW(pB1) AM_SW(nB1)
    AM_SW(t1B1) AM_SW(t2B1) AM_SW(viol) AM_SW(prev) AM_SW(J) AM_SW(lam) AM_SW(lamw) AM_SW(b) AM_SW(D)
#undef AM_SW
    std::swap(cap, o.cap);
}

int merge_push_down(am_ctx* c, bool integrated_only) {
    launch_push_down(c->stream, bodies_view(c), c->d_member_list.p, (int)c->member_list.size(),
                     integrated_only ? 1 : 0);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

int merge_member_transforms(am_ctx* c) {
    launch_member_transforms(c->stream, bodies_view(c), c->d_b2c.p, c->d_member_list.p, (int)c->member_list.size());
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

// CollisionProcessor.updateInCollections (:232-303): one PGS sweep over the
// external and the internal contacts with sub-body Jacobians, then the members
// advance their velocities and the top-level deltaV is cleared.  The sweep
// runs in colour order over physical body pairs (am_solve_order_pull gives
// the order right after this call).
// ---------------------------------------------------------------------------
// CollisionProcessor.getOrganizedContacts / getNextLayer (:346-441): the
// single sweep visits the body pairs breadth first from those holding a
// contact new this step: layer 0 those pairs, layer k+1 the pairs not yet
// reached that share a body (pinned ones included) with a pair of layer k,
// then the external pairs never reached, then the internal ones.  Groups
// (contiguous same-pair runs of the combined list, external first) stand for
// the pairs; the layers order the colouring of the single sweep.
// the bodies of the groups of layer k get mark k (every writer of a body
// writes the same value)
__global__ void k_bfs_mark(int n_groups, const int* ga, const int* gb, const int* layer, int k, int* mark) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups || layer[g] != k) return;
    mark[ga[g]] = k;
    mark[gb[g]] = k;
}
__global__ void k_bfs_seed(int n_groups, const int* group_start, const int* is_new, int n_ext, int* layer,
                           int* n_seeded) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    int l = -1;
    if (group_start[g] < n_ext)
        for (int k = group_start[g]; k < group_start[g + 1]; k++)
            if (is_new[k]) {
                l = 0;
                break;
            }
    layer[g] = l;
    if (l == 0) atomicAdd(n_seeded, 1);
}
// one breadth-first layer: a group not yet reached joins layer k + 1 when a
// pair of layer k shares one of its bodies, i.e. when one of its bodies has
// mark k (groups set this round read k + 1, never k, so the result does not
// depend on scheduling)
__global__ void k_bfs_round(int n_groups, const int* ga, const int* gb, const int* mark, int* layer, int k,
                            int* changed) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups || layer[g] >= 0) return;
    if (mark[ga[g]] == k || mark[gb[g]] == k) {
        layer[g] = k + 1;
        atomicAdd(changed, 1);
    }
}
__global__ void k_bfs_rest(int n_groups, const int* group_start, int n_ext, int* layer, int lmax) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups || layer[g] >= 0) return;
    layer[g] = group_start[g] < n_ext ? lmax + 1 : lmax + 2;
}

// Layers of the groups of the combined list (external contacts [0, n_ext)).
static int organized_layers(am_ctx* c, int n_ext) {
    hipStream_t s = c->stream;
    const int ng = c->n_groups, ne = c->n_ent;
    AM_HIP(c, c->g_layer.ensure(std::max(ng, 1)));
    AM_HIP(c, c->bfs_mark.ensure(ne + 1));
    int* sc = c->d_scalars.p;
    fill_int(s, c->bfs_mark.p, ne + 1, -1);
    dim3 grid((ng + 255) / 256), block(256);
    fill_int(s, sc + 12, 1, 0);
    hipLaunchKernelGGL(k_bfs_seed, grid, block, 0, s, ng, c->g_start.p, c->c_new.p, n_ext, c->g_layer.p, sc + 12);
    AM_HIP(c, hipGetLastError());
    int r = read_scalars(c, 12, 1);
    if (r) return r;
    int lmax = -1;
    if (c->h_scalars[12] > 0) {
        lmax = 0;
        for (int k = 0; k < ng; k++) {
            fill_int(s, sc + 12, 1, 0);
            hipLaunchKernelGGL(k_bfs_mark, grid, block, 0, s, ng, c->g_a.p, c->g_b.p, c->g_layer.p, k, c->bfs_mark.p);
            hipLaunchKernelGGL(k_bfs_round, grid, block, 0, s, ng, c->g_a.p, c->g_b.p, c->bfs_mark.p, c->g_layer.p,
                               k, sc + 12);
            AM_HIP(c, hipGetLastError());
            if ((r = read_scalars(c, 12, 1))) return r;
            if (c->h_scalars[12] == 0) break;
            lmax = k + 1;
        }
    }
    hipLaunchKernelGGL(k_bfs_rest, grid, block, 0, s, ng, c->g_start.p, n_ext, c->g_layer.p, lmax);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

int merge_update_in_collections(am_ctx* c, double dt, int* iters) {
    *iters = 0;
    const am_params& P = c->params;
    if (!P.update_contacts_in_collections || !has_collections(c)) return AM_OK;
    hipStream_t s = c->stream;
    // internal contacts of the collections that are awake, pair by pair
    std::vector<int> awake_idx;
    bool all_awake = true;
    for (const auto& ib : c->ibpc) all_awake = all_awake && !(c->flags[ib.coll] & AM_BODY_SLEEPING);
    if (!all_awake)
        for (const auto& ib : c->ibpc)
            if (!(c->flags[ib.coll] & AM_BODY_SLEEPING))
                for (int k = 0; k < ib.count; k++) awake_idx.push_back(ib.first + k);
    const int n_ext = c->n_contacts, n_int = all_awake ? c->n_ic : (int)awake_idx.size();
    int r = ensure_contact_capacity(c, n_ext + n_int, n_ext);
    if (r) return r;
    DBuf<int> d_awake;
    if (!all_awake) AM_HIP(c, upload(c, d_awake, awake_idx));
    launch_gather_contacts(s, contacts_view(c), n_ext, internal_view(c), all_awake ? nullptr : d_awake.p, 0, n_int, 0);
    c->n_contacts = n_ext + n_int;
    fill_int(s, c->d_scalars.p + 1, 1, c->n_contacts);
    if (c->n_contacts > 0) {
        launch_groups(s, c->n_contacts, c->d_scalars.p + 1, c->c_b1.p, c->c_b2.p, nullptr, nullptr, c->g_flags.p,
                      c->g_scanned.p, c->g_start.p, c->g_a.p, c->g_b.p, c->d_scalars.p + 2, c->d_scan_temp.p);
        AM_HIP(c, hipGetLastError());
        if ((r = read_scalars(c, 2, 1))) return r;
        c->n_groups = c->h_scalars[2];
        c->layout_valid = false;
        // Merging.organizeContacts: the breadth-first order of getOrganizedContacts
        if (P.organize_contacts) {
            if ((r = organized_layers(c, n_ext))) return r;
            c->layout_layer = c->g_layer.p;
        }
        r = do_solve(c, dt, 0, false, iters, nullptr);
        c->layout_layer = nullptr;
        if (r) return r;
        if (all_awake) launch_gather_contacts(s, internal_view(c), 0, contacts_view(c), nullptr, n_ext, n_int, 0);
        else launch_gather_contacts(s, internal_view(c), 0, contacts_view(c), d_awake.p, n_ext, n_int, 3);
        if (!all_awake) AM_HIP(c, hipStreamSynchronize(s));  // d_awake is freed on return
        AM_HIP(c, c->sweep_order.ensure(c->n_contacts));
        AM_HIP(c, hipMemcpyAsync(c->sweep_order.p, c->s_order.p, (size_t)c->n_contacts * 4, hipMemcpyDeviceToDevice,
                                 s));
        c->sweep_order_n = c->n_contacts;
    }
    c->n_contacts = n_ext;
    fill_int(s, c->d_scalars.p + 1, 1, n_ext);
    if ((r = groups_after_contacts(c))) return r;
    launch_advance(s, bodies_view(c), c->d_member_adv.p, (int)c->member_adv.size(), dt, false);
    launch_zero_dv(s, bodies_view(c), c->d_slot_entity.p, (int)c->slot_entity.size());
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

// The advanced poses the position-level metric reads, computed from the
// current state; L.xa stays null at the velocity level.
static int pose_level(am_ctx* c, double dt, PoseLevel* L) {
    L->xa = nullptr;
    L->Ra = nullptr;
    L->dt = dt;
    if (!c->params.metric_position_level) return AM_OK;
    if (!(dt > 0)) return fail(c, AM_ERR_INVALID, "the position-level metric needs dt > 0");
    AM_HIP(c, c->pose_xa.ensure((size_t)std::max(c->n_ent, 1) * 3));
    AM_HIP(c, c->pose_Ra.ensure((size_t)std::max(c->n_ent, 1) * 9));
    launch_advanced_poses(c->stream, bodies_view(c), dt, c->pose_xa.p, c->pose_Ra.p);
    AM_HIP(c, hipGetLastError());
    L->xa = c->pose_xa.p;
    L->Ra = c->pose_Ra.p;
    return AM_OK;
}

int merge_accumulate_unmerging(am_ctx* c, double dt) {
    PoseLevel L;
    int r = pose_level(c, dt, &L);
    if (r || c->ibpc.empty()) return r;
    launch_ibpc_accumulate(c->stream, bodies_view(c), c->d_ibpc.p, (int)c->ibpc.size(), internal_view(c),
                           merge_params(c->params), L);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

// clearBodyPairContacts + the BodyPairContact bookkeeping of
// updateBodyPairContacts (CollisionProcessor.java:145-226): the external pair
// table of this step, with histories carried over from the previous one.
int merge_body_pairs(am_ctx* c) {
    hipStream_t s = c->stream;
    std::swap(c->d_bpc.p, c->d_bpc_prev.p);
    std::swap(c->d_bpc.n, c->d_bpc_prev.n);
    c->n_bpc_prev = c->n_bpc;
    int mask = 0;
    if (c->n_bpc_prev > 0) {
        int ts = 16;
        while (ts < 2 * c->n_bpc_prev) ts *= 2;
        mask = ts - 1;
        AM_HIP(c, c->d_bpc_hash.ensure(ts));
        fill_int(s, c->d_bpc_hash.p, ts, -1);
        launch_bpc_hash(s, c->d_bpc_prev.p, c->n_bpc_prev, c->d_bpc_hash.p, mask);
    }
    c->n_bpc = 0;
    const int n = c->n_contacts;
    if (n == 0) return AM_OK;
    AM_HIP(c, c->d_bpc.ensure(n + 1));
    int* sc = c->d_scalars.p;
    launch_groups(s, n, sc + 1, c->c_b1.p, c->c_b2.p, nullptr, nullptr, c->bp_flags.p, c->bp_scanned.p,
                  c->bp_start.p, nullptr, nullptr, sc + 11, c->d_scan_temp.p);
    launch_bpc_build(s, contacts_view(c), sc + 1, c->bp_start.p, n, c->d_active.p, c->bp_keep.p, c->bp_keep_scan.p,
                     sc + 10, c->d_scan_temp.p, c->d_bpc_prev.p, c->n_bpc_prev, c->d_bpc_hash.p, mask, c->d_bpc.p,
                     c->params.threshold_breath);
    AM_HIP(c, hipGetLastError());
    int r = read_scalars(c, 10, 1);
    if (r) return r;
    c->n_bpc = c->h_scalars[10];
    return AM_OK;
}

int merge_accumulate_merging(am_ctx* c, double dt) {
    PoseLevel L;
    int r = pose_level(c, dt, &L);
    if (r) return r;
    launch_bpc_accumulate(c->stream, bodies_view(c), c->d_bpc.p, c->n_bpc, merge_params(c->params), L);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

namespace {

bool same_collection(const am_ctx* c, int a, int b) { return c->parent[a] >= 0 && c->parent[a] == c->parent[b]; }

// Uploads the internal pair records (after the host changed them).
int upload_ibpc(am_ctx* c) {
    for (auto& r : c->ibpc) r.coll = c->parent[r.b1];
    AM_HIP(c, upload(c, c->d_ibpc, c->ibpc));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

}  // namespace

// Merging.merge (Merging.java:73-163).
int merge_merge(am_ctx* c, int* n_merged) {
    *n_merged = 0;
    if (!c->params.enable_merging || c->n_bpc == 0) return AM_OK;
    hipStream_t s = c->stream;
    std::vector<BpcRec> T(c->n_bpc);
    AM_HIP(c, hipMemcpyAsync(T.data(), c->d_bpc.p, T.size() * sizeof(BpcRec), hipMemcpyDeviceToHost, s));
    AM_HIP(c, hipStreamSynchronize(s));
    // BodyPairContact.checkMergeCondition (:151-177): a pair whose two bodies
    // are both asleep merges without the other tests
    auto asleep = [&](int b) { return (c->flags[b] & AM_BODY_SLEEPING) != 0; };
    // kinematic bodies never sleep, so this test never pairs one
    bool any = false;
    for (const auto& r : T) any = any || r.flag || (asleep(r.b1) && asleep(r.b2));
    if (!any) return AM_OK;
    std::vector<unsigned char> act(c->n_contacts);
    AM_HIP(c, hipMemcpyAsync(act.data(), c->d_active.p, act.size(), hipMemcpyDeviceToHost, s));
    Event E;
    int r = begin_event(c, E);
    if (r) return r;
    // bodyPairContacts of each physical body (external pairs of this step)
    Adjacency adj;
    adj.build(c->n_phys, (int)T.size(), [&](int i) { return T[i].b1; },
              [&](int i) { return T[i].b2 != T[i].b1 ? T[i].b2 : -1; });
    std::vector<char> internal(T.size(), 0);
    std::vector<int> newly;
    auto internalize = [&](int i) {
        internal[i] = 1;
        newly.push_back(i);
    };
    // addIncompleteContacts (RigidCollection.java:1003-1015)
    auto add_incomplete = [&](int body) {
        for (int q = adj.start[body]; q < adj.start[body + 1]; q++) {
            const int i = adj.item[q];
            if (!internal[i] && same_collection(c, T[i].b1, T[i].b2)) internalize(i);
        }
    };
    // state of the collections this event changes (node-based: references
    // stay valid while others are added)
    std::unordered_map<int, Ent> cstate;
    auto coll = [&](int e) -> Ent& {
        auto it = cstate.find(e);
        if (it == cstate.end()) it = cstate.emplace(e, E.load(e)).first;
        return it->second;
    };
    auto save = [&](int e) { E.store_collection(e, coll(e)); };
    for (int i = 0; i < (int)T.size(); i++) {
        const BpcRec& b = T[i];
        if (internal[i]) continue;
        if (!b.flag && !(asleep(b.b1) && asleep(b.b2))) continue;
        if (same_collection(c, b.b1, b.b2)) continue;
        c->merging_event = true;
        (*n_merged)++;
        internalize(i);
        int p1 = c->parent[b.b1], p2 = c->parent[b.b2];
        if (p1 < 0 && p2 < 0) {
            E.top_remove(b.b1);
            E.top_remove(b.b2);
            Ent C;
            int ce = new_collection(E, b.b1, b.b2, C);
            cstate[ce] = C;
            save(ce);
            E.top_append(ce);
        } else if (p1 >= 0 && p2 >= 0) {
            int keep = p1, gone = p2;
            if (!(E.members[p1].size() > E.members[p2].size())) std::swap(keep, gone);
            E.top_remove(gone);
            std::vector<int> gone_bodies = E.members[gone];
            // the absorbed collection's own state is averaged in as one body
            Ent& K = coll(keep);
            add_collection(E, keep, K, gone);
            save(keep);
            for (int body : gone_bodies) add_incomplete(body);
            free_entity(E, gone);
            cstate.erase(gone);
        } else {
            int host = p1 >= 0 ? p1 : p2, body = p1 >= 0 ? b.b2 : b.b1;
            E.top_remove(body);
            Ent& K = coll(host);
            add_body(E, host, K, body);
            save(host);
            add_incomplete(body);
        }
    }
    // RigidCollection.addToInternalContact (:840-850): copies of the active
    // contacts of every pair that became internal, appended in that order
    std::vector<int> idx;
    for (int i : newly) {
        IbpcRec ib;
        ib.b1 = T[i].b1;
        ib.b2 = T[i].b2;
        ib.first = c->n_ic + (int)idx.size();
        ib.count = 0;
        for (int k = T[i].first; k < T[i].end; k++)
            if (act[k]) {
                idx.push_back(k);
                ib.count++;
            }
        ib.hist_n = 0;
        ib.flag = 0;
        ib.coll = -1;
        ib.pad = 0;
        c->ibpc.push_back(ib);
        T[i].b1 = -1;  // no longer an external pair
    }
    if (!idx.empty()) {
        AM_HIP(c, c->ic.ensure(c->n_ic + (int)idx.size(), c->n_ic, s));
        DBuf<int> d_idx;
        AM_HIP(c, upload(c, d_idx, idx));
        launch_gather_contacts(s, internal_view(c), c->n_ic, contacts_view(c), d_idx.p, 0, (int)idx.size(), 1);
        AM_HIP(c, hipStreamSynchronize(s));
        c->n_ic += (int)idx.size();
    }
    AM_HIP(c, hipMemcpyAsync(c->d_bpc.p, T.data(), T.size() * sizeof(BpcRec), hipMemcpyHostToDevice, s));
    if ((r = commit_event(c, E))) return r;
    return upload_ibpc(c);
}

// Merging.unmerge (Merging.java:215-273) and unmergeSelectedBpcs (:286-374).
// Contacts of pairs that leave their collection are appended to the current
// contact list with lambda_warm = lambda.
int merge_unmerge(am_ctx* c, double dt, int* n_events) {
    (void)dt;
    *n_events = 0;
    const am_params& P = c->params;
    if (!P.enable_unmerging || c->ibpc.empty()) return AM_OK;
    if (!P.enable_unmerge_relative_motion && !P.enable_unmerge_normal && !P.enable_unmerge_friction) return AM_OK;
    hipStream_t s = c->stream;
    AM_HIP(c, hipMemcpyAsync(c->ibpc.data(), c->d_ibpc.p, c->ibpc.size() * sizeof(IbpcRec), hipMemcpyDeviceToHost,
                             s));
    AM_HIP(c, hipStreamSynchronize(s));
    bool any = false;
    for (const auto& r : c->ibpc) any = any || r.flag;
    if (!any) return AM_OK;
    Event E;
    int r = begin_event(c, E);
    if (r) return r;
    const int nib = (int)c->ibpc.size();
    std::vector<char> incoll(nib, 1);
    std::vector<int> reexpose;  // internal pairs going back to the external list, in order
    std::vector<int> additions, removals;
    // internal pairs of each body (BodyPairContact lists of the members)
    Adjacency adj;
    adj.build(c->n_phys, nib, [&](int i) { return c->ibpc[i].b1; }, [&](int i) { return c->ibpc[i].b2; });
    // flagged pairs of each collection, in pair order
    std::unordered_map<int, std::vector<int>> cuts;
    for (int i = 0; i < nib; i++)
        if (c->ibpc[i].flag) cuts[c->ibpc[i].coll].push_back(i);
    // per-body marks, cleared after each collection (only its members are set)
    std::vector<char> handled(c->n_phys, 0), remained(c->n_phys, 0), in_sub(c->n_phys, 0);
    // the collections with flagged pairs, in top-level list order
    std::vector<int> order;
    for (const auto& kv : cuts) {
        const int ce = kv.first;
        if (ce >= 0 && ce < (int)E.top_pos.size() && E.top_pos[ce] >= 0 && c->desc[ce].geom == AM_GEOM_COLLECTION &&
            !(c->flags[ce] & AM_BODY_SLEEPING))
            order.push_back(ce);
    }
    std::sort(order.begin(), order.end(), [&](int a, int b) { return E.top_pos[a] < E.top_pos[b]; });
    // removeBodies of the collections that stay, run after the walk: they
    // touch only their own collection and members
    struct Removal {
        int ce;
        Ent C;
        std::vector<int> gone;
    };
    std::vector<Removal> removals_left;
    for (int ce : order) {
        const std::vector<int>& cut = cuts.find(ce)->second;
        for (int i : cut) incoll[i] = 0;
        Ent C = E.load(ce);
        const std::vector<int> bodies = E.members[ce];
        std::vector<int> new_bodies;
        bool remove_coll = true;
        for (int body : bodies) {
            if (handled[body]) continue;
            // buildNeighborBody (:381-393), depth first over pairs still inside
            std::vector<int> sub{body};
            in_sub[body] = 1;
            // recursive order: visit neighbours in pair order, descending immediately
            std::function<void(int)> dfs = [&](int x) {
                for (int q = adj.start[x]; q < adj.start[x + 1]; q++) {
                    const int i = adj.item[q];
                    if (!incoll[i]) continue;
                    int o = c->ibpc[i].b1 == x ? c->ibpc[i].b2 : c->ibpc[i].b1;
                    if (!in_sub[o] && !handled[o]) {
                        in_sub[o] = 1;
                        sub.push_back(o);
                        dfs(o);
                    }
                }
            };
            dfs(body);
            for (int b : sub) {
                handled[b] = 1;
                in_sub[b] = 0;
            }
            if (bodies.size() != sub.size()) {
                if (sub.size() < bodies.size() / 2 + 1) {
                    for (int b : sub) unmerge_body(E, C, b);
                    if (sub.size() > 1) {
                        Ent N;
                        int ne = new_collection(E, sub[0], sub[1], N);
                        std::vector<int> rest(sub.begin() + 2, sub.end());
                        if (!rest.empty()) add_bodies(E, ne, N, rest);
                        E.store_collection(ne, N);
                        new_bodies.push_back(ne);
                    } else {
                        new_bodies.push_back(sub[0]);
                    }
                } else {
                    remove_coll = false;
                    for (int b : sub) remained[b] = 1;
                }
            } else {
                remove_coll = false;
                for (int b : sub) remained[b] = 1;
                break;
            }
        }
        // reconnect (:334-354)
        for (int i : cut) {
            if (same_collection(c, c->ibpc[i].b1, c->ibpc[i].b2)) incoll[i] = 1;
            else reexpose.push_back(i);
        }
        if (!remove_coll) {
            std::vector<int> gone;
            for (int b : bodies)
                if (handled[b] && !remained[b]) gone.push_back(b);
            if (!gone.empty()) removals_left.push_back(Removal{ce, C, std::move(gone)});
        }
        for (int b : bodies) handled[b] = remained[b] = 0;
        if (!new_bodies.empty()) {
            c->merging_event = true;
            (*n_events)++;
            for (int e : new_bodies) additions.push_back(e);
            if (remove_coll) removals.push_back(ce);
        }
    }
    parallel_jobs(c->params.host_threads, (int)removals_left.size(),
                  [&](int k) { return (long)E.members[removals_left[k].ce].size(); },
                  [&](int k) {
                      Removal& R = removals_left[k];
                      remove_bodies(E, R.ce, R.C, R.gone);
                      E.store_collection(R.ce, R.C);
                  });
    for (int e : additions) E.top_append(e);
    for (int e : removals) {
        E.top_remove(e);
        free_entity(E, e);
    }
    // pairs whose bodies no longer share a collection go back to the contact list
    std::vector<int> ext_idx;
    for (int i : reexpose)
        for (int k = 0; k < c->ibpc[i].count; k++) ext_idx.push_back(c->ibpc[i].first + k);
    if (!ext_idx.empty()) {
        const int n0 = c->n_contacts, m = (int)ext_idx.size();
        if ((r = ensure_contact_capacity(c, n0 + m, n0))) return r;
        DBuf<int> d_idx;
        AM_HIP(c, upload(c, d_idx, ext_idx));
        launch_gather_contacts(s, contacts_view(c), n0, internal_view(c), d_idx.p, 0, m, 2);
        AM_HIP(c, hipStreamSynchronize(s));
        c->n_contacts = n0 + m;
        fill_int(s, c->d_scalars.p + 1, 1, c->n_contacts);
    }
    // compact the internal pairs still inside a collection
    std::vector<IbpcRec> kept;
    std::vector<int> keep_idx;
    for (int i = 0; i < nib; i++) {
        if (!incoll[i]) continue;
        IbpcRec ib = c->ibpc[i];
        int first = (int)keep_idx.size();
        for (int k = 0; k < ib.count; k++) keep_idx.push_back(ib.first + k);
        ib.first = first;
        kept.push_back(ib);
    }
    if ((int)keep_idx.size() != c->n_ic) {
        AM_HIP(c, c->ic_tmp.ensure((int)keep_idx.size(), 0, s));
        if (!keep_idx.empty()) {
            DBuf<int> d_idx;
            AM_HIP(c, upload(c, d_idx, keep_idx));
            launch_gather_contacts(s, c->ic_tmp.view(), 0, internal_view(c), d_idx.p, 0, (int)keep_idx.size(), 0);
            AM_HIP(c, hipStreamSynchronize(s));
        }
        c->ic.swap(c->ic_tmp);
        c->n_ic = (int)keep_idx.size();
    }
    c->ibpc = kept;
    if ((r = commit_event(c, E))) return r;
    if ((r = upload_ibpc(c))) return r;
    return groups_after_contacts(c);
}

// Merging.unmergeAll (Merging.java:168-206), the "unmerge all" request: every
// collection that is awake hands its bodies back to the top-level list (after
// the bodies that stay, in collection order) with the collection's velocity
// at their position.  Its body pairs become external with an empty motion
// history; their contacts are dropped (the reference's re-expose branch runs
// only while inCollection is still set, which it no longer is).
int merge_unmerge_all(am_ctx* c, int* n_dissolved) {
    *n_dissolved = 0;
    std::vector<int> colls;
    for (int e : c->slot_entity)
        if (c->desc[e].geom == AM_GEOM_COLLECTION && !(c->flags[e] & AM_BODY_SLEEPING)) colls.push_back(e);
    if (colls.empty()) return AM_OK;
    hipStream_t s = c->stream;
    Event E;
    int r = begin_event(c, E);
    if (r) return r;
    std::vector<char> gone(c->n_ent, 0);
    for (int ce : colls) gone[ce] = 1;
    // external pairs touching those collections restart their motion history
    if (c->n_bpc > 0) {
        std::vector<BpcRec> T(c->n_bpc);
        AM_HIP(c, hipMemcpyAsync(T.data(), c->d_bpc.p, T.size() * sizeof(BpcRec), hipMemcpyDeviceToHost, s));
        AM_HIP(c, hipStreamSynchronize(s));
        for (auto& t : T)
            if (t.b1 >= 0 && ((c->parent[t.b1] >= 0 && gone[c->parent[t.b1]]) ||
                              (c->parent[t.b2] >= 0 && gone[c->parent[t.b2]])))
                t.hist_n = 0;
        AM_HIP(c, hipMemcpyAsync(c->d_bpc.p, T.data(), T.size() * sizeof(BpcRec), hipMemcpyHostToDevice, s));
    }
    // internal pairs of the dissolved collections leave with their contacts
    std::vector<IbpcRec> kept;
    std::vector<int> keep_idx;
    for (const auto& ib : c->ibpc) {
        if (gone[ib.coll]) continue;
        IbpcRec k = ib;
        k.first = (int)keep_idx.size();
        for (int q = 0; q < ib.count; q++) keep_idx.push_back(ib.first + q);
        kept.push_back(k);
    }
    if ((int)keep_idx.size() != c->n_ic) {
        AM_HIP(c, c->ic_tmp.ensure((int)keep_idx.size(), 0, s));
        if (!keep_idx.empty()) {
            DBuf<int> d_idx;
            AM_HIP(c, upload(c, d_idx, keep_idx));
            launch_gather_contacts(s, c->ic_tmp.view(), 0, internal_view(c), d_idx.p, 0, (int)keep_idx.size(), 0);
            AM_HIP(c, hipStreamSynchronize(s));
        }
        c->ic.swap(c->ic_tmp);
        c->n_ic = (int)keep_idx.size();
    }
    c->ibpc = kept;
    std::vector<int> additions;
    for (int ce : colls) {
        Ent C = E.load(ce);
        for (int b : E.members[ce]) {
            unmerge_body(E, C, b);
            additions.push_back(b);
        }
    }
    for (int ce : colls) {
        E.top_remove(ce);
        free_entity(E, ce);
    }
    for (int b : additions) E.top_append(b);
    *n_dissolved = (int)colls.size();
    if ((r = commit_event(c, E))) return r;
    return upload_ibpc(c);
}

// ---------------------------------------------------------------------------
// Sleeping (Sleeping.java).  The flags live in d_flags (authoritative during
// a step) and in the host mirror c->flags; every change made on the device is
// pulled back and the derived lists (broad-phase skips, integration targets,
// members the single sweep advances) are rebuilt.

// Metric histories for every entity: kSleepHistStride doubles and a length.
int ensure_sleep_state(am_ctx* c) {
    hipStream_t s = c->stream;
    const int n = c->n_ent;
    if (c->sleep_cap < n) {
        const int keep = c->sleep_cap;
        AM_HIP(c, grow_keep(s, c->d_sleep_hist, (size_t)n, (size_t)kSleepHistStride, (size_t)keep));
        AM_HIP(c, grow_keep(s, c->d_sleep_n, (size_t)n, 1, (size_t)keep));
        AM_HIP(c, c->d_sleep_ext.ensure(n));
        fill_int(s, c->d_sleep_n.p + keep, n - keep, 0);
        fill_double(s, c->d_sleep_hist.p + (size_t)keep * kSleepHistStride, (n - keep) * kSleepHistStride, 0.0);
        c->sleep_cap = n;
    }
    if (!c->sleep_fresh.empty()) {
        std::vector<int> fresh;
        for (int e : c->sleep_fresh)
            if (e < n) fresh.push_back(e);
        c->sleep_fresh.clear();
        if (!fresh.empty()) {
            std::vector<int> hn(n);
            int r = pull_i(c, hn.data(), c->d_sleep_n.p, n);
            if (r) return r;
            AM_HIP(c, hipStreamSynchronize(s));
            for (int e : fresh) hn[e] = 0;
            AM_HIP(c, hipMemcpyAsync(c->d_sleep_n.p, hn.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
            AM_HIP(c, hipStreamSynchronize(s));
        }
    }
    return AM_OK;
}

namespace {

// RigidBody.wake (RigidBody.java:298-306): clears the flag (and the history)
// of the body, then of its collection.
void wake_flag(am_ctx* c, int e, std::vector<char>& woke) {
    if (c->flags[e] & AM_BODY_SLEEPING) {
        c->flags[e] &= ~AM_BODY_SLEEPING;
        c->desc[e].flags &= ~AM_BODY_SLEEPING;
        woke[e] = 1;
    }
}
void wake_one(am_ctx* c, int b, std::vector<char>& woke) {
    wake_flag(c, b, woke);
    if (c->parent[b] >= 0) wake_flag(c, c->parent[b], woke);
}

// Uploads flags and clears the histories of the woken entities.
int commit_wake(am_ctx* c, const std::vector<char>& woke, int* n_woken) {
    int n = 0;
    for (char w : woke) n += w != 0;
    if (n_woken) *n_woken = n;
    if (n == 0) return AM_OK;
    int r = ensure_sleep_state(c);
    if (r) return r;
    std::vector<int> hn(c->n_ent);
    if ((r = pull_i(c, hn.data(), c->d_sleep_n.p, c->n_ent))) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    for (int e = 0; e < c->n_ent; e++)
        if (woke[e]) hn[e] = 0;
    AM_HIP(c, hipMemcpyAsync(c->d_sleep_n.p, hn.data(), (size_t)c->n_ent * 4, hipMemcpyHostToDevice, c->stream));
    return rebuild_structure(c);
}

bool any_top_asleep(const am_ctx* c) {
    for (int e : c->slot_entity)
        if (c->flags[e] & AM_BODY_SLEEPING) return true;
    return false;
}

}  // namespace

// Sleeping.wake (Sleeping.java:107-139), after the warm start.
int sleep_wake(am_ctx* c, int* n_woken) {
    if (n_woken) *n_woken = 0;
    if (!c->params.enable_sleeping || !any_top_asleep(c)) return AM_OK;
    hipStream_t s = c->stream;
    std::vector<char> woke(c->n_ent, 0);
    const int n = c->n_contacts;
    if (n > 0) {
        fill_int(s, c->d_scalars.p + 20, 1, 0);
        launch_wake_count(s, bodies_view(c), c->c_b1.p, c->c_b2.p, n, c->d_scalars.p + 20);
        AM_HIP(c, hipGetLastError());
        int r = read_scalars(c, 20, 1);
        if (r) return r;
        if (c->h_scalars[20] > 0) {
            std::vector<int> b1(n), b2(n);
            if ((r = pull_i(c, b1.data(), c->c_b1.p, n))) return r;
            if ((r = pull_i(c, b2.data(), c->c_b2.p, n))) return r;
            AM_HIP(c, hipStreamSynchronize(s));
            // bodyPairContacts of every top-level body: the pairs of this step
            // between bodies that are not pinned
            auto top = [&](int b) { return c->parent[b] >= 0 ? c->parent[b] : b; };
            std::vector<std::vector<int>> pairs(c->n_ent);
            for (int k = 0; k < n; k++) {
                if ((c->flags[b1[k]] | c->flags[b2[k]]) & AM_BODY_PINNED) continue;
                pairs[top(b1[k])].push_back(k);
                if (top(b2[k]) != top(b1[k])) pairs[top(b2[k])].push_back(k);
            }
            const std::vector<int> order = c->slot_entity;
            for (int e : order) {
                if (!(c->flags[e] & AM_BODY_SLEEPING)) continue;
                for (int k : pairs[e]) {
                    wake_one(c, b1[k], woke);
                    wake_one(c, b2[k], woke);
                }
            }
        }
    }
    // body-body springs with one end asleep and neither end pinned
    for (const am_spring& sp : c->springs_host) {
        if (sp.type != AM_SPRING_BODYBODY) continue;
        const int t1 = c->parent[sp.body1] >= 0 ? c->parent[sp.body1] : sp.body1;
        const int t2 = c->parent[sp.body2] >= 0 ? c->parent[sp.body2] : sp.body2;
        const bool s1 = (c->flags[t1] & AM_BODY_SLEEPING) != 0, s2 = (c->flags[t2] & AM_BODY_SLEEPING) != 0;
        const bool p1 = (c->flags[t1] & AM_BODY_PINNED) != 0, p2 = (c->flags[t2] & AM_BODY_PINNED) != 0;
        if (s1 != s2 && !(p1 || p2)) {
            wake_one(c, sp.body1, woke);
            wake_one(c, sp.body2, woke);
        }
    }
    return commit_wake(c, woke, n_woken);
}

// Sleeping.wakeAll (Sleeping.java:144-150): every top-level body.
int sleep_wake_all(am_ctx* c) {
    std::vector<char> woke(c->n_ent, 0);
    for (int e : std::vector<int>(c->slot_entity)) wake_one(c, e, woke);
    return commit_wake(c, woke, nullptr);
}

// Sleeping.sleep (Sleeping.java:49-98), at the end of the step.
int sleep_sleep(am_ctx* c, int* n_asleep) {
    if (!c->params.enable_sleeping) {
        if (n_asleep) {
            int k = 0;
            for (int e : c->slot_entity) k += (c->flags[e] & AM_BODY_SLEEPING) != 0;
            *n_asleep = k;
        }
        return AM_OK;
    }
    int r = ensure_sleep_state(c);
    if (r) return r;
    hipStream_t s = c->stream;
    fill_int(s, c->d_sleep_ext.p, c->n_ent, 0);
    launch_sleep_external(s, bodies_view(c), c->c_b1.p, c->c_b2.p, c->c_lam.p, c->n_contacts, c->d_sleep_ext.p);
    fill_int(s, c->d_scalars.p + 21, 2, 0);
    launch_sleep_decide(s, bodies_view(c), c->d_flags.p, c->d_slot_entity.p, (int)c->slot_entity.size(),
                        c->d_sleep_ext.p, c->d_sleep_hist.p, c->d_sleep_n.p, c->params.sleep_step_accum,
                        c->params.sleep_threshold, c->d_scalars.p + 21);
    AM_HIP(c, hipGetLastError());
    if ((r = read_scalars(c, 21, 2))) return r;
    if (n_asleep) *n_asleep = c->h_scalars[22];
    if (c->h_scalars[21] > 0) {
        if ((r = pull_i(c, c->flags.data(), c->d_flags.p, c->n_ent))) return r;
        AM_HIP(c, hipStreamSynchronize(s));
        for (int e = 0; e < c->n_ent; e++)
            c->desc[e].flags = (c->desc[e].flags & ~AM_BODY_SLEEPING) | (c->flags[e] & AM_BODY_SLEEPING);
        return rebuild_structure(c);
    }
    return AM_OK;
}

}  // namespace am_rt
