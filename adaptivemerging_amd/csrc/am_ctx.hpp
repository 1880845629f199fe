// am_ctx.hpp - the context behind the C ABI: host mirrors, device buffers
// and the host-side helpers shared by the runtime's translation units.
#pragma once

#include <algorithm>
#include <new>
#include <cstring>
#include <string>
#include <vector>

#include "am_merge.hpp"
#include "am_mlp.hpp"
#include "am_solve.hpp"

namespace am_rt {

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Grows (never shrinks); contents are not preserved.
    hipError_t ensure(size_t want) {
        if (want <= n && p) return hipSuccess;
        release();
        size_t cap = std::max<size_t>(want, 1);
        hipError_t e = hipMalloc((void**)&p, cap * sizeof(T));
        if (e == hipSuccess) n = cap;
        else p = nullptr;
        return e;
    }
};


// Page-locked host memory: device copies to and from it run at full speed.
template <class T>
struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <class U>
    PinnedAlloc(const PinnedAlloc<U>&) {}
    T* allocate(size_t n) {
        void* p = nullptr;
        if (hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess)
            throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { (void)hipHostFree(p); }
    template <class U>
    bool operator==(const PinnedAlloc<U>&) const { return true; }
    template <class U>
    bool operator!=(const PinnedAlloc<U>&) const { return false; }
};
template <class T>
using PinnedVec = std::vector<T, PinnedAlloc<T>>;

// Grows b to cap elements of `stride` values, keeping the first `keep` elements.
template <class T>
hipError_t grow_keep(hipStream_t s, DBuf<T>& b, size_t cap, size_t stride, size_t keep) {
    if (keep == 0 || !b.p) return b.ensure(cap * stride);
    if (b.n >= cap * stride) return hipSuccess;
    DBuf<T> nb;
    hipError_t e = nb.ensure(cap * stride);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(nb.p, b.p, keep * stride * sizeof(T), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    std::swap(nb.p, b.p);
    std::swap(nb.n, b.n);
    return hipSuccess;
}

// A set of contact arrays (the fields of ContactsView) with its own capacity.
struct ContactStore {
    int cap = 0;
    DBuf<int> b1, b2, bv1, bv2, info, state, is_new;
    DBuf<double> pB1, nB1, t1B1, t2B1, viol, prev, J, lam, lamw, b, D;
    hipError_t ensure(int want, int keep, hipStream_t s);
    void swap(ContactStore& o);
    am::ContactsView view() {
        am::ContactsView C;
        C.cap = cap;
        C.b1 = b1.p;
        C.b2 = b2.p;
        C.bv1 = bv1.p;
        C.bv2 = bv2.p;
        C.info = info.p;
        C.pB1 = pB1.p;
        C.nB1 = nB1.p;
        C.t1B1 = t1B1.p;
        C.t2B1 = t2B1.p;
        C.viol = viol.p;
        C.prev_viol = prev.p;
        C.J = J.p;
        C.lam = lam.p;
        C.lamw = lamw.p;
        C.b = b.p;
        C.D = D.p;
        C.state = state.p;
        C.is_new = is_new.p;
        return C;
    }
};

}  // namespace am_rt

struct am_ctx {
    template <class T>
    using DBuf = am_rt::DBuf<T>;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    am_params params{};

    // scene (host mirrors)
    int n_ent = 0, n_phys = 0, n_nodes = 0;
    std::vector<am_body_desc> desc;
    std::vector<am_sphere_node> nodes;
    std::vector<am_body_desc> subs;     // sub-bodies of composites
    std::vector<double> sub_b2c;        // 12 per sub-body
    std::vector<int> node_sub;          // per node: sub-body (inside its composite) owning it, -1 otherwise
    std::vector<int> parent, slot, sub, slot_entity, member_first, member_count, members, flags;
    std::vector<int> planes, grid_bodies, node_body, top_pinned, top_sleeping, gravity_targets, integrate_targets;
    std::vector<int> plane_start;  // per scene: first of its planes in `planes` (planes grouped by scene)
    DBuf<int> d_plane_start;
    std::vector<int> grid_is_large, grid_large;  // per gridded body: too large for the grid; their gridded indices
    double cell_size = 1.0;
    int hash_mask = 0;

    // body buffers
    DBuf<am_body_desc> d_desc, d_subs;
    DBuf<double> d_sub_b2c;
    DBuf<int> d_node_sub;
    DBuf<am_sphere_node> d_nodes;
    DBuf<double> d_cw, d_x, d_R, d_v, d_w, d_f, d_t, d_jinv, d_dv, d_minv;
    DBuf<int> d_flags, d_parent, d_slot, d_sub, d_slot_entity, d_member_first, d_members, d_planes, d_grid,
        d_node_body, d_top_pinned, d_top_sleeping, d_gravity, d_integrate, d_grid_is_large, d_grid_large;

    // broad phase
    DBuf<double> d_sphere;
    DBuf<int> d_cell, d_bucket_count, d_bucket_start, d_bucket_cursor, d_bucket_items, d_pair_counts, d_pair_offsets;
    DBuf<uint64_t> d_keys_unsorted, d_keys;
    DBuf<unsigned char> d_sort_temp, d_scan_temp;
    size_t sort_temp_size = 0;

    // contacts
    int cap_c = 0;
    DBuf<int> c_b1, c_b2, c_bv1, c_bv2, c_info, c_state, c_new;
    DBuf<double> c_pB1, c_nB1, c_t1B1, c_t2B1, c_viol, c_prev, c_J, c_lam, c_lamw, c_b, c_D;
    DBuf<int> d_narrow_counts, d_narrow_offsets;

    // old contacts (warm start)
    int n_old = 0, table_mask = 0;
    DBuf<int> o_b1, o_b2, o_bv1, o_bv2, o_info, o_table;
    DBuf<double> o_pB1, o_lam, o_viol, o_prev;

    // groups, colouring, solve layout
    DBuf<int> g_flags, g_scanned, g_start, g_a, g_b, g_colour_a, g_colour_b, inc_count, inc_start, inc_cursor, inc_items,
        g_sizes, g_pos_start, s_order, s_colour_start;
    DBuf<uint64_t> g_keys, g_keys_sorted;
    DBuf<double> s_rec, s_lam;
    DBuf<int> s_meta, s_state, d_barrier, s_group_desc, s_pos_a;
    DBuf<unsigned long long> s_iter_max;
    // dataflow schedule: sorted position per group, ranks and degrees, previous-update
    // incidences, pass completion counts, deltaV slots, odd-pass buffers
    DBuf<int> g_spos, s_flow, s_pred;
    DBuf<unsigned long long> d_pass_done;
    DBuf<double> f_slot;
    DBuf<double> dv_alt, s_lam_alt;
    int flow_max_grid = 0;
    // poses advanced by dt for the position-level motion metric
    DBuf<double> pose_xa, pose_Ra;
    // springs (RigidBodySystem.springs)
    DBuf<am_spring> d_springs;
    std::vector<am_spring> springs_host;
    int n_springs = 0;
    // prescribed motions, forces held for the next external-force pass, simulation time
    std::vector<am_kinematic> kin_host;
    DBuf<am_kinematic> d_kin;
    std::vector<am::ImpulseRec> impulses;
    DBuf<am::ImpulseRec> d_imp;
    double sim_time = 0.0;
    // independent scenes (am_body_desc.scene): count, and the per-scene early-exit state of the last main solve
    int n_scenes = 1;
    bool scene_solve_valid = false;
    int scene_iters = 1;  // sweep cap of that solve (the stride of scene_max)
    DBuf<int> g_scene, scene_used;
    DBuf<unsigned long long> scene_done;
    DBuf<unsigned> scene_groups;
    DBuf<unsigned long long> scene_max;
    // sleeping (Sleeping.java): per-entity metric histories
    DBuf<double> d_sleep_hist;          // kSleepHistStride per entity, oldest first
    DBuf<int> d_sleep_n, d_sleep_ext;   // history length; active external contact this step
    int sleep_cap = 0;                  // entities the sleep buffers hold
    std::vector<int> sleep_fresh;       // new collection entities whose history must start empty

    // scalars on device: 0 n_pairs, 1 n_contacts, 2 n_groups, 3 total scratch, 4 overflow, 5 remaining,
    // 6 iters_used, 7 bad, 8 repaired, 9 n_sweep
    DBuf<int> d_scalars;
    int* h_scalars = nullptr;  // pinned mirror

    int n_pairs = 0, n_contacts = 0, n_groups = 0, n_colours = 0, max_colour_groups = 0;
    bool layout_valid = false;
    int last_iters = 0;
    double last_change = 0.0;
    std::vector<double> changes;
    int pgs_max_grid[5] = {0, 0, 0, 0, 0};  // co-resident workgroups of the coloured kernel, by block / 64
    hipEvent_t kev[3] = {nullptr, nullptr, nullptr};  // setup start, pgs start, pgs end
    double ms_pgs_kernel = 0.0, ms_setup_kernel = 0.0;
    int last_grid = 0, last_block = 0;  // launch shape of the last main solve's PGS kernel

    // learned warm start
    bool mlp_ready = false;
    DBuf<float> m_p, m_g, m_m, m_v;  // padded FP32 parameters, gradients, Adam moments
    DBuf<unsigned short> m_bf;       // padded BF16 weights
    DBuf<float> f_X, f_Y, d_loss;
    DBuf<double> f_scale, f_lam, s_start, dv_save;
    DBuf<unsigned long long> d_resid;
    DBuf<float> t_X, t_T;
    DBuf<int> t_idx;
    int t_n = 0, adam_step = 0;
    double ms_learned = 0.0, res_plain = 0.0, res_learned = 0.0;
    int used_learned = 0;
    hipEvent_t lev[2] = {nullptr, nullptr};

    // adaptive merging (Merging, RigidCollection, BodyPairContact)
    int total_steps = 0;
    bool merging_event = false;         // Merging.mergingEvent: set by the first merge or unmerge
    std::vector<int> free_entities;     // collection entity ids not in use
    std::vector<double> b2c;            // per entity: body-to-collection transform, R (9) then t (3)
    DBuf<double> d_b2c;
    DBuf<int> ev_list;                  // staging of an event's changed entities
    DBuf<double> ev_rec;
    // host mirrors of the entity state kept between events (an event takes
    // them and gives them back when it commits)
    am_rt::PinnedVec<double> ev_x, ev_R, ev_v, ev_w, ev_jinv;
    std::vector<std::vector<int>> ev_members;
    std::vector<int> member_list;       // every collection member (physical body)
    std::vector<int> member_adv;        // members that the single sweep integrates (not pinned)
    DBuf<int> d_member_list, d_member_adv;
    // external body-pair table, rebuilt after every main solve
    int n_bpc = 0, n_bpc_prev = 0, bpc_mask = 0;
    DBuf<am::BpcRec> d_bpc, d_bpc_prev;
    DBuf<int> d_bpc_hash, bp_flags, bp_scanned, bp_start, bp_keep, bp_keep_scan;
    DBuf<unsigned char> d_active;
    // internal body pairs and their contacts (RigidCollection.internalContacts)
    std::vector<am::IbpcRec> ibpc;
    DBuf<am::IbpcRec> d_ibpc;
    int n_ic = 0;
    am_rt::ContactStore ic, ic_tmp;
    int last_merged = 0, last_unmerged = 0;
    // breadth-first layers of the single sweep's groups (getOrganizedContacts)
    DBuf<int> g_layer, bfs_mark;
    DBuf<int> g_blocker;  // colouring: the neighbour a waiting group waits for
    const int* layout_layer = nullptr;  // set while the single sweep builds its layout
    DBuf<int> sweep_order;              // order of the last single sweep inside collections
    int sweep_order_n = -1;             // -1: none since the last solve or detection
    hipEvent_t sev[9] = {};               // am_step phase events
    double ms_merging = 0.0;
};

namespace am_rt {

int fail(am_ctx* c, int code, const std::string& m);
int apply_springs(am_ctx* c);
int apply_external_forces(am_ctx* c);
int kinematic_update(am_ctx* c, double t);
int apply_impulses(am_ctx* c);
int apply_viscous_decay(am_ctx* c);
int check_ctx(am_ctx* c);
am::BodiesView bodies_view(am_ctx* c);
am::TreeView tree_view(am_ctx* c);
am::ContactsView contacts_view(am_ctx* c);
am::ContactsView internal_view(am_ctx* c);
am::MergeParams merge_params(const am_params& p);
int read_scalars(am_ctx* c, int first, int n);
int ensure_contact_capacity(am_ctx* c, int want, int keep = 0);
// bodies_changed = false: only the collection structure changed (an event),
// so the lists derived from the physical bodies' geometry, the broad-phase
// grid and the member transforms (written by the event) are left as they are.
int rebuild_structure(am_ctx* c, bool bodies_changed = true);
int refresh_jinv(am_ctx* c);
int groups_after_contacts(am_ctx* c);
int do_solve(am_ctx* c, double dt, int post, bool main_solve, int* iters, double* last);
int apply_gravity(am_ctx* c);
int pull_d(am_ctx* c, double* dst, const double* src, size_t n);
int pull_i(am_ctx* c, int32_t* dst, const int* src, size_t n);

template <class T>
hipError_t upload(am_ctx* c, DBuf<T>& b, const std::vector<T>& v) {
    hipError_t e = b.ensure(std::max<size_t>(v.size(), 1));
    if (e != hipSuccess) return e;
    if (!v.empty()) return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream);
    return hipSuccess;
}

// merging phases (am_merge.hip)
int merge_update_in_collections(am_ctx* c, double dt, int* iters);
int merge_accumulate_unmerging(am_ctx* c, double dt);
int merge_unmerge(am_ctx* c, double dt, int* n_events);
int merge_body_pairs(am_ctx* c);
int merge_accumulate_merging(am_ctx* c, double dt);
int merge_merge(am_ctx* c, int* n_merged);
int merge_unmerge_all(am_ctx* c, int* n_dissolved);
int merge_push_down(am_ctx* c, bool integrated_only);
int merge_member_transforms(am_ctx* c);
int pull_contacts(am_ctx* c, const am::ContactsView& V, int count, am_contact_arrays* o, int32_t capacity);
bool has_collections(const am_ctx* c);
// Physical bodies hold ids [0, n_phys) (uploads never hold collections);
// collection entities are allocated after them, so a listed entity is a
// collection exactly when its id is n_phys or more.  Loops over every
// top-level entity use this instead of reading each descriptor.
inline bool is_collection(const am_ctx* c, int e) { return e >= c->n_phys; }
// sleeping phases (am_merge.hip)
int ensure_sleep_state(am_ctx* c);
int sleep_wake(am_ctx* c, int* n_woken);
int sleep_sleep(am_ctx* c, int* n_asleep);
int sleep_wake_all(am_ctx* c);

}  // namespace am_rt

#define AM_HIP(ctx, expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return am_rt::fail(ctx, AM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
