// am_api.hip - host runtime behind the C ABI of include/am.h: context,
// device buffers, the per-step orchestration of the kernels on one HIP stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "am_ctx.hpp"

using namespace am;
using namespace am_rt;

namespace am_rt {

int fail(am_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

BodiesView bodies_view(am_ctx* c) {
    BodiesView B;
    B.n = c->n_ent;
    B.n_phys = c->n_phys;
    B.desc = c->d_desc.p;
    B.x = c->d_x.p;
    B.R = c->d_R.p;
    B.v = c->d_v.p;
    B.w = c->d_w.p;
    B.f = c->d_f.p;
    B.t = c->d_t.p;
    B.jinv = c->d_jinv.p;
    B.dv = c->d_dv.p;
    B.minv = c->d_minv.p;
    B.flags = c->d_flags.p;
    B.parent = c->d_parent.p;
    B.slot = c->d_slot.p;
    B.sub = c->d_sub.p;
    B.sub_desc = c->d_subs.p;
    B.sub_b2c = c->d_sub_b2c.p;
    return B;
}
TreeView tree_view(am_ctx* c) { return TreeView{c->d_nodes.p, c->d_cw.p}; }
ContactsView contacts_view(am_ctx* c) {
    ContactsView C;
    C.cap = c->cap_c;
    C.b1 = c->c_b1.p;
    C.b2 = c->c_b2.p;
    C.bv1 = c->c_bv1.p;
    C.bv2 = c->c_bv2.p;
    C.info = c->c_info.p;
    C.pB1 = c->c_pB1.p;
    C.nB1 = c->c_nB1.p;
    C.t1B1 = c->c_t1B1.p;
    C.t2B1 = c->c_t2B1.p;
    C.viol = c->c_viol.p;
    C.prev_viol = c->c_prev.p;
    C.J = c->c_J.p;
    C.lam = c->c_lam.p;
    C.lamw = c->c_lamw.p;
    C.b = c->c_b.p;
    C.D = c->c_D.p;
    C.state = c->c_state.p;
    C.is_new = c->c_new.p;
    return C;
}
ContactsView internal_view(am_ctx* c) { return c->ic.view(); }
OldContactsView old_view(am_ctx* c) {
    OldContactsView O;
    O.n = c->n_old;
    O.b1 = c->o_b1.p;
    O.b2 = c->o_b2.p;
    O.bv1 = c->o_bv1.p;
    O.bv2 = c->o_bv2.p;
    O.info = c->o_info.p;
    O.pB1 = c->o_pB1.p;
    O.lam = c->o_lam.p;
    O.viol = c->o_viol.p;
    O.prev_viol = c->o_prev.p;
    O.table = c->o_table.p;
    O.table_mask = c->table_mask;
    return O;
}


// Reads n scalars from d_scalars[first..] (synchronises the stream).
int read_scalars(am_ctx* c, int first, int n) {
    AM_HIP(c, hipMemcpyAsync(c->h_scalars + first, c->d_scalars.p + first, n * sizeof(int), hipMemcpyDeviceToHost,
                             c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int ensure_scan_temp(am_ctx* c, int n) {
    size_t want = scan_temp_bytes(n) + 1024;
    AM_HIP(c, c->d_scan_temp.ensure(std::max(want, c->d_scan_temp.n)));
    return AM_OK;
}

// Capacity for `want` contacts; the first `keep` contacts survive a regrow.
int ensure_contact_capacity(am_ctx* c, int want, int keep) {
    if (want <= c->cap_c && c->c_b1.p) return AM_OK;
    int cap = std::max(want + want / 4, 1024);
    size_t k = (size_t)std::min(keep, c->cap_c);
    AM_HIP(c, grow_keep(c->stream, c->c_b1, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_b2, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_bv1, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_bv2, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_info, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_state, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_new, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_pB1, cap, 3, k));
    AM_HIP(c, grow_keep(c->stream, c->c_nB1, cap, 3, k));
    AM_HIP(c, grow_keep(c->stream, c->c_t1B1, cap, 3, k));
    AM_HIP(c, grow_keep(c->stream, c->c_t2B1, cap, 3, k));
    AM_HIP(c, grow_keep(c->stream, c->c_viol, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_prev, cap, 1, k));
    AM_HIP(c, grow_keep(c->stream, c->c_J, cap, 36, k));
    AM_HIP(c, grow_keep(c->stream, c->c_lam, cap, 3, k));
    AM_HIP(c, grow_keep(c->stream, c->c_lamw, cap, 3, k));
    AM_HIP(c, grow_keep(c->stream, c->c_b, cap, 3, k));
    AM_HIP(c, grow_keep(c->stream, c->c_D, cap, 3, k));
    AM_HIP(c, c->bp_flags.ensure(cap));
    AM_HIP(c, c->bp_scanned.ensure(cap));
    AM_HIP(c, c->bp_start.ensure(cap + 1));
    AM_HIP(c, c->bp_keep.ensure(cap + 1));
    AM_HIP(c, c->bp_keep_scan.ensure(cap + 1));
    AM_HIP(c, c->d_active.ensure(cap));
    // groups and solve layout are per contact too
    AM_HIP(c, c->g_flags.ensure(cap));
    AM_HIP(c, c->g_scanned.ensure(cap));
    AM_HIP(c, c->g_start.ensure(cap + 1));
    AM_HIP(c, c->g_a.ensure(cap));
    AM_HIP(c, c->g_b.ensure(cap));
    AM_HIP(c, c->g_colour_a.ensure(cap));
    AM_HIP(c, c->g_colour_b.ensure(cap));
    AM_HIP(c, c->g_keys.ensure(cap));
    AM_HIP(c, c->g_keys_sorted.ensure(cap));
    AM_HIP(c, c->g_spos.ensure(cap));
    AM_HIP(c, c->g_sizes.ensure(cap + 1));
    AM_HIP(c, c->g_pos_start.ensure(cap + 1));
    AM_HIP(c, c->s_order.ensure(cap));
    AM_HIP(c, c->s_pos_a.ensure(cap));
    AM_HIP(c, c->s_rec.ensure((size_t)cap * kRecDoubles));
    AM_HIP(c, c->s_lam.ensure((size_t)cap * 3));
    AM_HIP(c, c->s_meta.ensure((size_t)cap * 4));
    AM_HIP(c, c->s_state.ensure(cap));
    size_t st = sort_temp_bytes(cap);
    if (st > c->d_sort_temp.n) AM_HIP(c, c->d_sort_temp.ensure(st));
    int r = ensure_scan_temp(c, cap + 1);
    if (r) return r;
    c->cap_c = cap;
    return AM_OK;
}

void default_params(am_params* p) {
    std::memset(p, 0, sizeof *p);
    p->iterations = 30;
    p->iterations_in_collection = 1;
    p->tolerance = 1e-5;
    p->omega = 1.0;
    p->feedback_stiffness = 0.5;
    p->enable_post_stabilization = 0;
    p->enable_compliance = 1;
    p->compliance = 1e-3;
    p->restitution_override = 0;
    p->friction_override = 0;
    p->restitution = 0.5;
    p->friction = 0.1;
    p->sliding_threshold = 0.01;
    p->warm_start = 1;
    p->use_gravity = 1;
    p->gravity_amount = 1.0;
    p->gravity_angle_deg = 90.0;
    p->viscous_linear_decay = 1.0;
    p->viscous_angular_decay = 1.0;
    p->spring_stiffness_mod = 1.0;
    p->spring_damping_mod = 1.0;
    p->enable_sleeping = 1;
    p->sleep_step_accum = 10;
    p->sleep_threshold = 1e-5;
    p->enable_merging = 1;
    p->enable_merge_pinned = 1;
    p->enable_merge_stable_contact = 1;
    p->enable_merge_let_it_breathe = 1;
    p->enable_unmerging = 1;
    p->enable_unmerge_friction = 1;
    p->enable_unmerge_normal = 1;
    p->enable_unmerge_relative_motion = 1;
    p->update_contacts_in_collections = 1;
    p->step_accum_merging = 3;
    p->step_accum_unmerging = 3;
    p->steps_between_merge_events = 10;
    p->metric_position_level = 0;
    p->threshold_merge = 1e-2;
    p->threshold_unmerge = 2e-2;
    p->threshold_breath = 1e-5;
    p->pgs_block = 0;
    p->pgs_grid = 0;
    p->pgs_plain_launch = 0;
    p->pgs_schedule = 0;
    p->organize_contacts = 1;
    p->host_threads = 0;
}

// Derived host lists from the structure (parent/slot/sub/members).
int rebuild_structure(am_ctx* c, bool bodies_changed) {
    const int n = c->n_ent;
    if (bodies_changed) {
        c->planes.clear();
        c->grid_bodies.clear();
        for (int i = 0; i < c->n_phys; i++) {
            if (c->desc[i].geom == AM_GEOM_PLANE) c->planes.push_back(i);
            else c->grid_bodies.push_back(i);
        }
        // planes by scene, so a body's pair test visits only its scene's
        std::stable_sort(c->planes.begin(), c->planes.end(),
                         [&](int a, int b) { return c->desc[a].scene < c->desc[b].scene; });
        c->plane_start.assign(c->n_scenes + 1, 0);
        for (int p : c->planes) c->plane_start[c->desc[p].scene + 1]++;
        for (int k = 0; k < c->n_scenes; k++) c->plane_start[k + 1] += c->plane_start[k];
    }
    int n_slots = (int)c->slot_entity.size();
    if (n_slots >= kMaxSlots) return fail(c, AM_ERR_INVALID, "too many top-level bodies for the pair key");
    c->top_pinned.assign(std::max(n_slots, 1), 0);
    c->top_sleeping.assign(std::max(n_slots, 1), 0);
    c->gravity_targets.clear();
    c->integrate_targets.clear();
    for (int s = 0; s < n_slots; s++) {
        int e = c->slot_entity[s];
        c->top_pinned[s] = (c->flags[e] & AM_BODY_PINNED) ? 1 : 0;
        c->top_sleeping[s] = (c->flags[e] & AM_BODY_SLEEPING) ? 1 : 0;
        c->gravity_targets.push_back(e);
        if (is_collection(c, e))
            for (int q = 0; q < c->member_count[e]; q++) c->gravity_targets.push_back(c->members[c->member_first[e] + q]);
        if (!(c->flags[e] & (AM_BODY_PINNED | AM_BODY_SLEEPING))) c->integrate_targets.push_back(e);
    }
    for (int i = 0; i < n; i++)
        if (c->sub[i] >= kMaxSub) return fail(c, AM_ERR_INVALID, "collection too large for the pair key");
    AM_HIP(c, upload(c, c->d_parent, c->parent));
    AM_HIP(c, upload(c, c->d_slot, c->slot));
    AM_HIP(c, upload(c, c->d_sub, c->sub));
    AM_HIP(c, upload(c, c->d_slot_entity, c->slot_entity));
    AM_HIP(c, upload(c, c->d_member_first, c->member_first));
    AM_HIP(c, upload(c, c->d_members, c->members));
    if (bodies_changed) {
        AM_HIP(c, upload(c, c->d_planes, c->planes));
        AM_HIP(c, upload(c, c->d_plane_start, c->plane_start));
        AM_HIP(c, upload(c, c->d_grid, c->grid_bodies));
    }
    AM_HIP(c, upload(c, c->d_top_pinned, c->top_pinned));
    AM_HIP(c, upload(c, c->d_top_sleeping, c->top_sleeping));
    AM_HIP(c, upload(c, c->d_gravity, c->gravity_targets));
    AM_HIP(c, upload(c, c->d_integrate, c->integrate_targets));
    AM_HIP(c, upload(c, c->d_flags, c->flags));
    // collection members: all (push-down at clear, RigidCollection.clearBodies
    // visits every collection) and those the single sweep integrates (members
    // of awake collections that are not pinned, CollisionProcessor.java:288-296)
    c->member_list.clear();
    c->member_adv.clear();
    for (int e : c->slot_entity)
        if (is_collection(c, e))
            for (int q = 0; q < c->member_count[e]; q++) {
                int b = c->members[c->member_first[e] + q];
                c->member_list.push_back(b);
                if (!(c->flags[e] & AM_BODY_SLEEPING) && !(c->flags[b] & AM_BODY_PINNED)) c->member_adv.push_back(b);
            }
    AM_HIP(c, upload(c, c->d_member_list, c->member_list));
    AM_HIP(c, upload(c, c->d_member_adv, c->member_adv));
    if (!bodies_changed) {
        AM_HIP(c, c->inc_count.ensure(n + 1));
        AM_HIP(c, c->inc_start.ensure(n + 1));
        AM_HIP(c, c->inc_cursor.ensure(n + 1));
        c->layout_valid = false;
        return AM_OK;
    }
    if (c->b2c.size() < (size_t)n * 12) c->b2c.resize((size_t)n * 12, 0.0);
    AM_HIP(c, upload(c, c->d_b2c, c->b2c));
    // grid: cells sized for the bodies up to kLargeRatio times the median
    // bounding radius; the few larger ones (platforms, walls made of boxes)
    // stay out of the grid and are paired with every body by the pair kernel
    const int n_grid = (int)c->grid_bodies.size();
    std::vector<double> radii(n_grid);
    for (int k = 0; k < n_grid; k++) {
        const am_body_desc& d = c->desc[c->grid_bodies[k]];
        radii[k] = (d.geom == AM_GEOM_BOX || d.geom == AM_GEOM_COMPOSITE)
                       ? d.radius
                       : (d.tree_count > 0 ? c->nodes[d.tree_first].r : 0.0);
    }
    c->grid_is_large.assign(std::max(n_grid, 1), 0);
    c->grid_large.clear();
    if (n_grid > 0) {
        std::vector<double> sorted = radii;
        std::nth_element(sorted.begin(), sorted.begin() + n_grid / 2, sorted.end());
        const double limit = kLargeRatio * sorted[n_grid / 2];
        for (int k = 0; k < n_grid; k++)
            if (radii[k] > limit) c->grid_large.push_back(k);
        if ((int)c->grid_large.size() > kMaxLarge) c->grid_large.clear();  // not a few outliers: one grid for all
        for (int k : c->grid_large) c->grid_is_large[k] = 1;
    }
    double rmax = 0.0;
    for (int k = 0; k < n_grid; k++)
        if (!c->grid_is_large[k]) rmax = std::max(rmax, radii[k]);
    c->cell_size = std::max(2.0 * rmax, 1e-6);
    AM_HIP(c, upload(c, c->d_grid_is_large, c->grid_is_large));
    AM_HIP(c, upload(c, c->d_grid_large, c->grid_large));
    int tsize = 16;
    while (tsize < 2 * (int)c->grid_bodies.size()) tsize *= 2;
    c->hash_mask = tsize - 1;
    int ng = (int)c->grid_bodies.size();
    AM_HIP(c, c->d_sphere.ensure((size_t)std::max(ng, 1) * 4));
    AM_HIP(c, c->d_cell.ensure((size_t)std::max(ng, 1) * 4));
    AM_HIP(c, c->d_bucket_count.ensure(tsize));
    AM_HIP(c, c->d_bucket_start.ensure(tsize + 1));
    AM_HIP(c, c->d_bucket_cursor.ensure(tsize));
    AM_HIP(c, c->d_bucket_items.ensure(std::max(ng, 1)));
    const int nt = pair_tasks(std::max(ng, 1));
    AM_HIP(c, c->d_pair_counts.ensure((size_t)nt + 1));
    AM_HIP(c, c->d_pair_offsets.ensure((size_t)nt + 1));
    int r = ensure_scan_temp(c, std::max(tsize, nt) + 1);
    if (r) return r;
    // incidence arrays are per entity
    AM_HIP(c, c->inc_count.ensure(n + 1));
    AM_HIP(c, c->inc_start.ensure(n + 1));
    AM_HIP(c, c->inc_cursor.ensure(n + 1));
    c->layout_valid = false;
    return AM_OK;
}

}  // namespace am_rt

extern "C" {

const char* am_version(void) { return "libam 0.1 (gfx950, FP64 coloured PGS)"; }

int am_params_default(am_params* out) {
    if (!out) return AM_ERR_INVALID;
    default_params(out);
    return AM_OK;
}

int am_ctx_create(int32_t device, am_ctx** out) {
    if (!out) return AM_ERR_INVALID;
    *out = nullptr;
    am_ctx* c = new (std::nothrow) am_ctx();
    if (!c) return AM_ERR_INVALID;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0 || device < 0 || device >= n_dev) {
        delete c;
        return AM_ERR_HIP;
    }
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return AM_ERR_HIP;
    }
    default_params(&c->params);
    if (c->d_scalars.ensure(32) != hipSuccess || hipHostMalloc((void**)&c->h_scalars, 32 * sizeof(int)) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return AM_ERR_HIP;
    }
    std::memset(c->h_scalars, 0, 32 * sizeof(int));
    *out = c;
    return AM_OK;
}

int am_ctx_destroy(am_ctx* c) {
    if (!c) return AM_ERR_INVALID;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->h_scalars) (void)hipHostFree(c->h_scalars);
    for (auto& e : c->kev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->lev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->sev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return AM_OK;
}

const char* am_last_error(const am_ctx* c) { return c ? c->err.c_str() : "null context"; }

int am_params_set(am_ctx* c, const am_params* p) {
    if (!c || !p) return fail(c, AM_ERR_INVALID, "null argument");
    if (p->iterations < 0 || p->iterations_in_collection < 0) return fail(c, AM_ERR_INVALID, "negative iterations");
    if (p->pgs_block < 0 || p->pgs_block > 256 || (p->pgs_block % 64) != 0)
        return fail(c, AM_ERR_INVALID, "pgs_block must be 0 or a multiple of 64 up to 256");
    if (p->pgs_grid < 0) return fail(c, AM_ERR_INVALID, "pgs_grid must be >= 0");
    if (p->pgs_schedule != 0 && p->pgs_schedule != 1) return fail(c, AM_ERR_INVALID, "pgs_schedule must be 0 or 1");
    if (p->sleep_step_accum < 0 || p->sleep_step_accum > am::kMaxSleepAccum)
        return fail(c, AM_ERR_INVALID, "sleep_step_accum must be in 0..200");
    if (p->host_threads < 0) return fail(c, AM_ERR_INVALID, "host_threads must be >= 0");
    c->params = *p;
    return AM_OK;
}

}  // extern "C"

namespace am_rt {

int check_ctx(am_ctx* c) {
    if (!c) return AM_ERR_INVALID;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, AM_ERR_HIP, "hipSetDevice failed");
    return AM_OK;
}

// jinv = R jinv0 R^T (zero for pinned) for every entity.
int refresh_jinv(am_ctx* c) {
    launch_update_jinv(c->stream, bodies_view(c));
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

// Colouring and sweep layout for the current groups (depends only on the
// contact list and the solve bodies).
int build_layout(am_ctx* c) {
    if (c->layout_valid) return AM_OK;
    const int ng = c->n_groups;
    if (ng >= (1 << 28)) return fail(c, AM_ERR_INVALID, "more than 2^28 contact groups");
    hipStream_t s = c->stream;
    int* sc = c->d_scalars.p;
    fill_int(s, c->inc_count.p, c->n_ent + 1, 0);
    fill_int(s, c->inc_cursor.p, c->n_ent + 1, 0);
    if (ng > 0) {
        launch_incidence(s, false, ng, c->g_a.p, c->g_b.p, c->d_flags.p, c->inc_count.p, nullptr, nullptr, nullptr);
        exclusive_scan(s, c->inc_count.p, c->inc_start.p, c->n_ent + 1, sc + 3, c->d_scan_temp.p);
        AM_HIP(c, c->inc_items.ensure(2 * (size_t)ng + 1));
        launch_incidence(s, true, ng, c->g_a.p, c->g_b.p, c->d_flags.p, c->inc_count.p, c->inc_start.p,
                         c->inc_cursor.p, c->inc_items.p);
        fill_int(s, c->g_colour_a.p, ng, -1);
        AM_HIP(c, c->g_blocker.ensure(std::max(ng, 1)));
        fill_int(s, c->g_blocker.p, ng, -1);
        int* cin = c->g_colour_a.p;
        int* cout = c->g_colour_b.p;
        // Rounds run in batches between reads of a flag that some group is
        // still uncoloured: a round after the last one only copies the
        // colours, so the colouring does not depend on the batch size.
        constexpr int kRoundsPerCheck = 8;
        for (int round = 0; round < 4096; round += kRoundsPerCheck) {
            for (int k = 0; k < kRoundsPerCheck; k++) {
                fill_int(s, sc + 5, 1, 0);
                launch_colour_round(s, ng, c->g_a.p, c->g_b.p, c->d_flags.p, c->inc_start.p, c->inc_count.p,
                                    c->inc_items.p, cin, cout, sc + 5, c->layout_layer, c->g_blocker.p);
                std::swap(cin, cout);
            }
            int r = read_scalars(c, 5, 1);
            if (r) return r;
            if (c->h_scalars[5] == 0) break;
        }
        launch_colour_keys(s, ng, cin, c->g_start.p, c->g_keys.p);
        sort_u64(s, c->g_keys.p, c->g_keys_sorted.p, ng, c->d_sort_temp.p, c->d_sort_temp.n, 0, 64);
    }
    AM_HIP(c, c->s_colour_start.ensure(kMaxColours + 2));
    AM_HIP(c, c->s_group_desc.ensure((size_t)std::max(ng, 1) * 4));
    launch_colour_layout(s, ng, c->g_keys_sorted.p, c->g_start.p, c->g_a.p, c->g_b.p, c->s_colour_start.p,
                         c->g_sizes.p, c->g_pos_start.p, c->s_order.p, c->s_group_desc.p, c->s_pos_a.p,
                         c->g_spos.p, sc + 9, c->d_scan_temp.p);
    AM_HIP(c, c->s_flow.ensure((size_t)std::max(ng, 1) * 4));
    AM_HIP(c, c->s_pred.ensure((size_t)std::max(ng, 1) * 2));
    launch_flow_ranks(s, c->n_ent, ng, c->d_flags.p, c->inc_start.p, c->inc_count.p, c->inc_items.p, c->g_spos.p,
                      c->s_group_desc.p, c->s_flow.p, c->s_pred.p);
    AM_HIP(c, hipGetLastError());
    int cs[kMaxColours + 2];
    AM_HIP(c, hipMemcpyAsync(cs, c->s_colour_start.p, sizeof cs, hipMemcpyDeviceToHost, s));
    AM_HIP(c, hipStreamSynchronize(s));
    int ncol = 0, maxc = 0;
    for (int q = 0; q <= kMaxColours; q++) {
        int sz = cs[q + 1] - cs[q];
        if (sz > 0) ncol++;
        if (q < kMaxColours) maxc = std::max(maxc, sz);
    }
    c->n_colours = ncol;
    c->max_colour_groups = maxc;
    c->layout_valid = true;
    return AM_OK;
}

int ensure_old_capacity(am_ctx* c, int n) {
    size_t want = std::max(n, 1);
    if (c->o_b1.n < want) {
        size_t cap = want + want / 4;
        AM_HIP(c, c->o_b1.ensure(cap));
        AM_HIP(c, c->o_b2.ensure(cap));
        AM_HIP(c, c->o_bv1.ensure(cap));
        AM_HIP(c, c->o_bv2.ensure(cap));
        AM_HIP(c, c->o_info.ensure(cap));
        AM_HIP(c, c->o_pB1.ensure(cap * 3));
        AM_HIP(c, c->o_lam.ensure(cap * 3));
        AM_HIP(c, c->o_viol.ensure(cap));
        AM_HIP(c, c->o_prev.ensure(cap));
    }
    int tsize = 16;
    while (tsize < 2 * n + 16) tsize *= 2;
    AM_HIP(c, c->o_table.ensure(tsize));
    c->table_mask = tsize - 1;
    return AM_OK;
}

int groups_after_contacts(am_ctx* c) {
    if (c->n_contacts > 0) {
        launch_groups(c->stream, c->n_contacts, c->d_scalars.p + 1, c->c_b1.p, c->c_b2.p, c->d_slot.p, c->d_parent.p,
                      c->g_flags.p, c->g_scanned.p, c->g_start.p, c->g_a.p, c->g_b.p, c->d_scalars.p + 2,
                      c->d_scan_temp.p);
        AM_HIP(c, hipGetLastError());
        int r = read_scalars(c, 2, 1);
        if (r) return r;
        c->n_groups = c->h_scalars[2];
    } else {
        c->n_groups = 0;
    }
    c->layout_valid = false;
    return AM_OK;
}

int do_detect(am_ctx* c) {
    c->sweep_order_n = -1;
    hipStream_t s = c->stream;
    BodiesView B = bodies_view(c);
    TreeView T = tree_view(c);
    int* sc = c->d_scalars.p;
    launch_tree_world_centres(s, B, T, c->d_node_body.p, c->d_node_sub.p, c->n_nodes);
    const int ng = (int)c->grid_bodies.size();
    const int tsize = c->hash_mask + 1;
    launch_bounds(s, B, T, c->d_grid.p, ng, c->d_sphere.p, c->cell_size, c->d_cell.p, c->hash_mask,
                  c->d_grid_is_large.p);
    fill_int(s, c->d_bucket_count.p, tsize, 0);
    fill_int(s, c->d_bucket_cursor.p, tsize, 0);
    launch_buckets(s, c->d_cell.p, ng, c->d_bucket_count.p, nullptr, nullptr, nullptr, false);
    exclusive_scan(s, c->d_bucket_count.p, c->d_bucket_start.p, tsize, sc + 3, c->d_scan_temp.p);
    launch_buckets(s, c->d_cell.p, ng, c->d_bucket_count.p, c->d_bucket_start.p, c->d_bucket_cursor.p,
                   c->d_bucket_items.p, true);
    PairArgs A;
    A.B = B;
    A.T = T;
    A.grid_order = c->d_grid.p;
    A.n_grid = ng;
    A.sphere = c->d_sphere.p;
    A.cell = c->d_cell.p;
    A.bucket_start = c->d_bucket_start.p;
    A.bucket_count = c->d_bucket_count.p;
    A.bucket_items = c->d_bucket_items.p;
    A.mask = c->hash_mask;
    A.planes = c->d_planes.p;
    A.plane_start = c->d_plane_start.p;
    A.top_pinned = c->d_top_pinned.p;
    A.top_sleeping = c->d_top_sleeping.p;
    A.large = c->d_grid_large.p;
    A.n_large = (int)c->grid_large.size();
    fill_int(s, sc + 4, 1, 0);
    launch_pairs_impl(s, false, A, c->d_pair_counts.p, nullptr, nullptr, 0, sc + 4);
    exclusive_scan(s, c->d_pair_counts.p, c->d_pair_offsets.p, pair_tasks(ng), sc + 0, c->d_scan_temp.p);
    AM_HIP(c, hipGetLastError());
    int r = read_scalars(c, 0, 1);
    if (r) return r;
    int np = ng > 0 ? c->h_scalars[0] : 0;
    c->n_pairs = np;
    if (np > 0) {
        if ((int)c->d_keys.n < np) {
            size_t cap = (size_t)np + np / 4 + 64;
            AM_HIP(c, c->d_keys.ensure(cap));
            AM_HIP(c, c->d_keys_unsorted.ensure(cap));
            AM_HIP(c, c->d_narrow_counts.ensure(cap + 1));
            AM_HIP(c, c->d_narrow_offsets.ensure(cap + 1));
            size_t st = sort_temp_bytes((int)cap);
            if (st > c->d_sort_temp.n) AM_HIP(c, c->d_sort_temp.ensure(st));
            r = ensure_scan_temp(c, (int)cap + 1);
            if (r) return r;
        }
        launch_pairs_impl(s, true, A, nullptr, c->d_pair_offsets.p, c->d_keys_unsorted.p, (int)c->d_keys.n, sc + 4);
        sort_u64(s, c->d_keys_unsorted.p, c->d_keys.p, np, c->d_sort_temp.p, c->d_sort_temp.n, 0, 64);
        fill_int(s, sc + 0, 1, np);
        launch_narrow(s, false, B, T, c->d_keys.p, sc + 0, np, c->d_slot_entity.p, c->d_member_first.p,
                      c->d_members.p, c->d_narrow_counts.p, nullptr, contacts_view(c), sc + 4);
        exclusive_scan(s, c->d_narrow_counts.p, c->d_narrow_offsets.p, np, sc + 1, c->d_scan_temp.p);
        AM_HIP(c, hipGetLastError());
        r = read_scalars(c, 1, 1);
        if (r) return r;
        int nc = c->h_scalars[1];
        r = ensure_contact_capacity(c, nc);
        if (r) return r;
        c->n_contacts = nc;
        if (nc > 0)
            launch_narrow(s, true, B, T, c->d_keys.p, sc + 0, np, c->d_slot_entity.p, c->d_member_first.p,
                          c->d_members.p, nullptr, c->d_narrow_offsets.p, contacts_view(c), sc + 4);
    } else {
        c->n_contacts = 0;
        fill_int(s, sc + 1, 1, 0);
    }
    r = read_scalars(c, 4, 1);
    if (r) return r;
    if (c->h_scalars[4] & 2) return fail(c, AM_ERR_CAPACITY, "sphere-tree traversal stack overflow (tree too deep)");
    if (c->h_scalars[4] & 1) return fail(c, AM_ERR_CAPACITY, "contact or pair buffer overflow");
    if (c->n_contacts == 0) c->n_old = 0;  // CollisionProcessor.java:99-101
    return groups_after_contacts(c);
}

// ---- learned warm start helpers ------------------------------------------
constexpr int OW1 = 0, OW2 = kPadW1, OW3 = OW2 + kPadW2, OW4 = OW3 + kPadW3;
constexpr int OB1 = OW4 + kPadW4, OB2 = OB1 + kPadB1, OB3 = OB2 + kPadB2, OB4 = OB3 + kPadB3;
constexpr int kBfTotal = kPadW1 + kPadW2 + kPadW3 + kPadW4;

// flat (reference layout of include/am.h) <-> padded device layout
void pack_params(const float* flat, std::vector<float>& pad) {
    pad.assign(kPadTotal, 0.f);
    const float* q = flat;
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 64; j++) pad[OW1 + i * 64 + j] = *q++;
    for (int j = 0; j < 64; j++) pad[OB1 + j] = *q++;
    for (int i = 0; i < 64 * 64; i++) pad[OW2 + i] = *q++;
    for (int j = 0; j < 64; j++) pad[OB2 + j] = *q++;
    for (int i = 0; i < 64 * 64; i++) pad[OW3 + i] = *q++;
    for (int j = 0; j < 64; j++) pad[OB3 + j] = *q++;
    for (int i = 0; i < 64; i++)
        for (int j = 0; j < 3; j++) pad[OW4 + i * 16 + j] = *q++;
    for (int j = 0; j < 3; j++) pad[OB4 + j] = *q++;
}
void unpack_params(const std::vector<float>& pad, float* flat) {
    float* q = flat;
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 64; j++) *q++ = pad[OW1 + i * 64 + j];
    for (int j = 0; j < 64; j++) *q++ = pad[OB1 + j];
    for (int i = 0; i < 64 * 64; i++) *q++ = pad[OW2 + i];
    for (int j = 0; j < 64; j++) *q++ = pad[OB2 + j];
    for (int i = 0; i < 64 * 64; i++) *q++ = pad[OW3 + i];
    for (int j = 0; j < 64; j++) *q++ = pad[OB3 + j];
    for (int i = 0; i < 64; i++)
        for (int j = 0; j < 3; j++) *q++ = pad[OW4 + i * 16 + j];
    for (int j = 0; j < 3; j++) *q++ = pad[OB4 + j];
}

MlpWeights mlp_weights(am_ctx* c) {
    const unsigned short* bf = c->m_bf.p;
    const float* p = c->m_p.p;
    return MlpWeights{bf + OW1, bf + OW2, bf + OW3, bf + OW4, p + OB1, p + OB2, p + OB3, p + OB4};
}
MlpGrads mlp_grads(am_ctx* c) {
    float* g = c->m_g.p;
    return MlpGrads{g + OW1, g + OW2, g + OW3, g + OW4, g + OB1, g + OB2, g + OB3, g + OB4};
}

int mlp_alloc(am_ctx* c) {
    AM_HIP(c, c->m_p.ensure(kPadTotal));
    AM_HIP(c, c->m_g.ensure(kPadTotal));
    AM_HIP(c, c->m_m.ensure(kPadTotal));
    AM_HIP(c, c->m_v.ensure(kPadTotal));
    AM_HIP(c, c->m_bf.ensure(kBfTotal));
    AM_HIP(c, c->d_loss.ensure(4));
    return AM_OK;
}

int mlp_upload(am_ctx* c, const float* flat) {
    int r = mlp_alloc(c);
    if (r) return r;
    std::vector<float> pad;
    pack_params(flat, pad);
    hipStream_t s = c->stream;
    AM_HIP(c, hipMemcpyAsync(c->m_p.p, pad.data(), kPadTotal * 4, hipMemcpyHostToDevice, s));
    std::vector<float> z(kPadTotal, 0.f);
    AM_HIP(c, hipMemcpyAsync(c->m_g.p, z.data(), kPadTotal * 4, hipMemcpyHostToDevice, s));
    AM_HIP(c, hipMemcpyAsync(c->m_m.p, z.data(), kPadTotal * 4, hipMemcpyHostToDevice, s));
    AM_HIP(c, hipMemcpyAsync(c->m_v.p, z.data(), kPadTotal * 4, hipMemcpyHostToDevice, s));
    launch_to_bf16(s, c->m_p.p, c->m_bf.p, kBfTotal);
    AM_HIP(c, hipGetLastError());
    AM_HIP(c, hipStreamSynchronize(s));
    c->adam_step = 0;
    c->mlp_ready = true;
    return AM_OK;
}

int compute_features(am_ctx* c, double dt) {
    const int n = c->n_contacts;
    AM_HIP(c, c->f_X.ensure((size_t)std::max(n, 1) * kMlpIn));
    AM_HIP(c, c->f_Y.ensure((size_t)std::max(n, 1) * kMlpOut));
    AM_HIP(c, c->f_scale.ensure(std::max(n, 1)));
    AM_HIP(c, c->f_lam.ensure((size_t)std::max(n, 1) * 3));
    FeatureArgs F;
    F.B = bodies_view(c);
    F.C = contacts_view(c);
    F.n = n;
    F.g = c->params.gravity_amount != 0.0 ? std::fabs(c->params.gravity_amount) : 1.0;
    F.dt = dt > 0 ? dt : 0.01;
    F.out = c->f_X.p;
    F.scale = c->f_scale.p;
    launch_features(c->stream, F);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

int train_append(am_ctx* c) {
    const int n = c->n_contacts;
    if (n <= 0) return AM_OK;
    size_t need = (size_t)c->t_n + n;
    if (c->t_X.n < need * kMlpIn) {
        size_t cap = std::max(need + need / 2, (size_t)65536);
        DBuf<float> nx, nt;
        AM_HIP(c, nx.ensure(cap * kMlpIn));
        AM_HIP(c, nt.ensure(cap * kMlpOut));
        if (c->t_n > 0) {
            AM_HIP(c, hipMemcpyAsync(nx.p, c->t_X.p, (size_t)c->t_n * kMlpIn * 4, hipMemcpyDeviceToDevice, c->stream));
            AM_HIP(c, hipMemcpyAsync(nt.p, c->t_T.p, (size_t)c->t_n * kMlpOut * 4, hipMemcpyDeviceToDevice, c->stream));
        }
        AM_HIP(c, hipStreamSynchronize(c->stream));
        std::swap(c->t_X.p, nx.p);
        std::swap(c->t_X.n, nx.n);
        std::swap(c->t_T.p, nt.p);
        std::swap(c->t_T.n, nt.n);
    }
    AM_HIP(c, hipMemcpyAsync(c->t_X.p + (size_t)c->t_n * kMlpIn, c->f_X.p, (size_t)n * kMlpIn * 4,
                             hipMemcpyDeviceToDevice, c->stream));
    launch_targets(c->stream, c->c_lam.p, c->f_scale.p, c->f_X.p, n, c->t_T.p + (size_t)c->t_n * kMlpOut);
    AM_HIP(c, hipGetLastError());
    c->t_n += n;
    return AM_OK;
}

// Start residual of the lambdas currently in the records (sweep order).
void pgs_shape(am_ctx* c, const am_params& P, int* grid, int* block);

int start_residual(am_ctx* c, PgsArgs G, double* out) {
    hipStream_t s = c->stream;
    AM_HIP(c, hipMemcpyAsync(c->d_dv.p, c->dv_save.p, (size_t)c->n_ent * 48, hipMemcpyDeviceToDevice, s));
    G.iterations = 0;
    G.warm_start = 1;
    fill_int(s, c->d_barrier.p, 2, 0);
    // the warm-start pass alone, by the coloured kernel at its automatic
    // shape (its bits do not depend on the shape)
    am_params pc = c->params;
    pc.pgs_schedule = 1;
    pc.pgs_grid = 0;
    pc.pgs_block = 0;
    int grid = 1, block = 256;
    pgs_shape(c, pc, &grid, &block);
    AM_HIP(c, launch_pgs(s, G, grid, block, c->params.pgs_plain_launch != 0));
    AM_HIP(c, c->d_resid.ensure(1));
    fill_int(s, (int*)c->d_resid.p, 2, 0);
    launch_start_residual(s, c->s_rec.p, c->s_meta.p, c->d_dv.p, c->d_minv.p, c->d_jinv.p, c->n_contacts, G.omega,
                          G.compliance, c->d_resid.p);
    AM_HIP(c, hipGetLastError());
    unsigned long long bits = 0;
    AM_HIP(c, hipMemcpyAsync(&bits, c->d_resid.p, 8, hipMemcpyDeviceToHost, s));
    AM_HIP(c, hipStreamSynchronize(s));
    std::memcpy(out, &bits, 8);
    AM_HIP(c, hipMemcpyAsync(c->d_dv.p, c->dv_save.p, (size_t)c->n_ent * 48, hipMemcpyDeviceToDevice, s));
    return AM_OK;
}

// Launch shape of the PGS kernel under P (after build_layout): workgroups and
// threads per workgroup, clamped to what can be co-resident.
void pgs_shape(am_ctx* c, const am_params& P, int* grid, int* block) {
    bool ok = false;
    if (P.pgs_schedule == 0) {
        if (c->flow_max_grid == 0) {
            c->flow_max_grid = pgs_flow_max_blocks(&ok);
            if (!ok) c->flow_max_grid = 1;
        }
        // one wave per workgroup, each walking every W-th step of up to 4
        // groups: 512 waves up to ~20k groups, then one per 64 groups (the
        // settled 20k-box pile, ~78k groups, runs fastest at 1024 or more)
        int g = P.pgs_grid > 0 ? P.pgs_grid : std::max(std::min((c->n_groups + 3) / 4, 512), c->n_groups / 64);
        *grid = std::max(1, std::min(g, c->flow_max_grid));
        *block = 64;
        return;
    }
    const int b = P.pgs_block > 0 ? P.pgs_block : 256;
    int& mx = c->pgs_max_grid[b / 64];
    if (mx == 0) {
        mx = pgs_max_blocks(b, &ok);
        if (!ok) mx = 1;
    }
    // groups of a colour are dealt block-cyclically; about 20 groups of the
    // largest colour per workgroup (the best measured on the settling box pile)
    int g = P.pgs_grid > 0 ? P.pgs_grid : (c->max_colour_groups + 19) / 20;
    *grid = std::max(1, std::min(g, mx));
    *block = b;
}

int do_solve(am_ctx* c, double dt, int post, bool main_solve, int* iters, double* last) {
    if (main_solve) c->sweep_order_n = -1;
    const am_params& P = c->params;
    c->changes.clear();
    if (c->n_contacts == 0) {
        c->last_iters = 0;
        c->last_change = 0;
        if (iters) *iters = 0;
        if (last) *last = 0;
        return AM_OK;
    }
    int r = build_layout(c);
    if (r) return r;
    hipStream_t s = c->stream;
    SetupArgs S;
    S.B = bodies_view(c);
    S.C = contacts_view(c);
    S.n_contacts = c->d_scalars.p + 1;
    S.order = c->s_order.p;
    S.rec = c->s_rec.p;
    S.meta = c->s_meta.p;
    S.pos_a = c->s_pos_a.p;
    S.sstate = c->s_state.p;
    S.dt = dt;
    if (main_solve) S.feedback = (!P.enable_post_stabilization || post) ? P.feedback_stiffness : 0.0;
    else S.feedback = P.feedback_stiffness;
    S.compliance = P.enable_compliance ? P.compliance : 0.0;
    S.rest_val = P.restitution;
    S.fric_val = P.friction;
    S.rest_ovr = P.restitution_override;
    S.fric_ovr = P.friction_override;
    S.post = main_solve ? post : 0;
    S.in_collection = main_solve ? 0 : 1;
    S.lam_alt = nullptr;
    S.iters_used = nullptr;
    S.scene_used = nullptr;
    const bool flow = P.pgs_schedule == 0;
    fill_int(s, c->d_scalars.p + 1, 1, c->n_contacts);
    for (auto& e : c->kev)
        if (!e) AM_HIP(c, hipEventCreate(&e));
    for (auto& e : c->lev)
        if (!e) AM_HIP(c, hipEventCreate(&e));
    const bool learned = main_solve && P.learned_warm_start && c->mlp_ready;
    const bool collect = main_solve && P.collect_training;
    c->used_learned = 0;
    c->res_plain = 0;
    c->res_learned = 0;
    c->ms_learned = 0;
    AM_HIP(c, hipEventRecord(c->lev[0], s));
    if (learned || collect) {
        r = compute_features(c, dt);
        if (r) return r;
    }
    AM_HIP(c, hipEventRecord(c->kev[0], s));
    launch_setup(s, S, c->n_contacts);
    const int iterations = main_solve ? P.iterations : P.iterations_in_collection;
    AM_HIP(c, c->s_iter_max.ensure(std::max(iterations, 1)));
    fill_int(s, (int*)c->s_iter_max.p, 2 * std::max(iterations, 1), 0);
    PgsArgs G;
    G.colour_start = c->s_colour_start.p;
    G.group_desc = c->s_group_desc.p;
    G.meta = c->s_meta.p;
    G.rec = c->s_rec.p;
    G.sstate = c->s_state.p;
    G.dv = c->d_dv.p;
    G.minv = c->d_minv.p;
    G.jinv = c->d_jinv.p;
    G.bflags = c->d_flags.p;
    G.iterations = iterations;
    G.check_tolerance = main_solve ? 1 : 0;
    G.warm_start = 1;
    G.tolerance = P.tolerance;
    G.omega = P.omega;
    G.sliding = P.sliding_threshold;
    G.compliance = S.compliance;
    AM_HIP(c, c->d_barrier.ensure(2));
    fill_int(s, c->d_barrier.p, 2, 0);
    G.barrier = (unsigned*)c->d_barrier.p;
    G.error = c->d_barrier.p + 1;
    G.iter_max = c->s_iter_max.p;
    G.iters_used = c->d_scalars.p + 6;
    // several scenes in one context stop one by one, as each would alone
    G.per_scene = main_solve && c->n_scenes > 1 && P.tolerance > 0 ? 1 : 0;
    G.n_scenes = c->n_scenes;
    G.g_scene = nullptr;
    G.scene_max = nullptr;
    G.scene_done = nullptr;
    G.scene_groups = nullptr;
    G.scene_used = nullptr;
    if (main_solve) c->scene_solve_valid = false;
    if (G.per_scene) {
        const size_t ns = (size_t)c->n_scenes;
        AM_HIP(c, c->g_scene.ensure(std::max(c->n_groups, 1)));
        AM_HIP(c, c->scene_used.ensure(ns));
        AM_HIP(c, c->scene_groups.ensure(ns));
        AM_HIP(c, c->scene_max.ensure(ns * iterations));
        AM_HIP(c, c->scene_done.ensure(ns * (iterations + 1)));
        fill_int(s, (int*)c->scene_groups.p, (int)ns, 0);
        fill_int(s, (int*)c->scene_max.p, (int)(2 * ns * iterations), 0);
        fill_int(s, (int*)c->scene_done.p, (int)(2 * ns * (iterations + 1)), 0);
        launch_group_scene(s, c->s_group_desc.p, c->n_groups, c->s_order.p, c->c_b1.p, c->d_desc.p, c->n_scenes,
                           c->g_scene.p, c->scene_groups.p);
        G.g_scene = c->g_scene.p;
        G.scene_max = c->scene_max.p;
        G.scene_done = c->scene_done.p;
        G.scene_groups = c->scene_groups.p;
        G.scene_used = c->scene_used.p;
        S.scene_used = c->scene_used.p;
    }
    int grid = 1, block = 64;
    pgs_shape(c, P, &grid, &block);
    if (main_solve) {
        c->last_grid = grid;
        c->last_block = block;
    }
    if (learned) {
        // plain start residual, predicted start residual; keep the better
        AM_HIP(c, c->dv_save.ensure((size_t)c->n_ent * 6));
        AM_HIP(c, hipMemcpyAsync(c->dv_save.p, c->d_dv.p, (size_t)c->n_ent * 48, hipMemcpyDeviceToDevice, s));
        r = start_residual(c, G, &c->res_plain);
        if (r) return r;
        launch_mlp_forward(s, mlp_weights(c), c->f_X.p, c->n_contacts, c->f_Y.p);
        launch_apply_prediction(s, contacts_view(c), bodies_view(c), c->f_Y.p, c->f_scale.p, c->n_contacts,
                                P.friction_override, P.friction, c->f_lam.p);
        launch_rec_lambda(s, c->s_rec.p, c->s_order.p, c->f_lam.p, c->n_contacts);
        AM_HIP(c, hipGetLastError());
        r = start_residual(c, G, &c->res_learned);
        if (r) return r;
        if (c->res_learned < c->res_plain) {
            c->used_learned = 1;
            AM_HIP(c, hipMemcpyAsync(c->c_lam.p, c->f_lam.p, (size_t)c->n_contacts * 24, hipMemcpyDeviceToDevice, s));
        } else {
            launch_rec_lambda(s, c->s_rec.p, c->s_order.p, c->c_lam.p, c->n_contacts);
        }
        fill_int(s, c->d_barrier.p, 2, 0);
    }
    AM_HIP(c, hipEventRecord(c->lev[1], s));
    AM_HIP(c, c->s_start.ensure((size_t)std::max(c->n_contacts, 1) * 3));
    AM_HIP(c, hipMemcpyAsync(c->s_start.p, c->c_lam.p, (size_t)c->n_contacts * 24, hipMemcpyDeviceToDevice, s));
    if (flow) {
        const size_t n_dv = (size_t)c->n_ent * 6;
        AM_HIP(c, c->s_lam_alt.ensure((size_t)c->n_contacts * 4));
        AM_HIP(c, c->dv_alt.ensure(std::max(n_dv, (size_t)6)));
        const size_t n_slot = (size_t)2 * 2 * std::max(c->n_groups, 1) * kSlotDoubles;
        AM_HIP(c, c->f_slot.ensure(n_slot));
        AM_HIP(c, c->d_pass_done.ensure((size_t)(iterations + 2) * kPassWords));
        fill_u64(s, c->f_slot.p, n_slot, kSlotEmpty);
        fill_int(s, (int*)c->d_pass_done.p, 2 * (iterations + 2) * kPassWords, 0);
        AM_HIP(c, hipMemcpyAsync(c->dv_alt.p, c->d_dv.p, n_dv * sizeof(double), hipMemcpyDeviceToDevice, s));
        const int fgrid = grid;
        FlowArgs F;
        F.P = G;
        F.flow = c->s_flow.p;
        F.pred = c->s_pred.p;
        F.slot = c->f_slot.p;
        F.n_groups = c->n_groups;
        F.dv_alt = c->dv_alt.p;
        F.lam_alt = c->s_lam_alt.p;
        F.pass_done = c->d_pass_done.p;
        F.n_waves = (unsigned)fgrid;
        S.lam_alt = c->s_lam_alt.p;
        S.iters_used = G.iters_used;
        AM_HIP(c, hipEventRecord(c->kev[1], s));
        AM_HIP(c, launch_pgs_flow(s, F, fgrid, P.pgs_plain_launch != 0, (int)n_dv, bodies_view(c)));
        AM_HIP(c, hipEventRecord(c->kev[2], s));
    } else {
        AM_HIP(c, hipEventRecord(c->kev[1], s));
        AM_HIP(c, launch_pgs(s, G, grid, block, P.pgs_plain_launch != 0));
        AM_HIP(c, hipEventRecord(c->kev[2], s));
        launch_scene_finish(s, G);
    }
    if (G.per_scene) {
        c->scene_solve_valid = true;
        c->scene_iters = iterations;
    }
    launch_unpack(s, S, c->n_contacts);
    AM_HIP(c, hipGetLastError());
    r = read_scalars(c, 6, 1);
    if (r) return r;
    {
        float lms = 0;
        (void)hipEventElapsedTime(&lms, c->lev[0], c->lev[1]);
        c->ms_learned = learned ? lms : 0.0;
        int berr = 0;
        AM_HIP(c, hipMemcpy(&berr, c->d_barrier.p + 1, sizeof(int), hipMemcpyDeviceToHost));
        if (berr)
            return fail(c, AM_ERR_HIP, flow ? "PGS dataflow kernel stopped making progress (workgroups not co-resident?)"
                                            : "PGS grid barrier did not complete (blocks not co-resident)");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->kev[1], c->kev[2]);
        c->ms_pgs_kernel = ms;
        (void)hipEventElapsedTime(&ms, c->kev[0], c->kev[1]);
        c->ms_setup_kernel = ms;
    }
    int used = c->h_scalars[6];
    std::vector<unsigned long long> im(std::max(used, 1));
    if (used > 0)
        AM_HIP(c, hipMemcpy(im.data(), c->s_iter_max.p, used * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    double lc = 0;
    for (int i = 0; i < used; i++) {
        double m;
        std::memcpy(&m, &im[i], 8);
        if (main_solve) c->changes.push_back(std::log(m));
        lc = m;
    }
    c->last_iters = used;
    c->last_change = lc;
    if (iters) *iters = used;
    if (last) *last = lc;
    // training pairs only from solves that reached the tolerance: their
    // lambdas are the converged solution the network is to predict
    if (collect && used > 0 && lc < P.tolerance) {
        r = train_append(c);
        if (r) return r;
    }
    return AM_OK;
}

int pull_d(am_ctx* c, double* dst, const double* src, size_t n) {
    if (!dst || n == 0) return AM_OK;
    AM_HIP(c, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return AM_OK;
}
int pull_i(am_ctx* c, int32_t* dst, const int* src, size_t n) {
    if (!dst || n == 0) return AM_OK;
    AM_HIP(c, hipMemcpyAsync(dst, src, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    return AM_OK;
}

}  // namespace am_rt

extern "C" {

int am_scene_upload(am_ctx* c, const am_body_desc* bodies, int32_t n_bodies, const am_sphere_node* nodes,
                    int32_t n_nodes) {
    return am_scene_upload_ex(c, bodies, n_bodies, nodes, n_nodes, nullptr, nullptr, 0);
}

int am_scene_upload_ex(am_ctx* c, const am_body_desc* bodies, int32_t n_bodies, const am_sphere_node* nodes,
                       int32_t n_nodes, const am_body_desc* subs, const double* sub_b2c, int32_t n_subs) {
    int r = check_ctx(c);
    if (r) return r;
    if (n_bodies < 0 || n_nodes < 0 || (n_bodies > 0 && !bodies) || (n_nodes > 0 && !nodes))
        return fail(c, AM_ERR_INVALID, "invalid scene arrays");
    if (n_subs < 0 || (n_subs > 0 && (!subs || !sub_b2c))) return fail(c, AM_ERR_INVALID, "invalid sub-body arrays");
    try {
        c->desc.assign(bodies, bodies + n_bodies);
        c->nodes.assign(nodes, nodes + n_nodes);
        c->subs.assign(subs, subs + n_subs);
        c->sub_b2c.assign(sub_b2c, sub_b2c + (size_t)n_subs * 12);
        c->node_sub.assign(n_nodes, -1);
        c->n_ent = n_bodies;
        c->n_nodes = n_nodes;
        c->n_phys = 0;
        c->node_body.assign(n_nodes, -1);
        c->n_scenes = 1;
        c->scene_solve_valid = false;
        for (int i = 0; i < n_bodies; i++) {
            if (bodies[i].geom == AM_GEOM_COLLECTION)
                return fail(c, AM_ERR_INVALID, "collections are created by merging, not uploaded");
            if (bodies[i].scene < 0 || bodies[i].scene >= (1 << 20))
                return fail(c, AM_ERR_INVALID, "scene ids must be in [0, 2^20)");
            c->n_scenes = std::max(c->n_scenes, bodies[i].scene + 1);
        }
        // merging state starts empty
        c->total_steps = 0;
        c->n_springs = 0;
        c->springs_host.clear();
        c->kin_host.clear();
        c->impulses.clear();
        c->sim_time = 0.0;
        c->merging_event = false;
        // every body starts with an empty metric history
        c->sleep_cap = 0;
        c->sleep_fresh.clear();
        c->free_entities.clear();
        c->b2c.assign((size_t)n_bodies * 12, 0.0);
        c->n_bpc = c->n_bpc_prev = 0;
        c->ibpc.clear();
        c->n_ic = 0;
        for (int i = 0; i < n_bodies; i++) {
            const am_body_desc& d = bodies[i];
            c->n_phys = i + 1;
            if (d.geom == AM_GEOM_TREE) {
                if (d.tree_first < 0 || d.tree_count <= 0 || d.tree_first + d.tree_count > n_nodes)
                    return fail(c, AM_ERR_INVALID, "tree node range out of bounds");
                if (d.tree_count > 0xffff) return fail(c, AM_ERR_INVALID, "tree has more than 65535 nodes");
                for (int k = 0; k < d.tree_count; k++) {
                    const am_sphere_node& s = nodes[d.tree_first + k];
                    if (s.n_children < 0 || (s.n_children > 0 &&
                                             (s.first_child <= 0 || s.first_child + s.n_children > d.tree_count)))
                        return fail(c, AM_ERR_INVALID, "tree child range out of bounds");
                    c->node_body[d.tree_first + k] = i;
                }
            }
            if ((d.geom == AM_GEOM_PLANE) && !(d.flags & AM_BODY_PINNED))
                return fail(c, AM_ERR_INVALID, "planes must be pinned");
            if ((d.flags & AM_BODY_KINEMATIC) && !(d.flags & AM_BODY_PINNED))
                return fail(c, AM_ERR_INVALID, "kinematic bodies must be pinned");
            if (d.geom == AM_GEOM_COMPOSITE) {
                if (d.tree_first < 0 || d.tree_count <= 0 || d.tree_first + d.tree_count > n_subs)
                    return fail(c, AM_ERR_INVALID, "composite sub-body range out of bounds");
                if (d.tree_count > 0x7fff) return fail(c, AM_ERR_INVALID, "composite has more than 32767 sub-bodies");
                for (int q = 0; q < d.tree_count; q++) {
                    const am_body_desc& sd = subs[d.tree_first + q];
                    for (int k = 0; k < 12; k++)
                        if (!std::isfinite(sub_b2c[(size_t)(d.tree_first + q) * 12 + k]))
                            return fail(c, AM_ERR_NUMERIC, "non-finite sub-body transform");
                    if (sd.geom == AM_GEOM_BOX) continue;
                    if (sd.geom != AM_GEOM_TREE || sd.tree_first < 0 || sd.tree_count <= 0 ||
                        sd.tree_first + sd.tree_count > n_nodes || sd.tree_count > 0xffff)
                        return fail(c, AM_ERR_INVALID, "composite sub-bodies must be boxes or sphere trees");
                    for (int k = 0; k < sd.tree_count; k++) {
                        const am_sphere_node& s = nodes[sd.tree_first + k];
                        if (s.n_children < 0 || (s.n_children > 0 &&
                                                 (s.first_child <= 0 || s.first_child + s.n_children > sd.tree_count)))
                            return fail(c, AM_ERR_INVALID, "tree child range out of bounds");
                        c->node_body[sd.tree_first + k] = i;
                        c->node_sub[sd.tree_first + k] = q;
                    }
                }
            } else if (d.geom != AM_GEOM_BOX && d.geom != AM_GEOM_TREE && d.geom != AM_GEOM_PLANE) {
                return fail(c, AM_ERR_INVALID, "unknown body geometry");
            }
        }
        for (int g = 0; g < n_nodes; g++)
            if (c->node_body[g] < 0) c->node_body[g] = 0;
        c->flags.resize(n_bodies);
        std::vector<double> minv(n_bodies);
        for (int i = 0; i < n_bodies; i++) {
            c->flags[i] = bodies[i].flags;
            minv[i] = (bodies[i].flags & AM_BODY_PINNED) ? 0.0 : bodies[i].minv;
        }
        c->parent.assign(n_bodies, -1);
        c->slot.resize(n_bodies);
        c->sub.assign(n_bodies, 0);
        c->slot_entity.resize(n_bodies);
        for (int i = 0; i < n_bodies; i++) {
            c->slot[i] = i;
            c->slot_entity[i] = i;
        }
        c->member_first.assign(n_bodies, 0);
        c->member_count.assign(n_bodies, 0);
        c->members.clear();
        AM_HIP(c, upload(c, c->d_desc, c->desc));
        AM_HIP(c, upload(c, c->d_nodes, c->nodes));
        AM_HIP(c, upload(c, c->d_node_body, c->node_body));
        AM_HIP(c, upload(c, c->d_node_sub, c->node_sub));
        AM_HIP(c, upload(c, c->d_subs, c->subs));
        AM_HIP(c, upload(c, c->d_sub_b2c, c->sub_b2c));
        AM_HIP(c, upload(c, c->d_minv, minv));
        size_t n = std::max(n_bodies, 1);
        AM_HIP(c, c->d_cw.ensure((size_t)std::max(n_nodes, 1) * 3));
        AM_HIP(c, c->d_x.ensure(n * 3));
        AM_HIP(c, c->d_R.ensure(n * 9));
        AM_HIP(c, c->d_v.ensure(n * 3));
        AM_HIP(c, c->d_w.ensure(n * 3));
        AM_HIP(c, c->d_f.ensure(n * 3));
        AM_HIP(c, c->d_t.ensure(n * 3));
        AM_HIP(c, c->d_jinv.ensure(n * 9));
        AM_HIP(c, c->d_dv.ensure(n * 6));
        std::vector<double> zeros(n * 9, 0.0), ident(n * 9, 0.0);
        for (size_t i = 0; i < n; i++) ident[i * 9 + 0] = ident[i * 9 + 4] = ident[i * 9 + 8] = 1.0;
        AM_HIP(c, hipMemcpyAsync(c->d_x.p, zeros.data(), n * 24, hipMemcpyHostToDevice, c->stream));
        AM_HIP(c, hipMemcpyAsync(c->d_R.p, ident.data(), n * 72, hipMemcpyHostToDevice, c->stream));
        AM_HIP(c, hipMemcpyAsync(c->d_v.p, zeros.data(), n * 24, hipMemcpyHostToDevice, c->stream));
        AM_HIP(c, hipMemcpyAsync(c->d_w.p, zeros.data(), n * 24, hipMemcpyHostToDevice, c->stream));
        AM_HIP(c, hipMemcpyAsync(c->d_f.p, zeros.data(), n * 24, hipMemcpyHostToDevice, c->stream));
        AM_HIP(c, hipMemcpyAsync(c->d_t.p, zeros.data(), n * 24, hipMemcpyHostToDevice, c->stream));
        AM_HIP(c, hipMemcpyAsync(c->d_dv.p, zeros.data(), n * 48, hipMemcpyHostToDevice, c->stream));
        r = rebuild_structure(c);
        if (r) return r;
        r = refresh_jinv(c);
        if (r) return r;
        r = ensure_contact_capacity(c, std::max(1024, 4 * n_bodies));
        if (r) return r;
        r = ensure_old_capacity(c, 1024);
        if (r) return r;
        c->n_contacts = 0;
        c->n_old = 0;
        c->n_pairs = 0;
        c->n_groups = 0;
        c->layout_valid = false;
        AM_HIP(c, hipStreamSynchronize(c->stream));
    } catch (const std::bad_alloc&) {
        return fail(c, AM_ERR_INVALID, "host allocation failed");
    }
    return AM_OK;
}

int am_structure_pull(am_ctx* c, int32_t* parent, int32_t cap_entities, int32_t* top, int32_t cap_top,
                      int32_t* n_entities, int32_t* n_top) {
    int r = check_ctx(c);
    if (r) return r;
    if (n_entities) *n_entities = c->n_ent;
    if (n_top) *n_top = (int)c->slot_entity.size();
    if (parent)
        for (int e = 0; e < std::min(cap_entities, c->n_ent); e++) parent[e] = c->parent[e];
    if (top)
        for (int s_ = 0; s_ < std::min(cap_top, (int)c->slot_entity.size()); s_++) top[s_] = c->slot_entity[s_];
    return AM_OK;
}

int am_entities_pull(am_ctx* c, am_body_desc* out, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || n > c->n_ent || (n > 0 && !out)) return fail(c, AM_ERR_INVALID, "bad entity buffer");
    for (int e = 0; e < n; e++) out[e] = c->desc[e];
    return AM_OK;
}

int am_state_push(am_ctx* c, const double* x, const double* R, const double* v, const double* w, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || n > c->n_ent) return fail(c, AM_ERR_INVALID, "state size exceeds the entity count");
    size_t m = (size_t)n;
    for (size_t i = 0; i < m * 3; i++) {
        if ((x && !std::isfinite(x[i])) || (v && !std::isfinite(v[i])) || (w && !std::isfinite(w[i])))
            return fail(c, AM_ERR_NUMERIC, "non-finite body state");
    }
    if (R)
        for (size_t i = 0; i < m * 9; i++)
            if (!std::isfinite(R[i])) return fail(c, AM_ERR_NUMERIC, "non-finite rotation");
    if (m == 0) return AM_OK;
    if (x) AM_HIP(c, hipMemcpyAsync(c->d_x.p, x, m * 24, hipMemcpyHostToDevice, c->stream));
    if (R) AM_HIP(c, hipMemcpyAsync(c->d_R.p, R, m * 72, hipMemcpyHostToDevice, c->stream));
    if (v) AM_HIP(c, hipMemcpyAsync(c->d_v.p, v, m * 24, hipMemcpyHostToDevice, c->stream));
    if (w) AM_HIP(c, hipMemcpyAsync(c->d_w.p, w, m * 24, hipMemcpyHostToDevice, c->stream));
    r = refresh_jinv(c);
    if (r) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_state_pull(am_ctx* c, double* x, double* R, double* v, double* w, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || n > c->n_ent) return fail(c, AM_ERR_INVALID, "state size exceeds the entity count");
    size_t m = (size_t)n;
    if ((r = pull_d(c, x, c->d_x.p, m * 3))) return r;
    if ((r = pull_d(c, R, c->d_R.p, m * 9))) return r;
    if ((r = pull_d(c, v, c->d_v.p, m * 3))) return r;
    if ((r = pull_d(c, w, c->d_w.p, m * 3))) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_jinv_pull(am_ctx* c, double* jinv, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || n > c->n_ent || (n > 0 && !jinv)) return fail(c, AM_ERR_INVALID, "bad jinv buffer");
    if ((r = pull_d(c, jinv, c->d_jinv.p, (size_t)n * 9))) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_forces_push(am_ctx* c, const double* force, const double* torque, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || n > c->n_ent) return fail(c, AM_ERR_INVALID, "force size exceeds the entity count");
    if (force && n) AM_HIP(c, hipMemcpyAsync(c->d_f.p, force, (size_t)n * 24, hipMemcpyHostToDevice, c->stream));
    if (torque && n) AM_HIP(c, hipMemcpyAsync(c->d_t.p, torque, (size_t)n * 24, hipMemcpyHostToDevice, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_forces_pull(am_ctx* c, double* force, double* torque, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || n > c->n_ent) return fail(c, AM_ERR_INVALID, "force size exceeds the entity count");
    if ((r = pull_d(c, force, c->d_f.p, (size_t)n * 3))) return r;
    if ((r = pull_d(c, torque, c->d_t.p, (size_t)n * 3))) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_clear_bodies(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    launch_clear(c->stream, bodies_view(c));
    AM_HIP(c, hipGetLastError());
    // RigidCollection.clearBodies (RigidCollection.java:98-103)
    if (has_collections(c)) return merge_push_down(c, false);
    return AM_OK;
}

int am_springs_set(am_ctx* c, const am_spring* springs, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || (n > 0 && !springs)) return fail(c, AM_ERR_INVALID, "bad spring buffer");
    const int nb = c->n_phys;
    for (int32_t i = 0; i < n; i++) {
        const am_spring& sp = springs[i];
        if (sp.type < AM_SPRING_ZERO || sp.type > AM_SPRING_BODYBODY)
            return fail(c, AM_ERR_INVALID, "spring " + std::to_string(i) + ": unknown type");
        if (sp.body1 < 0 || sp.body1 >= nb)
            return fail(c, AM_ERR_INVALID, "spring " + std::to_string(i) + ": body1 out of range");
        if (sp.type == AM_SPRING_BODYBODY && (sp.body2 < 0 || sp.body2 >= nb))
            return fail(c, AM_ERR_INVALID, "spring " + std::to_string(i) + ": body2 out of range");
    }
    AM_HIP(c, c->d_springs.ensure(std::max(n, 1)));
    if (n > 0)
        AM_HIP(c, hipMemcpyAsync(c->d_springs.p, springs, (size_t)n * sizeof(am_spring), hipMemcpyHostToDevice,
                                 c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    c->springs_host.assign(springs, springs + n);
    c->n_springs = n;
    return AM_OK;
}

int am_apply_spring_forces(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    return apply_springs(c);
}

int am_wake(am_ctx* c, int32_t* n_woken) {
    int r = check_ctx(c);
    if (r) return r;
    int n = 0;
    r = sleep_wake(c, &n);
    if (n_woken) *n_woken = n;
    return r;
}

int am_sleep(am_ctx* c, int32_t* n_asleep) {
    int r = check_ctx(c);
    if (r) return r;
    int n = 0;
    r = sleep_sleep(c, &n);
    if (n_asleep) *n_asleep = n;
    return r;
}

int am_unmerge_all(am_ctx* c, int32_t* n_dissolved) {
    int r = check_ctx(c);
    if (r) return r;
    int n = 0;
    r = merge_unmerge_all(c, &n);
    if (n_dissolved) *n_dissolved = n;
    return r;
}

int am_wake_all(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    return sleep_wake_all(c);
}

int am_sleep_state_pull(am_ctx* c, uint8_t* sleeping, int32_t* hist_n, double* hist, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || n > c->n_ent) return fail(c, AM_ERR_INVALID, "entity count out of range");
    if (sleeping)
        for (int e = 0; e < n; e++) sleeping[e] = (c->flags[e] & AM_BODY_SLEEPING) ? 1 : 0;
    if (!hist_n && !hist) return AM_OK;
    if ((r = ensure_sleep_state(c))) return r;
    if (hist_n) r = pull_i(c, hist_n, c->d_sleep_n.p, n);
    if (!r && hist) r = pull_d(c, hist, c->d_sleep_hist.p, (size_t)n * am::kSleepHistStride);
    if (r) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_apply_gravity(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    return apply_gravity(c);
}

int am_impulse_add(am_ctx* c, int32_t body, const double* point_w, const double* force) {
    int r = check_ctx(c);
    if (r) return r;
    if (body < 0 || body >= c->n_phys || !point_w || !force) return fail(c, AM_ERR_INVALID, "bad impulse");
    am::ImpulseRec q{};
    q.body = body;
    for (int a = 0; a < 3; a++) {
        q.p[a] = point_w[a];
        q.f[a] = force[a];
        if (!std::isfinite(q.p[a]) || !std::isfinite(q.f[a])) return fail(c, AM_ERR_NUMERIC, "non-finite impulse");
    }
    c->impulses.push_back(q);
    return AM_OK;
}

int am_apply_impulses(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    return apply_impulses(c);
}

int am_apply_viscous_decay(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    return apply_viscous_decay(c);
}

int am_launch_shape_pull(am_ctx* c, int32_t* grid, int32_t* block) {
    int r = check_ctx(c);
    if (r) return r;
    if (!grid || !block) return fail(c, AM_ERR_INVALID, "null argument");
    *grid = c->last_grid;
    *block = c->last_block;
    return AM_OK;
}

int am_scene_iterations_pull(am_ctx* c, int32_t* used, double* last, int32_t capacity, int32_t* n_scenes) {
    int r = check_ctx(c);
    if (r) return r;
    if (n_scenes) *n_scenes = c->n_scenes;
    if (!used && !last) return AM_OK;
    if (capacity < c->n_scenes) return fail(c, AM_ERR_CAPACITY, "per-scene arrays too small");
    if (!c->scene_solve_valid) return fail(c, AM_ERR_STATE, "the last main solve did not stop scenes one by one");
    const int ns = c->n_scenes;
    std::vector<int> u(ns);
    AM_HIP(c, hipMemcpyAsync(u.data(), c->scene_used.p, (size_t)ns * 4, hipMemcpyDeviceToHost, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    const int iters = c->scene_iters;
    for (int sc = 0; sc < ns; sc++) {
        if (used) used[sc] = u[sc];
        if (last) {
            unsigned long long b = 0;
            if (u[sc] > 0)
                AM_HIP(c, hipMemcpy(&b, c->scene_max.p + (size_t)sc * iters + u[sc] - 1, 8, hipMemcpyDeviceToHost));
            std::memcpy(&last[sc], &b, 8);
        }
    }
    return AM_OK;
}

int am_kinematic_set(am_ctx* c, const am_kinematic* k, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || (n > 0 && !k)) return fail(c, AM_ERR_INVALID, "bad kinematic table");
    std::vector<char> seen(c->n_phys, 0);
    for (int i = 0; i < n; i++) {
        const int b = k[i].body;
        if (b < 0 || b >= c->n_phys) return fail(c, AM_ERR_INVALID, "kinematic body out of range");
        const int want = AM_BODY_PINNED | AM_BODY_KINEMATIC;
        if ((c->desc[b].flags & want) != want)
            return fail(c, AM_ERR_INVALID, "a kinematic body must be flagged AM_BODY_PINNED | AM_BODY_KINEMATIC");
        if (seen[b]) return fail(c, AM_ERR_INVALID, "a body has two prescribed motions");
        seen[b] = 1;
        const double* vals[] = {k[i].x0, k[i].dir};
        for (const double* v : vals)
            for (int a = 0; a < 3; a++)
                if (!std::isfinite(v[a])) return fail(c, AM_ERR_NUMERIC, "non-finite prescribed motion");
        if (!std::isfinite(k[i].amplitude) || !std::isfinite(k[i].omega) || !std::isfinite(k[i].phase))
            return fail(c, AM_ERR_NUMERIC, "non-finite prescribed motion");
    }
    try {
        c->kin_host.assign(k, k + n);
    } catch (const std::bad_alloc&) {
        return fail(c, AM_ERR_INVALID, "host allocation failed");
    }
    AM_HIP(c, upload(c, c->d_kin, c->kin_host));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_kinematic_update(am_ctx* c, double t) {
    int r = check_ctx(c);
    if (r) return r;
    if (!std::isfinite(t)) return fail(c, AM_ERR_NUMERIC, "non-finite time");
    return kinematic_update(c, t);
}

int am_sim_time_get(am_ctx* c, double* t) {
    int r = check_ctx(c);
    if (r) return r;
    if (!t) return fail(c, AM_ERR_INVALID, "null time");
    *t = c->sim_time;
    return AM_OK;
}

int am_sim_time_set(am_ctx* c, double t) {
    int r = check_ctx(c);
    if (r) return r;
    if (!std::isfinite(t)) return fail(c, AM_ERR_NUMERIC, "non-finite time");
    c->sim_time = t;
    return AM_OK;
}

}  // extern "C"

namespace am_rt {
int apply_springs(am_ctx* c) {
    launch_springs(c->stream, bodies_view(c), c->d_springs.p, c->n_springs, c->params.spring_stiffness_mod,
                   c->params.spring_damping_mod);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

// RigidBodySystem.applyExternalForces (RigidBodySystem.java:234-257): gravity,
// then the springs
int apply_external_forces(am_ctx* c) {
    int r = apply_gravity(c);
    if (r) return r;
    r = apply_springs(c);
    if (r) return r;
    return apply_impulses(c);
}

// RigidBodySystem.applyImpulse (:262-267): the held forces, once
int apply_impulses(am_ctx* c) {
    if (!c->impulses.empty()) {
        AM_HIP(c, upload(c, c->d_imp, c->impulses));
        launch_impulses(c->stream, bodies_view(c), c->d_imp.p, (int)c->impulses.size());
        AM_HIP(c, hipGetLastError());
        // the upload reads the host vector asynchronously: wait before it is cleared
        AM_HIP(c, hipStreamSynchronize(c->stream));
        c->impulses.clear();
    }
    return AM_OK;
}

// RigidBodySystem.applyViscousDecay (:173, :428-435); scaling by 1 changes no bit
int apply_viscous_decay(am_ctx* c) {
    const am_params& P = c->params;
    if (P.viscous_linear_decay == 1.0 && P.viscous_angular_decay == 1.0) return AM_OK;
    launch_viscous(c->stream, bodies_view(c), c->d_slot_entity.p, (int)c->slot_entity.size(),
                   P.viscous_linear_decay, P.viscous_angular_decay);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

int kinematic_update(am_ctx* c, double t) {
    if (c->kin_host.empty()) return AM_OK;
    launch_kinematic(c->stream, bodies_view(c), c->d_kin.p, (int)c->kin_host.size(), t);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

int apply_gravity(am_ctx* c) {
    if (!c->params.use_gravity) return AM_OK;
    double theta = c->params.gravity_angle_deg / 180.0 * M_PI;
    double g = c->params.gravity_amount;
    launch_gravity(c->stream, bodies_view(c), c->d_gravity.p, (int)c->gravity_targets.size(), g * std::cos(theta),
                   g * std::sin(theta), 0.0);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}
}  // namespace am_rt

extern "C" {

int am_deltav_pull(am_ctx* c, double* dv, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n < 0 || n > c->n_ent || (n > 0 && !dv)) return fail(c, AM_ERR_INVALID, "bad deltaV buffer");
    if ((r = pull_d(c, dv, c->d_dv.p, (size_t)n * 6))) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_update_contacts_map(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    const int n = c->n_contacts;
    r = ensure_old_capacity(c, n);
    if (r) return r;
    hipStream_t s = c->stream;
    if (n > 0) {
        AM_HIP(c, hipMemcpyAsync(c->o_b1.p, c->c_b1.p, n * 4, hipMemcpyDeviceToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->o_b2.p, c->c_b2.p, n * 4, hipMemcpyDeviceToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->o_bv1.p, c->c_bv1.p, n * 4, hipMemcpyDeviceToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->o_bv2.p, c->c_bv2.p, n * 4, hipMemcpyDeviceToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->o_info.p, c->c_info.p, n * 4, hipMemcpyDeviceToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->o_pB1.p, c->c_pB1.p, (size_t)n * 24, hipMemcpyDeviceToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->o_lam.p, c->c_lam.p, (size_t)n * 24, hipMemcpyDeviceToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->o_viol.p, c->c_viol.p, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
        AM_HIP(c, hipMemcpyAsync(c->o_prev.p, c->c_prev.p, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    }
    c->n_old = n;
    fill_int(s, c->o_table.p, c->table_mask + 1, -1);
    launch_map_insert(s, old_view(c), c->o_table.p);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

int am_detect(am_ctx* c, double dt, int32_t* n_contacts) {
    (void)dt;
    int r = check_ctx(c);
    if (r) return r;
    r = do_detect(c);
    if (r) return r;
    if (n_contacts) *n_contacts = c->n_contacts;
    return AM_OK;
}

int am_contacts_push(am_ctx* c, int32_t n, const int32_t* body1, const int32_t* body2, const double* point_w,
                     const double* normal_w, const double* violation, const int32_t* info, const int32_t* bv1,
                     const int32_t* bv2) {
    int r = check_ctx(c);
    if (r) return r;
    c->sweep_order_n = -1;
    if (n < 0 || (n > 0 && (!body1 || !body2 || !point_w || !normal_w || !violation)))
        return fail(c, AM_ERR_INVALID, "invalid contact arrays");
    for (int k = 0; k < n; k++) {
        if (body1[k] < 0 || body1[k] >= c->n_ent || body2[k] < 0 || body2[k] >= c->n_ent || body1[k] == body2[k])
            return fail(c, AM_ERR_INVALID, "contact body index out of range");
        if ((c->flags[body1[k]] & AM_BODY_PINNED) && (c->flags[body2[k]] & AM_BODY_PINNED))
            return fail(c, AM_ERR_INVALID, "contact between two pinned bodies");
    }
    r = ensure_contact_capacity(c, n);
    if (r) return r;
    std::vector<int32_t> zi(std::max(n, 1), 0), nb(std::max(n, 1), AM_BV_NONE);
    DBuf<int> t_b1, t_b2, t_info, t_bv1, t_bv2;
    DBuf<double> t_p, t_n, t_v;
    hipStream_t s = c->stream;
    if (n > 0) {
        AM_HIP(c, t_b1.ensure(n));
        AM_HIP(c, t_b2.ensure(n));
        AM_HIP(c, t_info.ensure(n));
        AM_HIP(c, t_bv1.ensure(n));
        AM_HIP(c, t_bv2.ensure(n));
        AM_HIP(c, t_p.ensure((size_t)n * 3));
        AM_HIP(c, t_n.ensure((size_t)n * 3));
        AM_HIP(c, t_v.ensure(n));
        AM_HIP(c, hipMemcpyAsync(t_b1.p, body1, n * 4, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(t_b2.p, body2, n * 4, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(t_info.p, info ? info : zi.data(), n * 4, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(t_bv1.p, bv1 ? bv1 : nb.data(), n * 4, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(t_bv2.p, bv2 ? bv2 : nb.data(), n * 4, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(t_p.p, point_w, (size_t)n * 24, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(t_n.p, normal_w, (size_t)n * 24, hipMemcpyHostToDevice, s));
        AM_HIP(c, hipMemcpyAsync(t_v.p, violation, (size_t)n * 8, hipMemcpyHostToDevice, s));
        fill_int(s, c->d_scalars.p + 4, 1, 0);
        launch_contacts_set(s, bodies_view(c), n, t_b1.p, t_b2.p, t_p.p, t_n.p, t_v.p, t_info.p, t_bv1.p, t_bv2.p,
                            contacts_view(c), c->d_scalars.p + 4);
        AM_HIP(c, hipGetLastError());
    }
    c->n_contacts = n;
    fill_int(s, c->d_scalars.p + 1, 1, n);
    r = groups_after_contacts(c);
    if (r) return r;
    AM_HIP(c, hipStreamSynchronize(s));
    return AM_OK;
}

}  // extern "C"

namespace am_rt {
int pull_contacts(am_ctx* c, const ContactsView& V, int count, am_contact_arrays* o, int32_t capacity) {
    int r;
    size_t n = (size_t)std::max(0, std::min<int>(capacity, count));
    if ((r = pull_i(c, o->body1, V.b1, n))) return r;
    if ((r = pull_i(c, o->body2, V.b2, n))) return r;
    if ((r = pull_i(c, o->bv1, V.bv1, n))) return r;
    if ((r = pull_i(c, o->bv2, V.bv2, n))) return r;
    if ((r = pull_i(c, o->info, V.info, n))) return r;
    if ((r = pull_d(c, o->point_b1, V.pB1, n * 3))) return r;
    if ((r = pull_d(c, o->normal_b1, V.nB1, n * 3))) return r;
    if ((r = pull_d(c, o->tangent1_b1, V.t1B1, n * 3))) return r;
    if ((r = pull_d(c, o->tangent2_b1, V.t2B1, n * 3))) return r;
    if ((r = pull_d(c, o->violation, V.viol, n))) return r;
    if ((r = pull_d(c, o->prev_violation, V.prev_viol, n))) return r;
    if ((r = pull_d(c, o->jacobian, V.J, n * 36))) return r;
    if ((r = pull_d(c, o->lambda, V.lam, n * 3))) return r;
    if ((r = pull_d(c, o->lambda_warm, V.lamw, n * 3))) return r;
    if ((r = pull_d(c, o->b, V.b, n * 3))) return r;
    if ((r = pull_d(c, o->diag, V.D, n * 3))) return r;
    if ((r = pull_i(c, o->state, V.state, n))) return r;
    if ((r = pull_i(c, o->new_this_step, V.is_new, n))) return r;
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}
}  // namespace am_rt

extern "C" {

int am_contacts_pull(am_ctx* c, am_contact_arrays* o, int32_t capacity, int32_t* n_out) {
    int r = check_ctx(c);
    if (r) return r;
    if (n_out) *n_out = c->n_contacts;
    if (!o) return AM_OK;
    return pull_contacts(c, contacts_view(c), c->n_contacts, o, capacity);
}

int am_internal_contacts_pull(am_ctx* c, am_contact_arrays* o, int32_t capacity, int32_t* n_out) {
    int r = check_ctx(c);
    if (r) return r;
    if (n_out) *n_out = c->n_ic;
    if (!o || c->n_ic == 0) return AM_OK;
    return pull_contacts(c, internal_view(c), c->n_ic, o, capacity);
}

int am_lambdas_push(am_ctx* c, const double* lambda, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n != c->n_contacts || (n > 0 && !lambda)) return fail(c, AM_ERR_INVALID, "lambda count mismatch");
    if (n > 0) {
        AM_HIP(c, hipMemcpyAsync(c->c_lam.p, lambda, (size_t)n * 24, hipMemcpyHostToDevice, c->stream));
        AM_HIP(c, hipMemcpyAsync(c->c_lamw.p, lambda, (size_t)n * 24, hipMemcpyHostToDevice, c->stream));
    }
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_warm_start(am_ctx* c, int32_t post, int32_t* bad, int32_t* repaired) {
    int r = check_ctx(c);
    if (r) return r;
    fill_int(c->stream, c->d_scalars.p + 7, 2, 0);
    if (c->n_contacts > 0) {
        fill_int(c->stream, c->d_scalars.p + 2, 1, c->n_groups);
        launch_warm_start(c->stream, bodies_view(c), contacts_view(c), old_view(c), c->d_scalars.p + 2, c->n_groups,
                          c->g_start.p, post, c->d_scalars.p + 7);
        AM_HIP(c, hipGetLastError());
    }
    r = read_scalars(c, 7, 2);
    if (r) return r;
    if (bad) *bad = c->h_scalars[7];
    if (repaired) *repaired = c->h_scalars[8];
    return AM_OK;
}

int am_redo_warm_start(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    launch_redo_warm(c->stream, contacts_view(c), c->d_scalars.p + 1, c->n_contacts, 0);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

int am_no_warm_start(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    launch_redo_warm(c->stream, contacts_view(c), c->d_scalars.p + 1, c->n_contacts, 1);
    AM_HIP(c, hipGetLastError());
    return AM_OK;
}

int am_solve(am_ctx* c, double dt, int32_t post, int32_t* iterations_used, double* last_change) {
    int r = check_ctx(c);
    if (r) return r;
    return do_solve(c, dt, post, true, iterations_used, last_change);
}

int am_autotune(am_ctx* c, double dt, int32_t sweeps, int32_t reps, int32_t apply, am_tune_trial* trials,
                int32_t capacity, int32_t* n_trials, int32_t* best) {
    int r = check_ctx(c);
    if (r) return r;
    if (sweeps < 1 || sweeps > 100000 || reps < 1 || reps > 100)
        return fail(c, AM_ERR_INVALID, "sweeps must be in 1..100000 and reps in 1..100");
    if (capacity < 0 || (capacity > 0 && !trials)) return fail(c, AM_ERR_INVALID, "null trials with capacity > 0");
    if (c->n_contacts == 0) return fail(c, AM_ERR_STATE, "no contacts to tune on: call am_detect or am_step first");
    if ((r = build_layout(c))) return r;
    hipStream_t s = c->stream;
    const size_t nc3 = (size_t)c->n_contacts * 3, nc = (size_t)c->n_contacts, n_dv = (size_t)c->n_ent * 6;

    // what a solve overwrites, restored before every trial and at the end
    DBuf<double> lam0, b0, D0, dv0, start0;
    DBuf<int> st0;
    AM_HIP(c, lam0.ensure(nc3));
    AM_HIP(c, b0.ensure(nc3));
    AM_HIP(c, D0.ensure(nc3));
    AM_HIP(c, start0.ensure(nc3));
    AM_HIP(c, dv0.ensure(n_dv));
    AM_HIP(c, st0.ensure(nc));
    AM_HIP(c, c->s_start.ensure(nc3));
    auto d2d = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
    };
    AM_HIP(c, d2d(lam0.p, c->c_lam.p, nc3 * 8));
    AM_HIP(c, d2d(b0.p, c->c_b.p, nc3 * 8));
    AM_HIP(c, d2d(D0.p, c->c_D.p, nc3 * 8));
    AM_HIP(c, d2d(start0.p, c->s_start.p, nc3 * 8));
    AM_HIP(c, d2d(dv0.p, c->d_dv.p, n_dv * 8));
    AM_HIP(c, d2d(st0.p, c->c_state.p, nc * 4));
    const am_params saved = c->params;
    const int last_iters = c->last_iters, sweep_order_n = c->sweep_order_n, used_learned = c->used_learned;
    const double last_change = c->last_change, ms_pgs = c->ms_pgs_kernel, ms_setup = c->ms_setup_kernel,
                 ms_learned = c->ms_learned, res_plain = c->res_plain, res_learned = c->res_learned;
    const std::vector<double> changes = c->changes;
    auto restore_inputs = [&]() -> hipError_t {
        hipError_t e = d2d(c->c_lam.p, lam0.p, nc3 * 8);
        if (e == hipSuccess) e = d2d(c->d_dv.p, dv0.p, n_dv * 8);
        return e;
    };

    // candidate shapes: the dataflow schedule at several wave counts, the
    // coloured schedule at each workgroup size and several grids; shapes that
    // clamp to one already listed are timed once
    struct Cand {
        int sched, grid, block, grid_used, block_used;
        double ms;
    };
    std::vector<Cand> cands;
    auto add = [&](int sched, int grid, int block) {
        am_params P = saved;
        P.pgs_schedule = sched;
        P.pgs_grid = grid;
        P.pgs_block = block;
        int g = 1, b = 64;
        pgs_shape(c, P, &g, &b);
        for (const Cand& k : cands)
            if (k.sched == sched && k.grid_used == g && k.block_used == b) return;
        cands.push_back({sched, grid, block, g, b, 0.0});
    };
    for (int g : {0, 64, 128, 192, 256, 384, 512, 768, 1024, 1280, 1536, 2048}) add(0, g, 0);
    for (int b : {0, 64, 128})
        for (int g : {0, 32, 64, 128, 256, 512}) add(1, g, b);

    std::vector<double> ref_lam(nc3), ref_dv(n_dv), got_lam(nc3), got_dv(n_dv);
    int err = AM_OK;
    for (size_t k = 0; k < cands.size() && err == AM_OK; k++) {
        am_params& P = c->params;
        P = saved;
        P.pgs_schedule = cands[k].sched;
        P.pgs_grid = cands[k].grid;
        P.pgs_block = cands[k].block;
        P.iterations = sweeps;
        P.tolerance = 0.0;  // no early exit: every trial runs the same sweeps
        P.learned_warm_start = 0;
        P.collect_training = 0;
        double best_ms = 1e300;
        for (int rep = 0; rep < reps && err == AM_OK; rep++) {
            hipError_t e = restore_inputs();
            if (e != hipSuccess) {
                err = fail(c, AM_ERR_HIP, hipGetErrorString(e));
                break;
            }
            int it = 0;
            if ((err = do_solve(c, dt, 0, true, &it, nullptr))) break;
            if (it != sweeps) {
                err = fail(c, AM_ERR_MISMATCH, "a tuning trial did not run the requested sweeps");
                break;
            }
            best_ms = std::min(best_ms, c->ms_pgs_kernel);
        }
        if (err) break;
        cands[k].ms = best_ms / sweeps;
        std::vector<double>& L = k == 0 ? ref_lam : got_lam;
        std::vector<double>& V = k == 0 ? ref_dv : got_dv;
        hipError_t e = hipMemcpyAsync(L.data(), c->c_lam.p, nc3 * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(V.data(), c->d_dv.p, n_dv * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            err = fail(c, AM_ERR_HIP, hipGetErrorString(e));
            break;
        }
        if (k > 0 && (std::memcmp(L.data(), ref_lam.data(), nc3 * 8) != 0 ||
                      std::memcmp(V.data(), ref_dv.data(), n_dv * 8) != 0)) {
            char msg[160];
            std::snprintf(msg, sizeof msg, "launch shape (schedule %d, grid %d, block %d) changed the solve's bits",
                          cands[k].sched, cands[k].grid_used, cands[k].block_used);
            err = fail(c, AM_ERR_MISMATCH, msg);
        }
    }

    // put back what the trials overwrote, also after a failed trial
    c->params = saved;
    hipError_t e = restore_inputs();
    if (e == hipSuccess) e = d2d(c->c_b.p, b0.p, nc3 * 8);
    if (e == hipSuccess) e = d2d(c->c_D.p, D0.p, nc3 * 8);
    if (e == hipSuccess) e = d2d(c->s_start.p, start0.p, nc3 * 8);
    if (e == hipSuccess) e = d2d(c->c_state.p, st0.p, nc * 4);
    hipError_t e2 = hipStreamSynchronize(s);  // before the saved copies are freed
    c->last_iters = last_iters;
    c->sweep_order_n = sweep_order_n;
    c->used_learned = used_learned;
    c->last_change = last_change;
    c->ms_pgs_kernel = ms_pgs;
    c->ms_setup_kernel = ms_setup;
    c->ms_learned = ms_learned;
    c->res_plain = res_plain;
    c->res_learned = res_learned;
    c->changes = changes;
    if (err) return err;
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail(c, AM_ERR_HIP, hipGetErrorString(e));

    // the fastest shape, unless the automatic dataflow shape (the first) is
    // within 2 % of it: equal shapes differ by about 1 % from run to run
    int kb = 0;
    for (size_t k = 1; k < cands.size(); k++)
        if (cands[k].ms < cands[kb].ms) kb = (int)k;
    if (cands[0].ms <= cands[kb].ms * 1.02) kb = 0;
    for (int k = 0; k < std::min(capacity, (int)cands.size()); k++) {
        am_tune_trial& t = trials[k];
        t.pgs_schedule = cands[k].sched;
        t.pgs_grid = cands[k].grid;
        t.pgs_block = cands[k].block;
        t.grid_used = cands[k].grid_used;
        t.block_used = cands[k].block_used;
        t.reserved = 0;
        t.ms_per_sweep = cands[k].ms;
    }
    if (n_trials) *n_trials = (int)cands.size();
    if (best) *best = kb;
    if (apply) {
        c->params.pgs_schedule = cands[kb].sched;
        c->params.pgs_grid = cands[kb].grid;
        c->params.pgs_block = cands[kb].block;
    }
    return AM_OK;
}

int am_update_in_collections(am_ctx* c, double dt, int32_t* iterations_used) {
    int r = check_ctx(c);
    if (r) return r;
    int it = 0;
    r = merge_update_in_collections(c, dt, &it);
    if (iterations_used) *iterations_used = it;
    return r;
}

int am_accumulate_for_unmerging(am_ctx* c, double dt) {
    int r = check_ctx(c);
    if (r) return r;
    return merge_accumulate_unmerging(c, dt);
}

int am_unmerge(am_ctx* c, double dt, int32_t* n_cut) {
    int r = check_ctx(c);
    if (r) return r;
    int n = 0;
    r = merge_unmerge(c, dt, &n);
    if (n_cut) *n_cut = n;
    return r;
}

int am_clear_top_level(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    launch_clear_targets(c->stream, bodies_view(c), c->d_slot_entity.p, (int)c->slot_entity.size());
    AM_HIP(c, hipGetLastError());
    return apply_external_forces(c);
}

int am_clear_body_pair_contacts(am_ctx* c) {
    int r = check_ctx(c);
    if (r) return r;
    return merge_body_pairs(c);
}

int am_accumulate_for_merging(am_ctx* c, double dt) {
    int r = check_ctx(c);
    if (r) return r;
    return merge_accumulate_merging(c, dt);
}

int am_merge(am_ctx* c, int32_t* n_merged) {
    int r = check_ctx(c);
    if (r) return r;
    int n = 0;
    r = merge_merge(c, &n);
    if (n_merged) *n_merged = n;
    return r;
}

int am_body_pairs_pull(am_ctx* c, am_body_pair* out, int32_t capacity, int32_t* n_external, int32_t* n_internal) {
    int r = check_ctx(c);
    if (r) return r;
    if (n_external) *n_external = c->n_bpc;
    if (n_internal) *n_internal = (int)c->ibpc.size();
    if (!out || capacity <= 0) return AM_OK;
    std::vector<BpcRec> T(c->n_bpc);
    std::vector<IbpcRec> I(c->ibpc.size());
    if (!T.empty())
        AM_HIP(c, hipMemcpyAsync(T.data(), c->d_bpc.p, T.size() * sizeof(BpcRec), hipMemcpyDeviceToHost, c->stream));
    if (!I.empty())
        AM_HIP(c, hipMemcpyAsync(I.data(), c->d_ibpc.p, I.size() * sizeof(IbpcRec), hipMemcpyDeviceToHost, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    int k = 0;
    for (const auto& t : T) {
        if (k >= capacity) break;
        am_body_pair& o = out[k++];
        std::memset(&o, 0, sizeof o);
        o.body1 = t.b1;
        o.body2 = t.b2;
        o.in_collection = 0;
        o.n_contacts = t.n_active;
        o.hist_n = t.hist_n;
        o.state_n = t.state_n;
        o.flag = t.flag;
        o.collection = -1;
        for (int q = 0; q < kMaxAccum; q++) {
            o.hist[q] = t.hist[q];
            o.states[q] = q < t.state_n ? (t.states >> (2 * q)) & 3 : 0;
        }
    }
    for (const auto& t : I) {
        if (k >= capacity) break;
        am_body_pair& o = out[k++];
        std::memset(&o, 0, sizeof o);
        o.body1 = t.b1;
        o.body2 = t.b2;
        o.in_collection = 1;
        o.n_contacts = t.count;
        o.hist_n = t.hist_n;
        o.flag = t.flag;
        o.collection = t.coll;
    }
    return AM_OK;
}

int am_solve_order_pull(am_ctx* c, int32_t* order, int32_t capacity, int32_t* n) {
    int r = check_ctx(c);
    if (r) return r;
    if (c->sweep_order_n >= 0) {  // the single sweep inside collections ran last
        if (n) *n = c->sweep_order_n;
        if (order) {
            if ((r = pull_i(c, order, c->sweep_order.p, (size_t)std::min(capacity, c->sweep_order_n)))) return r;
            AM_HIP(c, hipStreamSynchronize(c->stream));
        }
        return AM_OK;
    }
    if (c->n_contacts > 0) {
        r = build_layout(c);
        if (r) return r;
    }
    if (n) *n = c->n_contacts;
    if (order) {
        if ((r = pull_i(c, order, c->s_order.p, (size_t)std::min(capacity, c->n_contacts)))) return r;
        AM_HIP(c, hipStreamSynchronize(c->stream));
    }
    return AM_OK;
}

int am_colours_pull(am_ctx* c, int32_t* colour_of_contact, int32_t capacity, int32_t* n_colours) {
    int r = check_ctx(c);
    if (r) return r;
    if (c->n_contacts > 0) {
        r = build_layout(c);
        if (r) return r;
    }
    if (n_colours) *n_colours = c->n_colours;
    if (colour_of_contact && c->n_contacts > 0) {
        int cs[kMaxColours + 2];
        std::vector<int> order(c->n_contacts), pos(c->n_groups + 1);
        AM_HIP(c, hipMemcpy(cs, c->s_colour_start.p, sizeof cs, hipMemcpyDeviceToHost));
        AM_HIP(c, hipMemcpy(order.data(), c->s_order.p, order.size() * 4, hipMemcpyDeviceToHost));
        AM_HIP(c, hipMemcpy(pos.data(), c->g_pos_start.p, pos.size() * 4, hipMemcpyDeviceToHost));
        for (int col = 0; col <= kMaxColours; col++)
            for (int q = cs[col]; q < cs[col + 1]; q++)
                for (int k = pos[q]; k < pos[q + 1]; k++)
                    if (order[k] < capacity) colour_of_contact[order[k]] = col;
    }
    return AM_OK;
}

int am_changes_pull(am_ctx* c, double* changes, int32_t capacity, int32_t* n) {
    if (!c) return AM_ERR_INVALID;
    if (n) *n = (int32_t)c->changes.size();
    if (changes)
        for (int i = 0; i < (int)c->changes.size() && i < capacity; i++) changes[i] = c->changes[i];
    return AM_OK;
}

int am_advance_velocities(am_ctx* c, double dt) {
    int r = check_ctx(c);
    if (r) return r;
    launch_advance(c->stream, bodies_view(c), c->d_integrate.p, (int)c->integrate_targets.size(), dt, false);
    AM_HIP(c, hipGetLastError());
    // RigidCollection.advanceVelocities (RigidCollection.java:873-877)
    if (has_collections(c)) return merge_push_down(c, true);
    return AM_OK;
}

int am_advance_positions(am_ctx* c, double dt) {
    int r = check_ctx(c);
    if (r) return r;
    launch_advance(c->stream, bodies_view(c), c->d_integrate.p, (int)c->integrate_targets.size(), dt, true);
    AM_HIP(c, hipGetLastError());
    // RigidCollection.advancePositions (:879-883)
    if (has_collections(c)) return merge_member_transforms(c);
    return AM_OK;
}

int am_advance_positions_post(am_ctx* c, double dt) {
    int r = check_ctx(c);
    if (r) return r;
    launch_advance_post(c->stream, bodies_view(c), c->d_integrate.p, (int)c->integrate_targets.size(), dt);
    AM_HIP(c, hipGetLastError());
    // RigidCollection.advancePositionsPostStabilization (:890-893)
    if (has_collections(c)) return merge_member_transforms(c);
    return AM_OK;
}

int am_step(am_ctx* c, double dt, am_step_stats* st) {
    int r = check_ctx(c);
    if (r) return r;
    if (c->params.enable_post_stabilization &&
        (c->params.enable_merging || c->params.enable_unmerging || has_collections(c)))
        return fail(c, AM_ERR_INVALID, "am_step runs post-stabilization with merging and unmerging off only");
    for (auto& e : c->sev)
        if (!e) AM_HIP(c, hipEventCreate(&e));
    hipEvent_t* ev = c->sev;
    hipStream_t s = c->stream;
    const am_params& P = c->params;
    const bool merging = P.enable_merging || P.enable_unmerging || has_collections(c);
    c->total_steps++;
    AM_HIP(c, hipEventRecord(ev[0], s));
    // RigidBodySystem.clearBodies (:190-202); prescribed motions where the
    // animate tool acts (:109)
    if ((r = am_clear_bodies(c))) return r;
    if ((r = kinematic_update(c, c->sim_time))) return r;
    if ((r = apply_external_forces(c))) return r;
    if ((r = am_update_contacts_map(c))) return r;
    if ((r = do_detect(c))) return r;
    AM_HIP(c, hipEventRecord(ev[1], s));
    int bad = 0, rep = 0;
    if ((r = am_warm_start(c, 0, &bad, &rep))) return r;
    // Sleeping.wake (:128)
    if ((r = sleep_wake(c, nullptr))) return r;
    AM_HIP(c, hipEventRecord(ev[5], s));
    int sweep_it = 0, cut = 0, merged = 0;
    if (merging) {
        if ((r = merge_update_in_collections(c, dt, &sweep_it))) return r;
        if ((r = merge_accumulate_unmerging(c, dt))) return r;
        if ((r = merge_unmerge(c, dt, &cut))) return r;
        if (c->merging_event) {
            launch_clear_targets(s, bodies_view(c), c->d_slot_entity.p, (int)c->slot_entity.size());
            if ((r = apply_external_forces(c))) return r;
        }
    }
    AM_HIP(c, hipEventRecord(ev[6], s));
    if (P.warm_start) r = am_redo_warm_start(c);
    else r = am_no_warm_start(c);
    if (r) return r;
    AM_HIP(c, hipEventRecord(ev[2], s));
    int iters = 0;
    double last = 0;
    if ((r = do_solve(c, dt, 0, true, &iters, &last))) return r;
    AM_HIP(c, hipEventRecord(ev[3], s));
    if (P.enable_merging && (r = merge_body_pairs(c))) return r;
    AM_HIP(c, hipEventRecord(ev[7], s));
    if ((r = am_advance_velocities(c, dt))) return r;
    if (P.enable_merging && (r = merge_accumulate_merging(c, dt))) return r;
    if ((r = am_advance_positions(c, dt))) return r;
    // RigidBodySystem.postStabilization (:354-378): a second detection and a
    // velocity-free solve whose deltaV moves the positions
    if (P.enable_post_stabilization) {
        if ((r = am_clear_bodies(c))) return r;
        if ((r = am_update_contacts_map(c))) return r;
        if ((r = do_detect(c))) return r;
        int pb = 0, pr = 0;
        if ((r = am_warm_start(c, 1, &pb, &pr))) return r;
        int pit = 0;
        double plast = 0;
        if ((r = do_solve(c, dt, 1, true, &pit, &plast))) return r;
        if ((r = am_advance_positions_post(c, dt))) return r;
    }
    AM_HIP(c, hipEventRecord(ev[8], s));
    if (P.enable_merging && P.steps_between_merge_events > 0 && c->total_steps % P.steps_between_merge_events == 0)
        if ((r = merge_merge(c, &merged))) return r;
    // Sleeping.sleep (:170)
    int asleep = 0;
    if ((r = sleep_sleep(c, &asleep))) return r;
    if ((r = apply_viscous_decay(c))) return r;
    c->sim_time += dt;
    AM_HIP(c, hipEventRecord(ev[4], s));
    AM_HIP(c, hipEventSynchronize(ev[4]));
    c->last_merged = merged;
    c->last_unmerged = cut;
    if (st) {
        std::memset(st, 0, sizeof *st);
        st->n_contacts = c->n_contacts;
        st->n_pairs = c->n_pairs;
        st->n_groups = c->n_groups;
        st->n_colours = c->n_colours;
        st->iterations_used = iters;
        st->bad_warm_starts = bad;
        st->bad_warm_starts_repaired = rep;
        st->last_change = last;
        auto el = [&](int a, int b) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
            return (double)ms;
        };
        st->ms_detect = el(0, 1);
        st->ms_warm_start = el(1, 5) + el(6, 2);
        st->ms_solve = el(2, 3);
        st->ms_integrate = el(7, 8);
        st->ms_merging = el(5, 6) + el(3, 7) + el(8, 4);
        st->ms_total = el(0, 4);
        st->ms_pgs_kernel = c->n_contacts ? c->ms_pgs_kernel : 0.0;
        st->ms_setup_kernel = c->n_contacts ? c->ms_setup_kernel : 0.0;
        st->ms_learned = c->n_contacts ? c->ms_learned : 0.0;
        st->start_residual_plain = c->res_plain;
        st->start_residual_learned = c->res_learned;
        st->used_learned = c->used_learned;
        int nc = 0;
        for (int e : c->slot_entity)
            if (is_collection(c, e)) nc++;
        st->n_collections = nc;
        st->n_merged = merged;
        st->n_unmerged = cut;
        st->n_internal_contacts = c->n_ic;
        st->n_sleeping = asleep;
    }
    return AM_OK;
}

}  // extern "C"

namespace {
// Staging for am_motion_metric_noalloc: the pairs (2 ints each) then the
// results (1 double each) of up to AM_METRIC_NOALLOC_MAX pairs, 1 MiB in all.
// Every GPU the library runs on holds its own copy until the process ends.
constexpr size_t kNoAllocBytes = (size_t)AM_METRIC_NOALLOC_MAX * 16;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array, looked up once per
// device with that device current (the array's host-side name is not a
// device address).
int scratch_on_device(am_ctx* c, unsigned char** out) {
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, AM_ERR_INVALID, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    unsigned char*& p = g_scratch_ptr[c->device];
    if (!p) {
        void* addr = nullptr;
        hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !addr)
            return fail(c, AM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        p = (unsigned char*)addr;
    }
    *out = p;
    return AM_OK;
}
}  // namespace

extern "C" {

int am_motion_metric_noalloc(am_ctx* c, const int32_t* pairs, int32_t n_pairs, double* out) {
    int r = check_ctx(c);
    if (r) return r;
    if (n_pairs < 0 || (n_pairs > 0 && (!pairs || !out))) return fail(c, AM_ERR_INVALID, "bad metric arrays");
    if (n_pairs > AM_METRIC_NOALLOC_MAX)
        return fail(c, AM_ERR_CAPACITY, "am_motion_metric_noalloc takes at most AM_METRIC_NOALLOC_MAX (65536) pairs");
    unsigned char* base = nullptr;
    if ((r = scratch_on_device(c, &base))) return r;
    if (n_pairs == 0) return AM_OK;
    for (int k = 0; k < n_pairs; k++)
        if (pairs[2 * k] < 0 || pairs[2 * k] >= c->n_ent || pairs[2 * k + 1] >= c->n_ent)
            return fail(c, AM_ERR_INVALID, "metric entity out of range");
    int* dp = reinterpret_cast<int*>(base);
    double* dout = reinterpret_cast<double*>(base + (size_t)AM_METRIC_NOALLOC_MAX * 8);
    AM_HIP(c, hipMemcpyAsync(dp, pairs, (size_t)n_pairs * 8, hipMemcpyHostToDevice, c->stream));
    launch_metric(c->stream, bodies_view(c), dp, n_pairs, dout);
    AM_HIP(c, hipGetLastError());
    AM_HIP(c, hipMemcpyAsync(out, dout, (size_t)n_pairs * 8, hipMemcpyDeviceToHost, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_motion_metric(am_ctx* c, const int32_t* pairs, int32_t n_pairs, double* out) {
    int r = check_ctx(c);
    if (r) return r;
    if (n_pairs < 0 || (n_pairs > 0 && (!pairs || !out))) return fail(c, AM_ERR_INVALID, "bad metric arrays");
    if (n_pairs == 0) return AM_OK;
    for (int k = 0; k < n_pairs; k++)
        if (pairs[2 * k] < 0 || pairs[2 * k] >= c->n_ent || pairs[2 * k + 1] >= c->n_ent)
            return fail(c, AM_ERR_INVALID, "metric entity out of range");
    DBuf<int> dp;
    DBuf<double> dout;
    AM_HIP(c, dp.ensure((size_t)n_pairs * 2));
    AM_HIP(c, dout.ensure(n_pairs));
    AM_HIP(c, hipMemcpyAsync(dp.p, pairs, (size_t)n_pairs * 8, hipMemcpyHostToDevice, c->stream));
    launch_metric(c->stream, bodies_view(c), dp.p, n_pairs, dout.p);
    AM_HIP(c, hipGetLastError());
    AM_HIP(c, hipMemcpyAsync(out, dout.p, (size_t)n_pairs * 8, hipMemcpyDeviceToHost, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}


int am_mlp_init(am_ctx* c, uint64_t seed) {
    int r = check_ctx(c);
    if (r) return r;
    // He-uniform initialisation, biases 0; a small output layer so the first
    // predictions stay near zero (a splitmix64 stream seeded by `seed`)
    std::vector<float> flat(kMlpParams, 0.f);
    uint64_t st = seed ^ 0x9e3779b97f4a7c15ull;
    auto rnd = [&st]() {
        st += 0x9e3779b97f4a7c15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    size_t q = 0;
    const int dims[4][2] = {{24, 64}, {64, 64}, {64, 64}, {64, 3}};
    for (int layer = 0; layer < 4; layer++) {
        int fan_in = dims[layer][0], fan_out = dims[layer][1];
        double lim = std::sqrt(6.0 / fan_in) * (layer == 3 ? 0.1 : 1.0);
        for (int i = 0; i < fan_in * fan_out; i++) flat[q++] = (float)((2.0 * rnd() - 1.0) * lim);
        q += fan_out;  // biases stay 0
    }
    return mlp_upload(c, flat.data());
}

int am_mlp_set_params(am_ctx* c, const float* params, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (!params || n != kMlpParams) return fail(c, AM_ERR_INVALID, "expected AM_MLP_PARAMS floats");
    return mlp_upload(c, params);
}

int am_mlp_get_params(am_ctx* c, float* params, int32_t n) {
    int r = check_ctx(c);
    if (r) return r;
    if (!params || n != kMlpParams) return fail(c, AM_ERR_INVALID, "expected AM_MLP_PARAMS floats");
    if (!c->mlp_ready) return fail(c, AM_ERR_STATE, "MLP not initialised");
    std::vector<float> pad(kPadTotal);
    AM_HIP(c, hipMemcpyAsync(pad.data(), c->m_p.p, kPadTotal * 4, hipMemcpyDeviceToHost, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    unpack_params(pad, params);
    return AM_OK;
}

int am_mlp_forward(am_ctx* c, const float* x, int32_t n, float* y) {
    int r = check_ctx(c);
    if (r) return r;
    if (!c->mlp_ready) return fail(c, AM_ERR_STATE, "MLP not initialised");
    if (n < 0 || (n > 0 && (!x || !y))) return fail(c, AM_ERR_INVALID, "bad MLP arrays");
    if (n == 0) return AM_OK;
    DBuf<float> dx, dy;
    AM_HIP(c, dx.ensure((size_t)n * kMlpIn));
    AM_HIP(c, dy.ensure((size_t)n * kMlpOut));
    AM_HIP(c, hipMemcpyAsync(dx.p, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, c->stream));
    launch_mlp_forward(c->stream, mlp_weights(c), dx.p, n, dy.p);
    AM_HIP(c, hipGetLastError());
    AM_HIP(c, hipMemcpyAsync(y, dy.p, (size_t)n * kMlpOut * 4, hipMemcpyDeviceToHost, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_mlp_gradients(am_ctx* c, const float* x, const float* t, int32_t n, float* grads, float* loss) {
    int r = check_ctx(c);
    if (r) return r;
    if (!c->mlp_ready) return fail(c, AM_ERR_STATE, "MLP not initialised");
    if (n <= 0 || !x || !t || !grads) return fail(c, AM_ERR_INVALID, "bad MLP arrays");
    DBuf<float> dx, dt_;
    DBuf<int> di;
    AM_HIP(c, dx.ensure((size_t)n * kMlpIn));
    AM_HIP(c, dt_.ensure((size_t)n * kMlpOut));
    AM_HIP(c, di.ensure(n));
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    hipStream_t s = c->stream;
    AM_HIP(c, hipMemcpyAsync(dx.p, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, s));
    AM_HIP(c, hipMemcpyAsync(dt_.p, t, (size_t)n * kMlpOut * 4, hipMemcpyHostToDevice, s));
    AM_HIP(c, hipMemcpyAsync(di.p, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    fill_int(s, (int*)c->m_g.p, kPadTotal, 0);
    fill_int(s, (int*)c->d_loss.p, 1, 0);
    launch_mlp_backward(s, mlp_weights(c), dx.p, dt_.p, di.p, n, 1.0f / (float)n, mlp_grads(c), c->d_loss.p);
    AM_HIP(c, hipGetLastError());
    std::vector<float> pad(kPadTotal);
    float l = 0;
    AM_HIP(c, hipMemcpyAsync(pad.data(), c->m_g.p, kPadTotal * 4, hipMemcpyDeviceToHost, s));
    AM_HIP(c, hipMemcpyAsync(&l, c->d_loss.p, 4, hipMemcpyDeviceToHost, s));
    AM_HIP(c, hipStreamSynchronize(s));
    fill_int(s, (int*)c->m_g.p, kPadTotal, 0);
    unpack_params(pad, grads);
    if (loss) *loss = l;
    return AM_OK;
}

int am_features_pull(am_ctx* c, float* x, double* scale, int32_t capacity, int32_t* n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n) *n = c->n_contacts;
    if (c->n_contacts == 0 || (!x && !scale)) return AM_OK;
    r = compute_features(c, 0.01);
    if (r) return r;
    size_t m = (size_t)std::min(capacity, c->n_contacts);
    if (x) AM_HIP(c, hipMemcpyAsync(x, c->f_X.p, m * kMlpIn * 4, hipMemcpyDeviceToHost, c->stream));
    if (scale) AM_HIP(c, hipMemcpyAsync(scale, c->f_scale.p, m * 8, hipMemcpyDeviceToHost, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

int am_train_size(am_ctx* c, int32_t* n) {
    if (!c || !n) return AM_ERR_INVALID;
    *n = c->t_n;
    return AM_OK;
}

int am_train_clear(am_ctx* c) {
    if (!c) return AM_ERR_INVALID;
    c->t_n = 0;
    return AM_OK;
}

int am_train_epochs(am_ctx* c, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* last_loss) {
    int r = check_ctx(c);
    if (r) return r;
    if (!c->mlp_ready) return fail(c, AM_ERR_STATE, "MLP not initialised");
    if (epochs < 0 || batch <= 0 || c->t_n <= 0) return fail(c, AM_ERR_INVALID, "no training data or bad batch");
    const int n = c->t_n;
    hipStream_t s = c->stream;
    AM_HIP(c, c->t_idx.ensure(n));
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    uint64_t st = seed;
    float loss = 0;
    for (int e = 0; e < epochs; e++) {
        for (int i = n - 1; i > 0; i--) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            int j = (int)((st >> 33) % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        AM_HIP(c, hipMemcpyAsync(c->t_idx.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
        fill_int(s, (int*)c->d_loss.p, 1, 0);
        for (int b0 = 0; b0 < n; b0 += batch) {
            int m = std::min(batch, n - b0);
            launch_mlp_backward(s, mlp_weights(c), c->t_X.p, c->t_T.p, c->t_idx.p + b0, m, 1.0f / (float)m,
                                mlp_grads(c), c->d_loss.p);
            c->adam_step++;
            launch_adam(s, c->m_p.p, c->m_g.p, c->m_m.p, c->m_v.p, c->m_bf.p, kBfTotal, lr, 0.9f, 0.999f, 1e-8f,
                        c->adam_step, nullptr);
            launch_adam(s, c->m_p.p + kBfTotal, c->m_g.p + kBfTotal, c->m_m.p + kBfTotal, c->m_v.p + kBfTotal,
                        nullptr, kPadTotal - kBfTotal, lr, 0.9f, 0.999f, 1e-8f, c->adam_step, nullptr);
        }
        AM_HIP(c, hipGetLastError());
        AM_HIP(c, hipMemcpyAsync(&loss, c->d_loss.p, 4, hipMemcpyDeviceToHost, s));
        AM_HIP(c, hipStreamSynchronize(s));
        int nb = (n + batch - 1) / batch;
        loss /= (float)nb;
    }
    if (last_loss) *last_loss = loss;
    return AM_OK;
}

int am_start_lambdas_pull(am_ctx* c, double* lambda, int32_t capacity, int32_t* n) {
    int r = check_ctx(c);
    if (r) return r;
    if (n) *n = c->n_contacts;
    if (!lambda || c->n_contacts == 0 || !c->s_start.p) return AM_OK;
    size_t m = (size_t)std::min(capacity, c->n_contacts);
    AM_HIP(c, hipMemcpyAsync(lambda, c->s_start.p, m * 24, hipMemcpyDeviceToHost, c->stream));
    AM_HIP(c, hipStreamSynchronize(c->stream));
    return AM_OK;
}

}  // extern "C"
