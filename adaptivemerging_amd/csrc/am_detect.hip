// am_detect.hip - broad phase (sort-based uniform hash grid) and the two-pass
// narrow phase with the contact set-up (Contact.set) fused into the emit pass.
//
// Contact order is the reference's: the broad phase walks top-level entities
// i<j in list order and a collection's members in member order
// (CollisionProcessor.java:683-697, 768-791), so pairs are keyed
// (slot_lo, slot_hi, sub_lo, sub_hi), sorted, and each pair's contacts are
// written at the exclusive-scan offset of the per-pair counts.
#include "am_geom.hpp"

namespace am {

// ---------------------------------------------------------------------------
// BVSphere.updatecW for every node of every tree body (BVSphere.java:174-176).
// BVSphere.updatecW (BVSphere.java:174-176) with the owning body's transform;
// a node of a composite's sphere sub-body uses the sub-body's transform
// (RigidBodyGeomComposite.updateBodyPositionsFromParent).
__global__ void k_tree_world_centres(BodiesView B, TreeView T, const int* node_body, const int* node_sub,
                                     int n_nodes) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_nodes) return;
    int b = node_body[g];
    int s = node_sub[g];
    BodyPose P = s < 0 ? pose(B, b) : sub_pose(B, b, s);
    d3 c = xform(P.R, P.x, ld3(T.nodes[g].c));
    st3(&T.cw[g * 3], c);
}

void launch_tree_world_centres(hipStream_t s, BodiesView B, TreeView T, const int* node_body, const int* node_sub,
                               int n_nodes) {
    if (n_nodes <= 0) return;
    hipLaunchKernelGGL(k_tree_world_centres, dim3((n_nodes + 255) / 256), dim3(256), 0, s, B, T, node_body, node_sub,
                       n_nodes);
}

// ---------------------------------------------------------------------------
// Bounding sphere and grid cell of each gridded body.
AMD_INLINE int cell_coord(double v, double inv_h) {
    double c = floor(v * inv_h);
    c = fmin(fmax(c, -1.0e9), 1.0e9);
    return (int)c;
}
AMD_INLINE int cell_hash(int cx, int cy, int cz, int scene, int mask) {
    uint32_t h = (uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u ^ (uint32_t)cz * 83492791u ^ (uint32_t)scene * 2654435761u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    return (int)(h & (uint32_t)mask);
}

// sphere: 4 doubles per gridded body (centre, radius); cell_key: 4 ints (cx, cy, cz, bucket);
// a body too large for the grid gets bucket -1 and stays out of the buckets.
__global__ void k_bounds(BodiesView B, TreeView T, const int* order, int n, double* sphere, double inv_h,
                         int* cell, int mask, const int* is_large) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int b = order[k];
    const am_body_desc& d = B.desc[b];
    d3 c;
    double r;
    if (d.geom == AM_GEOM_BOX || d.geom == AM_GEOM_COMPOSITE) {
        // a composite's radius encloses all of its sub-bodies about its centre of mass
        c = ld3(&B.x[b * 3]);
        r = d.radius;
    } else {
        c = ld3(&T.cw[d.tree_first * 3]);
        r = T.nodes[d.tree_first].r;
    }
    sphere[k * 4 + 0] = c.x;
    sphere[k * 4 + 1] = c.y;
    sphere[k * 4 + 2] = c.z;
    sphere[k * 4 + 3] = r;
    int cx = cell_coord(c.x, inv_h), cy = cell_coord(c.y, inv_h), cz = cell_coord(c.z, inv_h);
    cell[k * 4 + 0] = cx;
    cell[k * 4 + 1] = cy;
    cell[k * 4 + 2] = cz;
    cell[k * 4 + 3] = is_large[k] ? -1 : cell_hash(cx, cy, cz, d.scene, mask);
}

void launch_bounds(hipStream_t s, BodiesView B, TreeView T, const int* order, int n, double* sphere, double cell,
                   int* cell_key, int hash_mask, const int* is_large) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bounds, dim3((n + 255) / 256), dim3(256), 0, s, B, T, order, n, sphere, 1.0 / cell, cell_key,
                       hash_mask, is_large);
}

__global__ void k_bucket_count(const int* cell, int n, int* count) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int h = cell[k * 4 + 3];
    if (h >= 0) atomicAdd(&count[h], 1);
}
__global__ void k_bucket_scatter(const int* cell, int n, const int* start, int* cursor, int* bucket_items) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int h = cell[k * 4 + 3];
    if (h < 0) return;
    int pos = start[h] + atomicAdd(&cursor[h], 1);
    bucket_items[pos] = k;
}

void launch_buckets(hipStream_t s, const int* cell, int n, int* count, const int* start, int* cursor, int* items,
                    bool scatter) {
    if (n <= 0) return;
    if (!scatter) hipLaunchKernelGGL(k_bucket_count, dim3((n + 255) / 256), dim3(256), 0, s, cell, n, count);
    else hipLaunchKernelGGL(k_bucket_scatter, dim3((n + 255) / 256), dim3(256), 0, s, cell, n, start, cursor, items);
}

AMD_INLINE uint64_t pair_key(int slot_a, int sub_a, int slot_b, int sub_b) {
    if (slot_a > slot_b) {
        int t = slot_a; slot_a = slot_b; slot_b = t;
        t = sub_a; sub_a = sub_b; sub_b = t;
    }
    return ((uint64_t)slot_a << 44) | ((uint64_t)slot_b << 24) | ((uint64_t)sub_a << 12) | (uint64_t)sub_b;
}

// Broad-phase skip rules on top-level entities (CollisionProcessor.java:692-693).
AMD_INLINE bool skip_pair(const PairArgs& A, int slot_a, int slot_b) {
    bool pa = A.top_pinned[slot_a], pb = A.top_pinned[slot_b];
    if (pa && pb) return true;
    if ((pa && A.top_sleeping[slot_b]) || (pb && A.top_sleeping[slot_a])) return true;
    return false;
}

// One thread per (gridded body, task): count (write=false) or write pairs.
// Tasks 0-26 are the 27 cells around the body's own: the pairs with grid
// bodies of a higher slot found there.  Task 27 holds the pairs with every
// large body (those stay out of the grid, whose cells are sized for the
// others) and all the plane pairs; a large body has only task 27, where it
// pairs with large bodies of a higher slot and with the planes.  Spreading a
// body's cells over threads keeps each thread's chain of dependent loads
// short.  The keys are sorted afterwards, so the emission order is free.
constexpr int kPairTasks = 28;
template <bool kWrite>
__global__ void k_pairs(PairArgs A, int* counts, const int* offsets, uint64_t* keys, int cap, int* overflow) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)A.n_grid * kPairTasks) return;
    const int k = (int)(t / kPairTasks), task = (int)(t % kPairTasks);
    const BodiesView& B = A.B;
    const int bi = A.grid_order[k];
    const int slot_i = B.slot[bi], sub_i = B.sub[bi];
    const am_body_desc& di = B.desc[bi];
    const double ci[4] = {A.sphere[k * 4 + 0], A.sphere[k * 4 + 1], A.sphere[k * 4 + 2], A.sphere[k * 4 + 3]};
    const bool large_i = A.cell[k * 4 + 3] < 0;
    int n = 0;
    const int base = kWrite ? offsets[t] : 0;
    auto emit = [&](int slot_j, int sub_j) {
        if (kWrite) {
            const int pos = base + n;
            if (pos < cap) keys[pos] = pair_key(slot_i, sub_i, slot_j, sub_j);
            else atomicOr(overflow, 1);
        }
        n++;
    };
    if (task < 27) {
        // a large body scans no cells: every grid body pairs with it in its task 27
        if (!large_i) {
            const int nx = A.cell[k * 4 + 0] + task % 3 - 1;
            const int ny = A.cell[k * 4 + 1] + (task / 3) % 3 - 1;
            const int nz = A.cell[k * 4 + 2] + task / 9 - 1;
            const int h = cell_hash(nx, ny, nz, di.scene, A.mask);
            const int s0 = A.bucket_start[h], cnt = A.bucket_count[h];
            for (int q = 0; q < cnt; q++) {
                const int kj = A.bucket_items[s0 + q];
                if (A.cell[kj * 4 + 0] != nx || A.cell[kj * 4 + 1] != ny || A.cell[kj * 4 + 2] != nz) continue;
                const int bj = A.grid_order[kj];
                if (B.desc[bj].scene != di.scene) continue;
                const int slot_j = B.slot[bj];
                if (slot_j <= slot_i) continue;
                if (skip_pair(A, slot_i, slot_j)) continue;
                const double ddx = A.sphere[kj * 4 + 0] - ci[0];
                const double ddy = A.sphere[kj * 4 + 1] - ci[1];
                const double ddz = A.sphere[kj * 4 + 2] - ci[2];
                const double rr = A.sphere[kj * 4 + 3] + ci[3];
                const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                // superset of every narrow-phase first test (<= with slack)
                if (d2 > rr * rr * (1.0 + 1e-9) + 1e-300) continue;
                emit(slot_j, B.sub[bj]);
            }
        }
    } else {
        for (int q = 0; q < A.n_large; q++) {
            const int kj = A.large[q];
            if (kj == k) continue;
            const int bj = A.grid_order[kj];
            if (B.desc[bj].scene != di.scene) continue;
            const int slot_j = B.slot[bj];
            if (slot_j == slot_i || (large_i && slot_j < slot_i)) continue;
            if (skip_pair(A, slot_i, slot_j)) continue;
            const double ddx = A.sphere[kj * 4 + 0] - ci[0];
            const double ddy = A.sphere[kj * 4 + 1] - ci[1];
            const double ddz = A.sphere[kj * 4 + 2] - ci[2];
            const double rr = A.sphere[kj * 4 + 3] + ci[3];
            const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (d2 > rr * rr * (1.0 + 1e-9) + 1e-300) continue;
            emit(slot_j, B.sub[bj]);
        }
        for (int q = A.plane_start[di.scene]; q < A.plane_start[di.scene + 1]; q++) {
            const int p = A.planes[q];
            const am_body_desc& pd = B.desc[p];
            const int slot_p = B.slot[p];
            if (slot_p == slot_i) continue;
            if (skip_pair(A, slot_i, slot_p)) continue;
            const d3 nrm = ld3(pd.plane_n);
            if (di.geom == AM_GEOM_BOX || di.geom == AM_GEOM_COMPOSITE) {
                // first test of BoxPlane.dBoxPlane (for a composite: its enclosing sphere)
                if (dot(ld3(&B.x[bi * 3]), nrm) + pd.plane_d > di.radius) continue;
            } else {
                // first (root) test of collideSphereTreeAndPlane
                const d3 c = ld3(&A.T.cw[di.tree_first * 3]);
                const double d = nrm.x * c.x + nrm.y * c.y + nrm.z * c.z + pd.plane_d - A.T.nodes[di.tree_first].r;
                if (!(d < 0)) continue;
            }
            emit(slot_p, B.sub[p]);
        }
    }
    if (!kWrite) counts[t] = n;
}

int pair_tasks(int n_grid) { return n_grid * kPairTasks; }

void launch_pairs_impl(hipStream_t s, bool write, const PairArgs& A, int* counts, const int* offsets, uint64_t* keys,
                       int cap, int* overflow) {
    if (A.n_grid <= 0) return;
    const long n = (long)A.n_grid * kPairTasks;
    dim3 g((unsigned)((n + 255) / 256)), b(256);
    if (write) hipLaunchKernelGGL(k_pairs<true>, g, b, 0, s, A, counts, offsets, keys, cap, overflow);
    else hipLaunchKernelGGL(k_pairs<false>, g, b, 0, s, A, counts, offsets, keys, cap, overflow);
}

// ---------------------------------------------------------------------------
// Narrow phase.
struct CountSink {
    int n = 0;
    AMD_INLINE void emit(int, int, d3, d3, int, int, int, double) { n++; }
};

// Contact.set (Contact.java:159-233): tangent frame, Jacobian against the
// solve bodies (collection parent when merged), frame stored in body-1
// coordinates, multipliers cleared, newThisTimeStep set.
struct WriteSink {
    const BodiesView* B;
    ContactsView C;
    int base;
    int n = 0;
    int* overflow;
    AMD_INLINE void emit(int b1, int b2, d3 p, d3 nW, int bv1, int bv2, int info, double viol) {
        int idx = base + n;
        n++;
        if (idx >= C.cap) {
            atomicOr(overflow, 1);
            return;
        }
        C.b1[idx] = b1;
        C.b2[idx] = b2;
        C.bv1[idx] = bv1;
        C.bv2[idx] = bv2;
        C.info[idx] = info;
        C.viol[idx] = viol;
        C.prev_viol[idx] = 0.0;
        C.state[idx] = AM_STATE_CLEAR;
        C.is_new[idx] = 1;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            C.lam[idx * 3 + r] = 0.0;
            C.lamw[idx * 3 + r] = 0.0;
            C.b[idx * 3 + r] = 0.0;
            C.D[idx * 3 + r] = 0.0;
        }
        double anx = fabs(nW.x), any = fabs(nW.y), anz = fabs(nW.z);
        d3 t1;
        if (anx < any && anx < anz) t1 = mk(1, 0, 0);
        else if (any < anz) t1 = mk(0, 1, 0);
        else t1 = mk(0, 0, 1);
        d3 t2 = normalized(cross(nW, t1));
        t1 = cross(t2, nW);
        int a = B->parent[b1] >= 0 ? B->parent[b1] : b1;
        int b = B->parent[b2] >= 0 ? B->parent[b2] : b2;
        d3 r1 = sub(p, ld3(&B->x[a * 3]));
        d3 r2 = sub(p, ld3(&B->x[b * 3]));
        double* J = &C.J[(size_t)idx * 36];
        const d3 dirs[3] = {nW, t1, t2};
#pragma unroll
        for (int row = 0; row < 3; row++) {
            d3 d = dirs[row];
            st3(J + row * 12 + 0, mk(-1.0 * d.x, -1.0 * d.y, -1.0 * d.z));
            st3(J + row * 12 + 3, cross(d, r1));
            st3(J + row * 12 + 6, d);
            st3(J + row * 12 + 9, cross(r2, d));
        }
        m33 R = ldm(&B->R[b1 * 9]);
        d3 x = ld3(&B->x[b1 * 3]);
        st3(&C.pB1[idx * 3], inv_xform(R, x, p));
        st3(&C.nB1[idx * 3], mul_t(R, nW));
        st3(&C.t1B1[idx * 3], mul_t(R, t1));
        st3(&C.t2B1[idx * 3], mul_t(R, t2));
    }
};

struct NarrowArgs {
    BodiesView B;
    TreeView T;
    const uint64_t* keys;
    const int* n_pairs;
    const int* slot_entity;
    const int* member_first;
    const int* members;
};

AMD_INLINE int body_of(const NarrowArgs& A, int slot, int sub) {
    int e = A.slot_entity[slot];
    if (A.B.desc[e].geom == AM_GEOM_COLLECTION) return A.members[A.member_first[e] + sub];
    return e;
}

template <bool kWrite>
__global__ __launch_bounds__(128) void k_narrow(NarrowArgs A, int* counts, const int* offsets, ContactsView C,
                                                int* overflow) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *A.n_pairs) return;
    uint64_t key = A.keys[k];
    int slot_lo = (int)(key >> 44) & (kMaxSlots - 1);
    int slot_hi = (int)(key >> 24) & (kMaxSlots - 1);
    int sub_lo = (int)(key >> 12) & (kMaxSub - 1);
    int sub_hi = (int)key & (kMaxSub - 1);
    int b1 = body_of(A, slot_lo, sub_lo);
    int b2 = body_of(A, slot_hi, sub_hi);
    if (kWrite) {
        WriteSink s;
        s.B = &A.B;
        s.C = C;
        s.base = offsets[k];
        s.overflow = overflow;
        narrow_dispatch(A.B, A.T, b1, b2, s, overflow);
    } else {
        CountSink s;
        narrow_dispatch(A.B, A.T, b1, b2, s, overflow);
        counts[k] = s.n;
    }
}

void launch_narrow(hipStream_t s, bool write, BodiesView B, TreeView T, const uint64_t* keys, const int* n_pairs,
                   int pair_cap, const int* slot_entity, const int* member_first, const int* members, int* counts,
                   const int* offsets, ContactsView C, int* overflow) {
    if (pair_cap <= 0) return;
    NarrowArgs A{B, T, keys, n_pairs, slot_entity, member_first, members};
    dim3 g((pair_cap + 127) / 128), b(128);
    if (write) hipLaunchKernelGGL(k_narrow<true>, g, b, 0, s, A, counts, offsets, C, overflow);
    else hipLaunchKernelGGL(k_narrow<false>, g, b, 0, s, A, counts, offsets, C, overflow);
}

// ---------------------------------------------------------------------------
// Contacts supplied by the caller (am_contacts_push): Contact.set on each.
__global__ void k_contacts_set(BodiesView B, int n, const int* b1, const int* b2, const double* pw, const double* nw,
                               const double* viol, const int* info, const int* bv1, const int* bv2, ContactsView C,
                               int* overflow) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    WriteSink s;
    s.B = &B;
    s.C = C;
    s.base = k;
    s.overflow = overflow;
    s.emit(b1[k], b2[k], ld3(&pw[k * 3]), ld3(&nw[k * 3]), bv1[k], bv2[k], info[k], viol[k]);
}

void launch_contacts_set(hipStream_t s, BodiesView B, int n, const int* b1, const int* b2, const double* pw,
                         const double* nw, const double* viol, const int* info, const int* bv1, const int* bv2,
                         ContactsView C, int* overflow) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_contacts_set, dim3((n + 127) / 128), dim3(128), 0, s, B, n, b1, b2, pw, nw, viol, info, bv1,
                       bv2, C, overflow);
}

}  // namespace am
