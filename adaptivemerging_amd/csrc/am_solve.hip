// am_solve.hip - solve groups, greedy Jones-Plassmann colouring of the
// group-body graph, the persistent coloured projected Gauss-Seidel kernel,
// the warm start, integration, gravity and the relative-motion metric.
//
// Coloured Gauss-Seidel: contacts are grouped by the pair of solve bodies they
// act on (a "group"; contacts of one body pair are contiguous in the contact
// list).  Groups that share a dynamic body get different colours.  A sweep
// runs the colours in order; inside a colour every group is independent and
// one lane processes the contacts of its group sequentially.  The result is
// bit-identical to a sequential PGS sweep (PGS.java:100-193) over the contacts
// in (colour, group, contact) order, because each contact reads and writes
// exactly the deltaV values that order would give it.
#include <hip/hip_cooperative_groups.h>

#include "am_solve.hpp"

namespace cg = cooperative_groups;

namespace am {

// ---------------------------------------------------------------------------
// small utility kernels
__global__ void k_fill_int(int* p, int n, int v) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}
__global__ void k_fill_double(double* p, int n, double v) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}
void fill_int(hipStream_t s, int* p, int n, int v) {
    if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}
void fill_double(hipStream_t s, double* p, int n, double v) {
    if (n > 0) hipLaunchKernelGGL(k_fill_double, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}

// ---------------------------------------------------------------------------
// Groups: runs of contacts with the same unordered slot pair.
__global__ void k_group_flags(int cap, const int* n_contacts, const int* cb1, const int* cb2, const int* slot,
                              int* flags) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int n = *n_contacts;
    if (k >= cap) return;
    if (k >= n) {
        flags[k] = 0;
        return;
    }
    int f = 1;
    if (k > 0) {
        int a = slot[cb1[k]], b = slot[cb2[k]];
        int pa = slot[cb1[k - 1]], pb = slot[cb2[k - 1]];
        int lo = min(a, b), hi = max(a, b), plo = min(pa, pb), phi = max(pa, pb);
        f = (lo != plo || hi != phi) ? 1 : 0;
    }
    flags[k] = f;
}

__global__ void k_group_build(const int* n_contacts, const int* flags, const int* scanned, const int* cb1,
                              const int* cb2, const int* parent, int* group_start, int* group_a, int* group_b) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int n = *n_contacts;
    if (k >= n) return;
    if (flags[k]) {
        int g = scanned[k];
        group_start[g] = k;
        int a = cb1[k], b = cb2[k];
        group_a[g] = parent[a] >= 0 ? parent[a] : a;
        group_b[g] = parent[b] >= 0 ? parent[b] : b;
    }
    if (k == n - 1) group_start[scanned[k] + flags[k]] = n;
}

void launch_groups(hipStream_t s, int cap, const int* n_contacts, const int* cb1, const int* cb2, const int* slot,
                   const int* parent, int* flags, int* scanned, int* group_start, int* group_a, int* group_b,
                   int* d_n_groups, void* scan_temp) {
    if (cap <= 0) return;
    dim3 g((cap + 255) / 256), b(256);
    hipLaunchKernelGGL(k_group_flags, g, b, 0, s, cap, n_contacts, cb1, cb2, slot, flags);
    exclusive_scan(s, flags, scanned, cap, d_n_groups, scan_temp);
    hipLaunchKernelGGL(k_group_build, g, b, 0, s, n_contacts, flags, scanned, cb1, cb2, parent, group_start, group_a,
                       group_b);
}

// ---------------------------------------------------------------------------
// Colouring.
AMD_INLINE bool dynamic_body(const int* flags, int e) { return (flags[e] & AM_BODY_PINNED) == 0; }

__global__ void k_incidence_count(int n_groups, const int* ga, const int* gb, const int* bflags, int* inc_count) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    if (dynamic_body(bflags, ga[g])) atomicAdd(&inc_count[ga[g]], 1);
    if (dynamic_body(bflags, gb[g])) atomicAdd(&inc_count[gb[g]], 1);
}
__global__ void k_incidence_fill(int n_groups, const int* ga, const int* gb, const int* bflags, const int* inc_start,
                                 int* inc_cursor, int* inc_items) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    int a = ga[g], b = gb[g];
    if (dynamic_body(bflags, a)) inc_items[inc_start[a] + atomicAdd(&inc_cursor[a], 1)] = g;
    if (dynamic_body(bflags, b)) inc_items[inc_start[b] + atomicAdd(&inc_cursor[b], 1)] = g;
}

AMD_INLINE uint32_t prio(uint32_t g) {
    uint32_t h = g * 0x9E3779B1u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    return h;
}
AMD_INLINE bool higher(uint32_t h, uint32_t g) {
    uint32_t ph = prio(h), pg = prio(g);
    return ph > pg || (ph == pg && h > g);
}

// One Jones-Plassmann round: an uncoloured group whose priority is the highest
// among its uncoloured neighbours takes the smallest colour unused by its
// coloured neighbours (colours read from the previous round's array, so the
// result does not depend on scheduling).
__global__ void k_colour_round(int n_groups, const int* ga, const int* gb, const int* bflags, const int* inc_start,
                               const int* inc_count, const int* inc_items, const int* col_in, int* col_out,
                               int* remaining) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    int c = col_in[g];
    if (c >= 0) {
        col_out[g] = c;
        return;
    }
    uint64_t used = 0;
    bool ready = true;
    const int bodies[2] = {ga[g], gb[g]};
    for (int s = 0; s < 2 && ready; s++) {
        int e = bodies[s];
        if (!dynamic_body(bflags, e)) continue;
        int st = inc_start[e], cnt = inc_count[e];
        for (int q = 0; q < cnt; q++) {
            int h = inc_items[st + q];
            if (h == g) continue;
            int ch = col_in[h];
            if (ch < 0) {
                if (higher((uint32_t)h, (uint32_t)g)) {
                    ready = false;
                    break;
                }
            } else if (ch < kMaxColours) {
                used |= (uint64_t)1 << ch;
            }
        }
    }
    if (!ready) {
        col_out[g] = -1;
        atomicAdd(remaining, 1);
        return;
    }
    int pick = kMaxColours;  // sequential tail if all 64 parallel colours are taken
    if (~used != 0) pick = __ffsll((long long)~used) - 1;
    col_out[g] = pick;
}

__global__ void k_colour_keys(int n_groups, const int* colour, uint64_t* keys) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    keys[g] = ((uint64_t)(uint32_t)colour[g] << 32) | (uint32_t)g;
}

// colour_start[c] = first sorted position with colour >= c, c = 0..kMaxColours+1
__global__ void k_colour_starts(int n_groups, const uint64_t* sorted, int* colour_start) {
    int c = threadIdx.x;
    if (c > kMaxColours + 1) return;
    int lo = 0, hi = n_groups;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((int)(sorted[mid] >> 32) < c) lo = mid + 1;
        else hi = mid;
    }
    colour_start[c] = lo;
}

__global__ void k_sorted_sizes(int n_groups, const uint64_t* sorted, const int* group_start, int* sizes) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == n_groups) sizes[q] = 0;
    if (q >= n_groups) return;
    int g = (int)(uint32_t)sorted[q];
    sizes[q] = group_start[g + 1] - group_start[g];
}

__global__ void k_solve_order(int n_groups, const uint64_t* sorted, const int* group_start, const int* pos_start,
                              int* order) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_groups) return;
    int g = (int)(uint32_t)sorted[q];
    int s = group_start[g], e = group_start[g + 1], o = pos_start[q];
    for (int k = s; k < e; k++) order[o + (k - s)] = k;
}

// ---------------------------------------------------------------------------
// Per-contact solve records in sweep order (Contact.computeB,
// Contact.computeJMinvJt, the friction rule of PGS.java:128-142).
AMD_INLINE double contact_mu(const am_body_desc& d1, const am_body_desc& d2, int fovr, double fval) {
    if (fovr) return fval;
    double f1 = d1.friction, f2 = d2.friction;
    if (f1 < 0.2 || f2 < 0.2) return jmin(f1, f2);
    if (f1 > 1. || f2 > 1.) return jmax(f1, f2);
    return (f1 + f2) / 2.;
}

__global__ void k_setup(SetupArgs S) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *S.n_contacts) return;
    const BodiesView& B = S.B;
    const ContactsView& C = S.C;
    int c = S.order[k];
    int b1 = C.b1[c], b2 = C.b2[c];
    int a = (B.parent[b1] >= 0 && !S.in_collection) ? B.parent[b1] : b1;
    int b = (B.parent[b2] >= 0 && !S.in_collection) ? B.parent[b2] : b2;
    double* J = &C.J[(size_t)c * 36];
    // CollisionProcessor.updateJacobiansThatNeedUpdating (:326-334): contacts
    // touching a collection rebuild the world frame from body-1 coordinates.
    if (B.parent[b1] >= 0 || B.parent[b2] >= 0) {
        m33 R = ldm(&B.R[b1 * 9]);
        d3 x = ld3(&B.x[b1 * 3]);
        d3 pW = xform(R, x, ld3(&C.pB1[c * 3]));
        d3 nW = mul(R, ld3(&C.nB1[c * 3]));
        d3 t1 = mul(R, ld3(&C.t1B1[c * 3]));
        d3 t2 = mul(R, ld3(&C.t2B1[c * 3]));
        d3 r1 = sub(pW, ld3(&B.x[a * 3]));
        d3 r2 = sub(pW, ld3(&B.x[b * 3]));
        const d3 dirs[3] = {nW, t1, t2};
#pragma unroll
        for (int row = 0; row < 3; row++) {
            d3 d = dirs[row];
            st3(J + row * 12 + 0, mk(-1.0 * d.x, -1.0 * d.y, -1.0 * d.z));
            st3(J + row * 12 + 3, cross(d, r1));
            st3(J + row * 12 + 6, d);
            st3(J + row * 12 + 9, cross(r2, d));
        }
    }
    double* rec = &S.rec[(size_t)k * kRecDoubles];
    double Jr[36];
#pragma unroll
    for (int i = 0; i < 36; i++) Jr[i] = J[i];
    // b (Contact.computeB, Contact.java:279-326)
    double bn = 0, bt1 = 0, bt2 = 0;
    if (S.post) {
        bn = S.feedback * C.viol[c];
        bt1 = 0.;
        bt2 = 0.;
    } else {
        double restitution = (B.desc[b1].restitution + B.desc[b2].restitution) / 2.;
        if (S.rest_ovr) restitution = S.rest_val;
        const int ids[2] = {a, b};
#pragma unroll
        for (int s = 0; s < 2; s++) {
            int e = ids[s];
            const double* Jn = Jr + 0 + s * 6;
            const double* Jt1 = Jr + 12 + s * 6;
            const double* Jt2 = Jr + 24 + s * 6;
            double sc = B.minv[e] * S.dt;
            d3 tmp = scale_add(sc, ld3(&B.f[e * 3]), ld3(&B.v[e * 3]));
            bn += dot(tmp, ld3(Jn));
            bt1 += dot(tmp, ld3(Jt1));
            bt2 += dot(tmp, ld3(Jt2));
            tmp = mul(ldm(&B.jinv[e * 9]), ld3(&B.t[e * 3]));
            tmp = mk(tmp.x * S.dt, tmp.y * S.dt, tmp.z * S.dt);
            tmp = add(tmp, ld3(&B.w[e * 3]));
            bn += dot(tmp, ld3(Jn + 3));
            bt1 += dot(tmp, ld3(Jt1 + 3));
            bt2 += dot(tmp, ld3(Jt2 + 3));
            double bounce = dot(ld3(&B.v[e * 3]), ld3(Jn)) + dot(ld3(&B.w[e * 3]), ld3(Jn + 3));
            bounce *= restitution;
            bn += bounce;
        }
        bn += S.feedback * C.viol[c];
    }
    // D (Contact.computeJMinvJt, Contact.java:340-354)
    m33 ja_inv = ldm(&B.jinv[a * 9]);
    m33 jb_inv = ldm(&B.jinv[b * 9]);
    double D[3];
#pragma unroll
    for (int row = 0; row < 3; row++) {
        const double* ja = Jr + row * 12;
        const double* jb = Jr + row * 12 + 6;
        d3 t1 = mul(ja_inv, ld3(ja + 3));
        d3 t2 = mul(jb_inv, ld3(jb + 3));
        D[row] = B.minv[a] * dot(ld3(ja), ld3(ja)) + dot(ld3(ja + 3), t1) + B.minv[b] * dot(ld3(jb), ld3(jb)) +
                 dot(ld3(jb + 3), t2);
    }
#pragma unroll
    for (int i = 0; i < 36; i++) rec[i] = Jr[i];
    rec[36] = bn;
    rec[37] = bt1;
    rec[38] = bt2;
    rec[39] = D[0];
    rec[40] = D[1];
    rec[41] = D[2];
    rec[42] = D[0] + S.compliance;
    rec[43] = D[1] + S.compliance;
    rec[44] = D[2] + S.compliance;
    rec[45] = contact_mu(B.desc[b1], B.desc[b2], S.fric_ovr, S.fric_val);
    rec[46] = 0.0;
    rec[47] = 0.0;
#pragma unroll
    for (int r = 0; r < 3; r++) S.slam[k * 3 + r] = C.lam[c * 3 + r];
    S.rec_ab[k * 2 + 0] = a;
    S.rec_ab[k * 2 + 1] = b;
    bool clamp = !(B.flags[b1] & AM_BODY_MAGNET_ACTIVE) && !(B.flags[b2] & AM_BODY_MAGNET_ACTIVE);
    S.rec_flags[k] = clamp ? 1 : 0;
    S.sstate[k] = C.state[c];
}

void launch_setup(hipStream_t s, const SetupArgs& S, int cap) {
    if (cap <= 0) return;
    hipLaunchKernelGGL(k_setup, dim3((cap + 127) / 128), dim3(128), 0, s, S);
}

// Results back to list order.
__global__ void k_unpack(SetupArgs S) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *S.n_contacts) return;
    int c = S.order[k];
    const double* rec = &S.rec[(size_t)k * kRecDoubles];
    ContactsView C = S.C;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        C.lam[c * 3 + r] = S.slam[k * 3 + r];
        C.b[c * 3 + r] = rec[36 + r];
        C.D[c * 3 + r] = rec[39 + r];
    }
    C.state[c] = S.sstate[k];
}

void launch_unpack(hipStream_t s, const SetupArgs& S, int cap) {
    if (cap <= 0) return;
    hipLaunchKernelGGL(k_unpack, dim3((cap + 255) / 256), dim3(256), 0, s, S);
}

// ---------------------------------------------------------------------------
// The sweep.
struct BodyRegs {
    double dv[6];
    double minv;
    m33 jinv;
    bool dyn;
};

AMD_INLINE void load_body(const PgsArgs& P, int e, BodyRegs& r) {
#pragma unroll
    for (int i = 0; i < 6; i++) r.dv[i] = P.dv[e * 6 + i];
    r.minv = P.minv[e];
    r.jinv = ldm(&P.jinv[e * 9]);
    r.dyn = (P.bflags[e] & AM_BODY_PINNED) == 0;
}
AMD_INLINE void store_body(const PgsArgs& P, int e, const BodyRegs& r) {
    if (!r.dyn) return;
#pragma unroll
    for (int i = 0; i < 6; i++) P.dv[e * 6 + i] = r.dv[i];
}
AMD_INLINE void swap_regs(BodyRegs& a, BodyRegs& b) {
    BodyRegs t = a;
    a = b;
    b = t;
}

// PGS.updateDeltaVwithLambdai (PGS.java:221-245)
AMD_INLINE void apply_row(const double* row, double lambda, BodyRegs& A, BodyRegs& Bb) {
    double s1 = A.minv * lambda;
    A.dv[0] = s1 * row[0] + A.dv[0];
    A.dv[1] = s1 * row[1] + A.dv[1];
    A.dv[2] = s1 * row[2] + A.dv[2];
    d3 t = mul(A.jinv, mk(row[3], row[4], row[5]));
    A.dv[3] = lambda * t.x + A.dv[3];
    A.dv[4] = lambda * t.y + A.dv[4];
    A.dv[5] = lambda * t.z + A.dv[5];
    double s2 = Bb.minv * lambda;
    Bb.dv[0] = s2 * row[6] + Bb.dv[0];
    Bb.dv[1] = s2 * row[7] + Bb.dv[1];
    Bb.dv[2] = s2 * row[8] + Bb.dv[2];
    t = mul(Bb.jinv, mk(row[9], row[10], row[11]));
    Bb.dv[3] = lambda * t.x + Bb.dv[3];
    Bb.dv[4] = lambda * t.y + Bb.dv[4];
    Bb.dv[5] = lambda * t.z + Bb.dv[5];
}
// Contact.getJdv (Contact.java:362-378)
AMD_INLINE double jdv(const double* row, const BodyRegs& A, const BodyRegs& Bb) {
    double da = (row[0] * A.dv[0] + row[1] * A.dv[1] + row[2] * A.dv[2]) + (row[3] * A.dv[3] + row[4] * A.dv[4] + row[5] * A.dv[5]);
    double db = (row[6] * Bb.dv[0] + row[7] * Bb.dv[1] + row[8] * Bb.dv[2]) +
                (row[9] * Bb.dv[3] + row[10] * Bb.dv[4] + row[11] * Bb.dv[5]);
    return da + db;
}

// mode 0: confidentWarmStart (PGS.java:250-256); mode 1: one PGS pass over the group.
AMD_INLINE void process_group(const PgsArgs& P, int q, int mode, bool last_iter, double& mx) {
    const int k0 = P.pos_start[q], k1 = P.pos_start[q + 1];
    const int A = P.rec_ab[k0 * 2 + 0], Bi = P.rec_ab[k0 * 2 + 1];
    BodyRegs ra, rb;
    load_body(P, A, ra);
    load_body(P, Bi, rb);
    for (int k = k0; k < k1; k++) {
        const bool swapped = P.rec_ab[k * 2 + 0] != A;
        if (swapped) swap_regs(ra, rb);
        const double* rec = &P.rec[(size_t)k * kRecDoubles];
        double row[12];
        double lam[3] = {P.slam[k * 3 + 0], P.slam[k * 3 + 1], P.slam[k * 3 + 2]};
        if (mode == 0) {
#pragma unroll 1
            for (int r = 0; r < 3; r++) {
#pragma unroll
                for (int i = 0; i < 12; i++) row[i] = rec[r * 12 + i];
                apply_row(row, lam[r], ra, rb);
            }
        } else {
            const bool clamp = P.rec_flags[k] != 0;
            const double mu = rec[45];
            const double omega = P.omega;
#pragma unroll 1
            for (int r = 0; r < 3; r++) {
#pragma unroll
                for (int i = 0; i < 12; i++) row[i] = rec[r * 12 + i];
                double Jd = jdv(row, ra, rb);
                double prev = lam[r];
                double l = (rec[39 + r] * lam[r] - omega * (rec[36 + r] + Jd)) / rec[42 + r];
                if (clamp) {
                    if (r == 0) {
                        l = jmax(0, l);
                    } else {
                        double limit = mu * lam[0];
                        l = jmax(l, -limit);
                        l = jmin(l, limit);
                    }
                }
                lam[r] = l;
                double diff = l - prev;
                apply_row(row, diff, ra, rb);
                mx = jmax(mx, fabs(diff));
            }
            P.slam[k * 3 + 0] = lam[0];
            P.slam[k * 3 + 1] = lam[1];
            P.slam[k * 3 + 2] = lam[2];
            if (last_iter) {
                // Contact.updateContactState (Contact.java:385-398)
#pragma unroll
                for (int i = 0; i < 12; i++) row[i] = rec[12 + i];
                double w1 = rec[37] + jdv(row, ra, rb);
#pragma unroll
                for (int i = 0; i < 12; i++) row[i] = rec[24 + i];
                double w2 = rec[38] + jdv(row, ra, rb);
                int st;
                if (fabs(lam[0]) <= 1e-14) st = AM_STATE_BROKEN;
                else if (fabs(w1) > P.sliding) st = AM_STATE_ONEDGE;
                else if (fabs(w2) > P.sliding) st = AM_STATE_ONEDGE;
                else st = AM_STATE_CLEAR;
                P.sstate[k] = st;
            }
        }
        if (swapped) swap_regs(ra, rb);
    }
    store_body(P, A, ra);
    store_body(P, Bi, rb);
}

AMD_INLINE double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = jmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <bool kSingleBlock>
AMD_INLINE void grid_barrier() {
    if (kSingleBlock) {
        __syncthreads();
    } else {
        __threadfence();
        cg::this_grid().sync();
        __threadfence();
    }
}

// Persistent kernel: warm-start pass then up to `iterations` sweeps; one grid
// barrier per colour.  The per-sweep max |delta lambda| is reduced with an
// integer atomic max on the bit pattern (non-negative doubles order as their
// bits) into a slot per sweep.
template <bool kSingleBlock>
__global__ void k_pgs(PgsArgs P) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gstride = gridDim.x * blockDim.x;
    const int* cs = P.colour_start;
    // Colours with no group are skipped by every thread alike (same reads), so
    // barriers stay uniform.  The tail colour (kMaxColours) runs on one lane.
    int last_col = -1;
    for (int c = 0; c < kMaxColours; c++)
        if (cs[c + 1] > cs[c]) last_col = c;
    const bool has_tail = cs[kMaxColours + 1] > cs[kMaxColours];
    if (P.warm_start) {
        double dummy = 0;
        for (int c = 0; c <= last_col; c++) {
            if (cs[c + 1] == cs[c]) continue;
            for (int q = cs[c] + gtid; q < cs[c + 1]; q += gstride) process_group(P, q, 0, false, dummy);
            grid_barrier<kSingleBlock>();
        }
        if (has_tail) {
            if (gtid == 0)
                for (int q = cs[kMaxColours]; q < cs[kMaxColours + 1]; q++) process_group(P, q, 0, false, dummy);
            grid_barrier<kSingleBlock>();
        }
    }
    int it = 0;
    for (; it < P.iterations; it++) {
        const bool last = (it == P.iterations - 1);
        double mx = 0;
        for (int c = 0; c <= last_col; c++) {
            if (cs[c + 1] == cs[c]) continue;
            for (int q = cs[c] + gtid; q < cs[c + 1]; q += gstride) process_group(P, q, 1, last, mx);
            if (c == last_col && !has_tail) {
                double wm = wave_max(mx);
                if ((threadIdx.x & 63) == 0 && wm > 0)
                    atomicMax(&P.iter_max[it], (unsigned long long)__double_as_longlong(wm));
            }
            grid_barrier<kSingleBlock>();
        }
        if (has_tail) {
            if (gtid == 0)
                for (int q = cs[kMaxColours]; q < cs[kMaxColours + 1]; q++) process_group(P, q, 1, last, mx);
            double wm = wave_max(mx);
            if ((threadIdx.x & 63) == 0 && wm > 0)
                atomicMax(&P.iter_max[it], (unsigned long long)__double_as_longlong(wm));
            grid_barrier<kSingleBlock>();
        }
        if (last_col < 0 && !has_tail) grid_barrier<kSingleBlock>();
        double m = __longlong_as_double((long long)__hip_atomic_load(&P.iter_max[it], __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_AGENT));
        if (P.check_tolerance && m < P.tolerance) {
            it++;
            break;
        }
    }
    if (gtid == 0) *P.iters_used = it;
}

int pgs_max_blocks(int block, bool* ok) {
    int dev = 0, n_cu = 0, per_cu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pgs<false>, block, 0);
    *ok = per_cu > 0;
    return n_cu * (per_cu > 0 ? per_cu : 1);
}

hipError_t launch_pgs(hipStream_t s, const PgsArgs& P, int grid, int block) {
    if (grid <= 1) {
        hipLaunchKernelGGL(k_pgs<true>, dim3(1), dim3(block), 0, s, P);
        return hipGetLastError();
    }
    void* args[] = {(void*)&P};
    return hipLaunchCooperativeKernel((const void*)k_pgs<false>, dim3(grid), dim3(block), args, 0, s);
}

// ---------------------------------------------------------------------------
// Warm start (CollisionProcessor.java:443-665).
AMD_INLINE uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
struct CKey {
    int lo, hi, bvlo, bvhi, info;
};
// Contact.hashCode/equals (Contact.java:543-572): swap-invariant identity.
AMD_INLINE CKey make_key(int b1, int b2, int bv1, int bv2, int info) {
    CKey k;
    if (b1 < b2) {
        k.lo = b1; k.hi = b2; k.bvlo = bv1; k.bvhi = bv2;
    } else {
        k.lo = b2; k.hi = b1; k.bvlo = bv2; k.bvhi = bv1;
    }
    k.info = info;
    return k;
}
AMD_INLINE uint64_t key_hash(const CKey& k) {
    uint64_t h = mix64(((uint64_t)(uint32_t)k.lo << 32) | (uint32_t)k.hi);
    h = mix64(h ^ (((uint64_t)(uint32_t)k.bvlo << 32) | (uint32_t)k.bvhi));
    return mix64(h ^ (uint64_t)(uint32_t)k.info);
}
AMD_INLINE bool key_eq_old(const CKey& k, const OldContactsView& O, int i) {
    CKey o = make_key(O.b1[i], O.b2[i], O.bv1[i], O.bv2[i], O.info[i]);
    return k.lo == o.lo && k.hi == o.hi && k.bvlo == o.bvlo && k.bvhi == o.bvhi && k.info == o.info;
}

// updateContactsMap (CollisionProcessor.java:443-448): insert every old
// contact; for equal keys the later list position wins, like HashMap.put.
__global__ void k_map_insert(OldContactsView O, int* table) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= O.n) return;
    CKey k = make_key(O.b1[i], O.b2[i], O.bv1[i], O.bv2[i], O.info[i]);
    int s = (int)(key_hash(k) & (uint64_t)O.table_mask);
    for (;;) {
        int cur = atomicCAS(&table[s], -1, i);
        if (cur == -1) return;
        if (key_eq_old(k, O, cur)) {
            atomicMax(&table[s], i);
            return;
        }
        s = (s + 1) & O.table_mask;
    }
}

void launch_map_insert(hipStream_t s, const OldContactsView& O, int* table) {
    if (O.n <= 0) return;
    hipLaunchKernelGGL(k_map_insert, dim3((O.n + 255) / 256), dim3(256), 0, s, O, table);
}

AMD_INLINE int old_lookup(const OldContactsView& O, const CKey& k) {
    if (O.n <= 0) return -1;
    int s = (int)(key_hash(k) & (uint64_t)O.table_mask);
    for (;;) {
        int i = O.table[s];
        if (i < 0) return -1;
        if (key_eq_old(k, O, i)) return i;
        s = (s + 1) & O.table_mask;
    }
}

AMD_INLINE void take_old(ContactsView& C, int c, OldContactsView& O, int o, int post, bool zero_old) {
    C.is_new[c] = 0;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        double l = O.lam[o * 3 + r];
        C.lam[c * 3 + r] = l;
        C.lamw[c * 3 + r] = l;
    }
    if (zero_old) {
        O.lam[o * 3 + 0] = 0.0;
        O.lam[o * 3 + 1] = 0.0;
        O.lam[o * 3 + 2] = 0.0;
    }
    C.prev_viol[c] = post ? O.prev_viol[o] : O.viol[o];
}

AMD_INLINE d3 world_point(const BodiesView& B, int body, d3 pB1) {
    return xform(ldm(&B.R[body * 9]), ld3(&B.x[body * 3]), pB1);
}

// One lane per group, its contacts in list order: the box-box "take and zero"
// only involves contacts of the same body pair, so this order reproduces the
// reference's.
__global__ void k_warm_start(BodiesView B, ContactsView C, OldContactsView O, const int* n_groups,
                             const int* group_start, int post, int* counters) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= *n_groups) return;
    int bad = 0, rep = 0;
    for (int c = group_start[g]; c < group_start[g + 1]; c++) {
        int b1 = C.b1[c], b2 = C.b2[c];
        bool boxbox = B.desc[b1].geom == AM_GEOM_BOX && B.desc[b2].geom == AM_GEOM_BOX;
        CKey key = make_key(b1, b2, C.bv1[c], C.bv2[c], C.info[c]);
        if (!boxbox) {
            int o = old_lookup(O, key);
            if (o >= 0) take_old(C, c, O, o, post, false);
            else C.is_new[c] = 1;
            continue;
        }
        d3 pNew = world_point(B, b1, ld3(&C.pB1[c * 3]));
        int o = old_lookup(O, key);
        if (o >= 0) {
            d3 pOld = world_point(B, O.b1[o], ld3(&O.pB1[o * 3]));
            double dist = distance(pNew, pOld);
            if (dist > 0.05) {
                bad++;
                int myInfo = C.info[c];
                double best = dist;
                int bestInfo = myInfo, bestOld = o;
                for (int info = 0; info < 9; info++) {
                    if (info == myInfo) continue;
                    key.info = info;
                    int oo = old_lookup(O, key);
                    if (oo < 0) break;
                    pOld = world_point(B, O.b1[oo], ld3(&O.pB1[oo * 3]));
                    dist = distance(pNew, pOld);
                    if (dist < best) {
                        best = dist;
                        bestInfo = info;
                        bestOld = oo;
                    }
                }
                if (bestInfo != myInfo) rep++;
                o = bestOld;
                dist = best;
            }
            if (dist < 0.05) take_old(C, c, O, o, post, true);
            else C.is_new[c] = 1;
        } else {
            double best = 1.7976931348623157e308;
            int bestInfo = -1, bestOld = -1;
            for (int info = 0; info < 9; info++) {
                key.info = info;
                int oo = old_lookup(O, key);
                if (oo < 0) break;
                d3 pOld = world_point(B, O.b1[oo], ld3(&O.pB1[oo * 3]));
                double dist = distance(pNew, pOld);
                if (dist < best) {
                    best = dist;
                    bestInfo = info;
                    bestOld = oo;
                }
            }
            if (bestInfo != -1) {
                rep++;
                if (best < 0.05) take_old(C, c, O, bestOld, post, true);
            } else {
                C.is_new[c] = 1;
            }
        }
    }
    if (bad) atomicAdd(&counters[0], bad);
    if (rep) atomicAdd(&counters[1], rep);
}

void launch_warm_start(hipStream_t s, BodiesView B, ContactsView C, OldContactsView O, const int* n_groups, int cap,
                       const int* group_start, int post, int* counters) {
    if (cap <= 0) return;
    hipLaunchKernelGGL(k_warm_start, dim3((cap + 127) / 128), dim3(128), 0, s, B, C, O, n_groups, group_start, post,
                       counters);
}

// redoWarmStart / noWarmStart (CollisionProcessor.java:453-470)
__global__ void k_redo_warm(ContactsView C, const int* n, int zero) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= *n) return;
#pragma unroll
    for (int r = 0; r < 3; r++) C.lam[c * 3 + r] = zero ? 0.0 : C.lamw[c * 3 + r];
}
void launch_redo_warm(hipStream_t s, ContactsView C, const int* n, int cap, int zero) {
    if (cap <= 0) return;
    hipLaunchKernelGGL(k_redo_warm, dim3((cap + 255) / 256), dim3(256), 0, s, C, n, zero);
}

// ---------------------------------------------------------------------------
// Forces and integration.
// RigidBody.clear (RigidBody.java:226-230)
__global__ void k_clear(BodiesView B) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        B.f[i * 3 + a] = 0.0;
        B.t[i * 3 + a] = 0.0;
    }
#pragma unroll
    for (int a = 0; a < 6; a++) B.dv[i * 6 + a] = 0.0;
}
void launch_clear(hipStream_t s, BodiesView B) {
    if (B.n > 0) hipLaunchKernelGGL(k_clear, dim3((B.n + 255) / 256), dim3(256), 0, s, B);
}

// RigidBodySystem.applyGravityForce (RigidBodySystem.java:276-290): listed
// entities and the members of listed collections.
__global__ void k_gravity(BodiesView B, const int* targets, int n, d3 dir) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = targets[k];
    double m = -B.desc[i].mass;
    B.f[i * 3 + 0] = B.f[i * 3 + 0] + m * dir.x;
    B.f[i * 3 + 1] = B.f[i * 3 + 1] + m * dir.y;
    B.f[i * 3 + 2] = B.f[i * 3 + 2] + m * dir.z;
}
void launch_gravity(hipStream_t s, BodiesView B, const int* targets, int n, double gx, double gy, double gz) {
    if (n > 0) hipLaunchKernelGGL(k_gravity, dim3((n + 255) / 256), dim3(256), 0, s, B, targets, n, d3{gx, gy, gz});
}

// RigidBody.advanceVelocities (RigidBody.java:409-417)
__global__ void k_advance_velocities(BodiesView B, const int* targets, int n, double dt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = targets[k];
    double s = dt * B.minv[i];
    d3 v = scale_add(s, ld3(&B.f[i * 3]), ld3(&B.v[i * 3]));
    v = add(v, ld3(&B.dv[i * 6]));
    st3(&B.v[i * 3], v);
    d3 dom = mul(ldm(&B.jinv[i * 9]), ld3(&B.t[i * 3]));
    dom = mk(dom.x * dt, dom.y * dt, dom.z * dt);
    d3 w = add(ld3(&B.w[i * 3]), dom);
    w = add(w, ld3(&B.dv[i * 6 + 3]));
    st3(&B.w[i * 3], w);
}

// RigidBody.advancePositions (RigidBody.java:427-441): Rodrigues rotation
// (expRodrigues, RigidBody.java:382-401) and Matrix3d.normalizeCP.
__global__ void k_advance_positions(BodiesView B, const int* targets, int n, double dt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = targets[k];
    d3 v = ld3(&B.v[i * 3]);
    d3 x = scale_add(dt, v, ld3(&B.x[i * 3]));
    st3(&B.x[i * 3], x);
    d3 om = ld3(&B.w[i * 3]);
    double t = length(om) * dt;
    d3 d = normalized(om);
    m33 R = ldm(&B.R[i * 9]);
    if (t > 1e-8) {
        double c = cos(t), s = sin(t), c1 = 1 - c;
        double wX = d.x, wY = d.y, wZ = d.z;
        m33 dR;
        dR.m[0] = c + wX * wX * c1;
        dR.m[3] = wZ * s + wX * wY * c1;
        dR.m[6] = -wY * s + wX * wZ * c1;
        dR.m[1] = -wZ * s + wX * wY * c1;
        dR.m[4] = c + wY * wY * c1;
        dR.m[7] = wX * s + wY * wZ * c1;
        dR.m[2] = wY * s + wX * wZ * c1;
        dR.m[5] = -wX * s + wY * wZ * c1;
        dR.m[8] = c + wZ * wZ * c1;
        dR = mmul(dR, R);
        double mag = 1.0 / sqrt(dR.m[0] * dR.m[0] + dR.m[3] * dR.m[3] + dR.m[6] * dR.m[6]);
        R.m[0] = dR.m[0] * mag;
        R.m[3] = dR.m[3] * mag;
        R.m[6] = dR.m[6] * mag;
        mag = 1.0 / sqrt(dR.m[1] * dR.m[1] + dR.m[4] * dR.m[4] + dR.m[7] * dR.m[7]);
        R.m[1] = dR.m[1] * mag;
        R.m[4] = dR.m[4] * mag;
        R.m[7] = dR.m[7] * mag;
        R.m[2] = R.m[3] * R.m[7] - R.m[4] * R.m[6];
        R.m[5] = R.m[1] * R.m[6] - R.m[0] * R.m[7];
        R.m[8] = R.m[0] * R.m[4] - R.m[1] * R.m[3];
        stm(&B.R[i * 9], R);
    }
    if (B.desc[i].geom != AM_GEOM_COLLECTION) stm(&B.jinv[i * 9], rm0rt(R, ldm(B.desc[i].jinv0)));
}

void launch_advance(hipStream_t s, BodiesView B, const int* targets, int n, double dt, bool positions) {
    if (n <= 0) return;
    if (positions) hipLaunchKernelGGL(k_advance_positions, dim3((n + 127) / 128), dim3(128), 0, s, B, targets, n, dt);
    else hipLaunchKernelGGL(k_advance_velocities, dim3((n + 255) / 256), dim3(256), 0, s, B, targets, n, dt);
}

// jinv = R jinv0 R^T for non-pinned non-collection entities, zero for pinned.
__global__ void k_update_jinv(BodiesView B) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    if (B.flags[i] & AM_BODY_PINNED) {
#pragma unroll
        for (int a = 0; a < 9; a++) B.jinv[i * 9 + a] = 0.0;
        return;
    }
    if (B.desc[i].geom == AM_GEOM_COLLECTION) return;
    stm(&B.jinv[i * 9], rm0rt(ldm(&B.R[i * 9]), ldm(B.desc[i].jinv0)));
}
void launch_update_jinv(hipStream_t s, BodiesView B) {
    if (B.n > 0) hipLaunchKernelGGL(k_update_jinv, dim3((B.n + 255) / 256), dim3(256), 0, s, B);
}

// ---------------------------------------------------------------------------
// MotionMetricProcessor.getLargestVelocityNorm (MotionMetricProcessor.java:39-73)
AMD_INLINE d3 spatial_velocity(const BodiesView& B, int b, d3 pW) {
    d3 r = sub(pW, ld3(&B.x[b * 3]));
    return add(cross(ld3(&B.w[b * 3]), r), ld3(&B.v[b * 3]));
}
__global__ void k_metric(BodiesView B, const int* pairs, int n, double* out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int a = pairs[k * 2], b = pairs[k * 2 + 1];
    double largest = 0;
    if (b < 0) {
        m33 R = ldm(&B.R[a * 9]);
        d3 x = ld3(&B.x[a * 3]);
        for (int q = 0; q < 8; q++) {
            d3 pW = xform(R, x, ld3(&B.desc[a].bbox[q * 3]));
            largest = jmax(length(spatial_velocity(B, a, pW)), largest);
        }
    } else {
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
    }
    out[k] = largest;
}
void launch_metric(hipStream_t s, BodiesView B, const int* pairs, int n, double* out) {
    if (n > 0) hipLaunchKernelGGL(k_metric, dim3((n + 127) / 128), dim3(128), 0, s, B, pairs, n, out);
}

// ---------------------------------------------------------------------------
// Colouring driver pieces
void launch_incidence(hipStream_t s, bool fill, int n_groups, const int* ga, const int* gb, const int* bflags,
                      int* inc_count, const int* inc_start, int* inc_cursor, int* inc_items) {
    if (n_groups <= 0) return;
    dim3 g((n_groups + 255) / 256), b(256);
    if (!fill) hipLaunchKernelGGL(k_incidence_count, g, b, 0, s, n_groups, ga, gb, bflags, inc_count);
    else hipLaunchKernelGGL(k_incidence_fill, g, b, 0, s, n_groups, ga, gb, bflags, inc_start, inc_cursor, inc_items);
}
void launch_colour_round(hipStream_t s, int n_groups, const int* ga, const int* gb, const int* bflags,
                         const int* inc_start, const int* inc_count, const int* inc_items, const int* col_in,
                         int* col_out, int* remaining) {
    if (n_groups <= 0) return;
    hipLaunchKernelGGL(k_colour_round, dim3((n_groups + 255) / 256), dim3(256), 0, s, n_groups, ga, gb, bflags,
                       inc_start, inc_count, inc_items, col_in, col_out, remaining);
}
void launch_colour_keys(hipStream_t s, int n_groups, const int* colour, uint64_t* keys) {
    if (n_groups > 0)
        hipLaunchKernelGGL(k_colour_keys, dim3((n_groups + 255) / 256), dim3(256), 0, s, n_groups, colour, keys);
}
void launch_colour_layout(hipStream_t s, int n_groups, const uint64_t* sorted, const int* group_start,
                          int* colour_start, int* sizes, int* pos_start, int* order, int* d_total, void* scan_temp) {
    hipLaunchKernelGGL(k_colour_starts, dim3(1), dim3(128), 0, s, n_groups, sorted, colour_start);
    if (n_groups <= 0) return;
    dim3 g((n_groups + 255) / 256), b(256);
    hipLaunchKernelGGL(k_sorted_sizes, dim3((n_groups + 256) / 256), b, 0, s, n_groups, sorted, group_start, sizes);
    exclusive_scan(s, sizes, pos_start, n_groups + 1, d_total, scan_temp);
    hipLaunchKernelGGL(k_solve_order, g, b, 0, s, n_groups, sorted, group_start, pos_start, order);
}

}  // namespace am
