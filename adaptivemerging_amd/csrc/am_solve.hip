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
#include "am_solve.hpp"

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
__global__ void k_fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}
void fill_u64(hipStream_t s, void* p, size_t n, unsigned long long v) {
    if (n > 0)
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (unsigned long long*)p, n,
                           v);
}

// ---------------------------------------------------------------------------
// Groups: runs of contacts with the same unordered slot pair (solve bodies are
// the top-level entities), or with the same unordered physical body pair when
// slot and parent are null (the single sweep inside collections).
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
        // slot == nullptr: physical body pairs (computeInCollection)
        int a = slot ? slot[cb1[k]] : cb1[k], b = slot ? slot[cb2[k]] : cb2[k];
        int pa = slot ? slot[cb1[k - 1]] : cb1[k - 1], pb = slot ? slot[cb2[k - 1]] : cb2[k - 1];
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
        if (group_a) group_a[g] = (parent && parent[a] >= 0) ? parent[a] : a;
        if (group_b) group_b[g] = (parent && parent[b] >= 0) ? parent[b] : b;
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
//
// With layers (the single sweep's breadth-first layers), a group also waits
// for its neighbours of lower layers and takes a colour above all of theirs,
// so along every body the sweep visits the groups layer by layer; only
// neighbours of its own layer compete for colours.  Colours then may pass
// the parallel ones: those groups run in the sequential tail in colour order.
__global__ void k_colour_round(int n_groups, const int* ga, const int* gb, const int* bflags, const int* inc_start,
                               const int* inc_count, const int* inc_items, const int* col_in, int* col_out,
                               int* remaining, const int* layer, int* blocker) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    int c = col_in[g];
    if (c >= 0) {
        col_out[g] = c;
        return;
    }
    const int lg = layer ? layer[g] : 0;
    int lo = 0;  // the colour must be at least lo (above every lower-layer neighbour)
    // The neighbour that kept this group waiting in an earlier round keeps it
    // waiting while it is uncoloured (priorities and layers do not change),
    // so the incidence lists are scanned again only once it has its colour.
    const int bl = blocker[g];
    bool ready = !(bl >= 0 && col_in[bl] < 0);
    const int bodies[2] = {ga[g], gb[g]};
    if (layer) {
        for (int s = 0; s < 2 && ready; s++) {
            int e = bodies[s];
            if (!dynamic_body(bflags, e)) continue;
            int st = inc_start[e], cnt = inc_count[e];
            for (int q = 0; q < cnt; q++) {
                int h = inc_items[st + q];
                if (h == g || layer[h] >= lg) continue;
                int ch = col_in[h];
                if (ch < 0) {
                    ready = false;
                    blocker[g] = h;
                    break;
                }
                lo = max(lo, ch + 1);
            }
        }
    }
    uint64_t used = 0;  // colours lo .. lo + 63 taken by same-layer neighbours
    for (int s = 0; s < 2 && ready; s++) {
        int e = bodies[s];
        if (!dynamic_body(bflags, e)) continue;
        int st = inc_start[e], cnt = inc_count[e];
        for (int q = 0; q < cnt; q++) {
            int h = inc_items[st + q];
            if (h == g || (layer && layer[h] != lg)) continue;
            int ch = col_in[h];
            if (ch < 0) {
                if (higher((uint32_t)h, (uint32_t)g)) {
                    ready = false;
                    blocker[g] = h;
                    break;
                }
            } else if (ch >= lo && ch - lo < 64) {
                used |= (uint64_t)1 << (ch - lo);
            }
        }
    }
    int pick = -1;
    if (ready) {
        if (!layer) {
            pick = kMaxColours;  // sequential tail if all 64 parallel colours are taken
            if (~used != 0) pick = __ffsll((long long)~used) - 1;
        } else {
            pick = lo + (~used != 0 ? __ffsll((long long)~used) - 1 : 64);
        }
    }
    col_out[g] = pick;
    // some group is still uncoloured: one flag store per wave
    const unsigned long long waiting = __ballot(!ready);
    if (waiting && (int)(threadIdx.x & 63) == __ffsll((long long)__ballot(1)) - 1) *remaining = 1;
}

// Sort key of a group: colour, then (inside a parallel colour, where groups
// commute) larger groups first, then group index.  Equal-sized groups end up
// side by side, so a step of consecutive groups wastes few contact slots.
constexpr uint32_t kKeyGroupMask = 0x0FFFFFFFu;
__global__ void k_colour_keys(int n_groups, const int* colour, const int* group_start, uint64_t* keys) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    const int c = colour[g];
    const uint32_t size = (uint32_t)min(group_start[g + 1] - group_start[g], 15);
    const uint32_t rank = c < kMaxColours ? 15u - size : 0u;
    keys[g] = ((uint64_t)(uint32_t)c << 32) | (rank << 28) | (uint32_t)g;
}

// colour_start[c] = first sorted position with colour >= c, c = 0..kMaxColours;
// colour_start[kMaxColours + 1] = n_groups (the tail holds every colour from
// kMaxColours on, in colour order)
__global__ void k_colour_starts(int n_groups, const uint64_t* sorted, int* colour_start) {
    int c = threadIdx.x;
    if (c > kMaxColours + 1) return;
    if (c == kMaxColours + 1) {
        colour_start[c] = n_groups;
        return;
    }
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
    int g = (int)((uint32_t)sorted[q] & kKeyGroupMask);
    sizes[q] = group_start[g + 1] - group_start[g];
}

__global__ void k_solve_order(int n_groups, const uint64_t* sorted, const int* group_start, const int* pos_start,
                              const int* group_a, const int* group_b, int* order, int* group_desc, int* pos_a,
                              int* spos) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_groups) return;
    int g = (int)((uint32_t)sorted[q] & kKeyGroupMask);
    spos[g] = q;
    int s = group_start[g], e = group_start[g + 1], o = pos_start[q];
    for (int k = s; k < e; k++) {
        order[o + (k - s)] = k;
        pos_a[o + (k - s)] = group_a[g];
    }
    reinterpret_cast<int4*>(group_desc)[q] = make_int4(o, o + (e - s), group_a[g], group_b[g]);
}

// Dataflow ranks: for every dynamic body a with deg_a incident groups, the rank
// of each of its groups among them in sorted (colour-major) order.  flow[q] =
// {rank_a, deg_a, rank_b, deg_b} for sorted group q with solve bodies A, B
// (deg 0 for a pinned body: never waited on, never written).  pred[2 q + s]
// is the incidence (2 q' + s') of the update of that body just before this
// one: the group ranked one lower, or for rank 0 the body's last group, whose
// update of the previous pass comes first.
__global__ void k_flow_ranks(int n_ent, const int* bflags, const int* inc_start, const int* inc_count,
                             const int* inc_items, const int* spos, const int* group_desc, int* flow, int* pred) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_ent || !dynamic_body(bflags, a)) return;
    const int s = inc_start[a], d = inc_count[a];
    auto incidence = [&](int q) { return 2 * q + (group_desc[4 * q + 2] == a ? 0 : 1); };
    for (int i = 0; i < d; i++) {
        const int qi = spos[inc_items[s + i]];
        int r = 0, before = -1, last = -1;
        for (int j = 0; j < d; j++) {
            const int qj = spos[inc_items[s + j]];
            if (qj < qi) {
                r++;
                before = max(before, qj);
            }
            last = max(last, qj);
        }
        const int side = group_desc[4 * qi + 2] == a ? 0 : 1;
        flow[4 * qi + 2 * side] = r;
        flow[4 * qi + 2 * side + 1] = d;
        pred[2 * qi + side] = incidence(before >= 0 ? before : last);
    }
}

void launch_flow_ranks(hipStream_t s, int n_ent, int n_groups, const int* bflags, const int* inc_start,
                       const int* inc_count, const int* inc_items, const int* spos, const int* group_desc,
                       int* flow, int* pred) {
    fill_int(s, flow, 4 * n_groups, 0);
    fill_int(s, pred, 2 * n_groups, 0);
    if (n_ent > 0 && n_groups > 0)
        hipLaunchKernelGGL(k_flow_ranks, dim3((n_ent + 255) / 256), dim3(256), 0, s, n_ent, bflags, inc_start,
                           inc_count, inc_items, spos, group_desc, flow, pred);
}

// The division (D lambda - omega (b + J.dv)) / (D + C) of a row.  The
// denominator is fixed during a solve, so the denominator-only half of the
// hardware division sequence (reciprocal estimate and two Newton steps) is
// done once per row in setup; per sweep the quotient is then q0 = n y,
// r = fma(-d, q0, n), q = fma(r, y, q0), the same operations the sequence
// performs when v_div_scale leaves its operands unscaled and v_div_fixup has
// nothing to fix.  Both |n| and |d| in [2^-250, 2^250] guarantees that (no
// exponent gap near 768, no denormal quotient or reciprocal, nonzero finite
// operands); anything else takes the full division.  The result is the
// correctly rounded quotient either way.
AMD_INLINE double div_recip(double den) {
    const double a = fabs(den);
    if (!(a >= 0x1p-250 && a <= 0x1p250)) return 0.0;  // 0: always divide
    double y = __builtin_amdgcn_rcp(den);
    double e = __builtin_fma(-den, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-den, y, 1.0);
    return __builtin_fma(y, e, y);
}
AMD_INLINE double div_row(double num, double den, double y) {
    const double a = fabs(num);
    const bool fast = y != 0.0 && a >= 0x1p-250 && a <= 0x1p250;
    const double q0 = num * y;
    const double r = __builtin_fma(-den, q0, num);
    double q = __builtin_fma(r, y, q0);
    if (__builtin_expect(!__all(fast), 0)) {
        if (!fast) q = num / den;
    }
    return q;
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
    // D (Contact.computeJMinvJt, Contact.java:340-354); the jinv products are
    // kept for the sweep's deltaV updates (PGS.updateDeltaVwithLambdai).
    m33 ja_inv = ldm(&B.jinv[a * 9]);
    m33 jb_inv = ldm(&B.jinv[b * 9]);
    double D[3];
    d3 ta[3], tb[3];
#pragma unroll
    for (int row = 0; row < 3; row++) {
        const double* ja = Jr + row * 12;
        const double* jb = Jr + row * 12 + 6;
        ta[row] = mul(ja_inv, ld3(ja + 3));
        tb[row] = mul(jb_inv, ld3(jb + 3));
        D[row] = B.minv[a] * dot(ld3(ja), ld3(ja)) + dot(ld3(ja + 3), ta[row]) + B.minv[b] * dot(ld3(jb), ld3(jb)) +
                 dot(ld3(jb + 3), tb[row]);
    }
    // Records are oriented so that their first body is the group's first
    // body A; a contact listed the other way round stores jb.w, ja.v (= -jb.v
    // exactly), ja.w and the swapped jinv products.  Every product of the
    // update is unchanged and J.dv only swaps its two (commuting) halves, so
    // the sweep stays bit-exact without swapping body registers per contact.
    const bool flip = a != S.pos_a[k];
    double* rec = &S.rec[(size_t)k * kRecDoubles];
#pragma unroll
    for (int row = 0; row < 3; row++) {
        const double* J0 = Jr + row * 12;
        double* q = rec + row * kRecRow;
        if (!flip) {
            st3(q + 0, ld3(J0 + 3));   // ja.w
            st3(q + 3, ld3(J0 + 6));   // jb.v
            st3(q + 6, ld3(J0 + 9));   // jb.w
            st3(q + 9, ta[row]);
            st3(q + 12, tb[row]);
        } else {
            st3(q + 0, ld3(J0 + 9));   // jb.w
            st3(q + 3, ld3(J0 + 0));   // ja.v = -jb.v
            st3(q + 6, ld3(J0 + 3));   // ja.w
            st3(q + 9, tb[row]);
            st3(q + 12, ta[row]);
        }
    }
    const int ea = flip ? b : a, eb = flip ? a : b;
    rec[kRecB + 0] = bn;
    rec[kRecB + 1] = bt1;
    rec[kRecB + 2] = bt2;
    rec[kRecD + 0] = D[0];
    rec[kRecD + 1] = D[1];
    rec[kRecD + 2] = D[2];
    rec[kRecMu] = contact_mu(B.desc[b1], B.desc[b2], S.fric_ovr, S.fric_val);
#pragma unroll
    for (int r = 0; r < 3; r++) rec[kRecLambda + r] = C.lam[c * 3 + r];
    bool clamp = !(B.flags[b1] & AM_BODY_MAGNET_ACTIVE) && !(B.flags[b2] & AM_BODY_MAGNET_ACTIVE);
    rec[kRecAB] = __longlong_as_double((long long)(((uint64_t)(uint32_t)eb << 32) | (uint32_t)ea));
    bool recips = true;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double den = rec[kRecD + r] + S.compliance;
        rec[kRecDen + r] = den;
        rec[kRecY + r] = div_recip(den);
        recips = recips && rec[kRecY + r] != 0.0;
    }
    // 0: not clamped; 1: clamped, every row's reciprocal stored (the
    // dataflow kernel's fast update may take it); 2: clamped otherwise
    rec[kRecClamp] = !clamp ? 0.0 : recips ? 1.0 : 2.0;
    rec[kRecDoubles - 1] = 0.0;
    reinterpret_cast<int4*>(S.meta)[k] = make_int4(ea, eb, clamp ? 1 : 0, 0);
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
    double* rec = &S.rec[(size_t)k * kRecDoubles];
    ContactsView C = S.C;
    const int used = S.scene_used ? S.scene_used[S.B.desc[S.C.b1[c]].scene] : *S.iters_used;
    if (S.lam_alt && (used & 1)) {
        // dataflow schedule: an odd last sweep left lambda in the odd-sweep buffer
#pragma unroll
        for (int r = 0; r < 3; r++) rec[kRecLambda + r] = S.lam_alt[4 * k + r];
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        C.lam[c * 3 + r] = rec[kRecLambda + r];
        C.b[c * 3 + r] = rec[kRecB + r];
        C.D[c * 3 + r] = rec[kRecD + r];
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
    bool dyn;
};

// deltaV crosses workgroups between colours: it is read and written with
// agent-scope relaxed atomics (coherent at device level, global_load/store
// ... sc1), so the barrier needs no L2 write-back or invalidation and the
// read-only contact records stay resident in each XCD's L2 across sweeps.
AMD_INLINE double ld_coherent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
AMD_INLINE void st_coherent(double* p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

AMD_INLINE void load_body(const PgsArgs& P, int e, BodyRegs& r) {
#pragma unroll
    for (int i = 0; i < 6; i++) r.dv[i] = ld_coherent(&P.dv[e * 6 + i]);
    r.minv = P.minv[e];
    r.dyn = (P.bflags[e] & AM_BODY_PINNED) == 0;
}
AMD_INLINE void store_body(const PgsArgs& P, int e, const BodyRegs& r) {
    if (!r.dyn) return;
#pragma unroll
    for (int i = 0; i < 6; i++) st_coherent(&P.dv[e * 6 + i], r.dv[i]);
}

// One row of a contact record: aw = ja.w, bv = jb.v (ja.v = -jb.v),
// bw = jb.w, ta = jinv_a ja.w, tb = jinv_b jb.w.
struct Row {
    d3 aw, bv, bw, ta, tb;
};
AMD_INLINE Row row_of(const double* rec, int r) {
    const double* q = rec + r * kRecRow;
    return Row{mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), mk(q[6], q[7], q[8]), mk(q[9], q[10], q[11]),
               mk(q[12], q[13], q[14])};
}

// PGS.updateDeltaVwithLambdai (PGS.java:221-245)
AMD_INLINE void apply_row(const Row& R, double lambda, BodyRegs& A, BodyRegs& Bb) {
    double s1 = A.minv * lambda;
    A.dv[0] = s1 * -R.bv.x + A.dv[0];
    A.dv[1] = s1 * -R.bv.y + A.dv[1];
    A.dv[2] = s1 * -R.bv.z + A.dv[2];
    A.dv[3] = lambda * R.ta.x + A.dv[3];
    A.dv[4] = lambda * R.ta.y + A.dv[4];
    A.dv[5] = lambda * R.ta.z + A.dv[5];
    double s2 = Bb.minv * lambda;
    Bb.dv[0] = s2 * R.bv.x + Bb.dv[0];
    Bb.dv[1] = s2 * R.bv.y + Bb.dv[1];
    Bb.dv[2] = s2 * R.bv.z + Bb.dv[2];
    Bb.dv[3] = lambda * R.tb.x + Bb.dv[3];
    Bb.dv[4] = lambda * R.tb.y + Bb.dv[4];
    Bb.dv[5] = lambda * R.tb.z + Bb.dv[5];
}
// Contact.getJdv (Contact.java:362-378)
AMD_INLINE double jdv(const Row& R, const BodyRegs& A, const BodyRegs& Bb) {
    double da = (-R.bv.x * A.dv[0] + -R.bv.y * A.dv[1] + -R.bv.z * A.dv[2]) +
                (R.aw.x * A.dv[3] + R.aw.y * A.dv[4] + R.aw.z * A.dv[5]);
    double db = (R.bv.x * Bb.dv[0] + R.bv.y * Bb.dv[1] + R.bv.z * Bb.dv[2]) +
                (R.bw.x * Bb.dv[3] + R.bw.y * Bb.dv[4] + R.bw.z * Bb.dv[5]);
    return da + db;
}

constexpr int kRec2 = kRecDoubles / 2;  // 16-byte words per record
constexpr int kRecLoad = 58;            // the coloured sweep reads up to the clamp flag
struct Rec {
    double v[kRecLoad];
};
AMD_INLINE void load_rec(const double2* src, Rec& r) {
#pragma unroll
    for (int i = 0; i < kRecLoad / 2; i++) {
        double2 t = src[i];
        r.v[2 * i] = t.x;
        r.v[2 * i + 1] = t.y;
    }
}

// One group (the contacts of one body pair): mode 0 is confidentWarmStart
// (PGS.java:250-256), mode 1 one PGS pass (PGS.java:100-184).  recs points at
// the group's first record; lambda and the contact state are written back at
// sweep position k.
AMD_INLINE void process_group(const PgsArgs& P, int4 gd, const double2* recs, int mode, bool last_iter, double& mx) {
    const int k0 = gd.x, k1 = gd.y, A = gd.z, Bi = gd.w;
    BodyRegs ra, rb;
    load_body(P, A, ra);
    load_body(P, Bi, rb);
    bool nan = false;
    for (int k = k0; k < k1; k++) {
        Rec cur;
        load_rec(recs + (size_t)(k - k0) * kRec2, cur);
        const double* rec = cur.v;
        double lam[3] = {rec[kRecLambda + 0], rec[kRecLambda + 1], rec[kRecLambda + 2]};
        if (mode == 0) {
#pragma unroll
            for (int r = 0; r < 3; r++) apply_row(row_of(rec, r), lam[r], ra, rb);
        } else {
            const bool clamp = rec[kRecClamp] != 0.0;
            const double mu = rec[kRecMu];
            const double omega = P.omega;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                Row R = row_of(rec, r);
                double Jd = jdv(R, ra, rb);
                double prev = lam[r];
                double Dr = rec[kRecD + r];
                double l = (Dr * lam[r] - omega * (rec[kRecB + r] + Jd)) / (Dr + P.compliance);
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
                apply_row(R, diff, ra, rb);
                // Math.max over |diff| (PGS.java:125); a NaN is carried separately
                mx = fmax(mx, fabs(diff));
                nan = nan || diff != diff;
            }
            double* dst = &P.rec[(size_t)k * kRecDoubles + kRecLambda];
            dst[0] = lam[0];
            dst[1] = lam[1];
            dst[2] = lam[2];
            if (last_iter) {
                // Contact.updateContactState (Contact.java:385-398)
                double w1 = rec[kRecB + 1] + jdv(row_of(rec, 1), ra, rb);
                double w2 = rec[kRecB + 2] + jdv(row_of(rec, 2), ra, rb);
                int st;
                if (fabs(lam[0]) <= 1e-14) st = AM_STATE_BROKEN;
                else if (fabs(w1) > P.sliding) st = AM_STATE_ONEDGE;
                else if (fabs(w2) > P.sliding) st = AM_STATE_ONEDGE;
                else st = AM_STATE_CLEAR;
                P.sstate[k] = st;
            }
        }
    }
    if (nan) mx = __longlong_as_double(0x7ff8000000000000ll);
    store_body(P, A, ra);
    store_body(P, Bi, rb);
}

AMD_INLINE double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = jmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Grid barrier for a co-resident (cooperatively launched) grid.  The caller
// has already waited for its deltaV stores (s_waitcnt) and passed a workgroup
// barrier, so lane 0 arrives on a monotonically increasing counter right
// away and spins on a coherent load until all blocks of this instance have
// arrived.  A spin limit turns a barrier that cannot complete into an error
// flag.
template <bool kSingleBlock>
AMD_INLINE bool grid_barrier(const PgsArgs& P, unsigned& epoch) {
    if (kSingleBlock) return true;
    __shared__ int bail;
    epoch += gridDim.x;
    if (threadIdx.x == 0) {
        bail = 0;
        __hip_atomic_fetch_add(P.barrier, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        long spins = 0;
        while (__hip_atomic_load(P.barrier, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            if (++spins > 20000000L) {
                atomicOr(P.error, 1);
                bail = 1;
                break;
            }
            // another workgroup gave up: leave now instead of at our own limit
            if ((spins & 1023) == 0 && __hip_atomic_load(P.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                break;
            __builtin_amdgcn_s_sleep(1);
        }
        if (__hip_atomic_load(P.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) bail = 1;
    }
    __syncthreads();
    return bail == 0;
}

// Persistent kernel: warm-start pass then up to `iterations` sweeps; one grid
// barrier per non-empty colour.  Groups of a colour are dealt block-cyclically
// (group j of the colour goes to block j % grid, lane j / grid) so every
// workgroup, hence every CU in use, gets a share.  The per-sweep max |delta
// lambda| is reduced with an integer atomic max on the bit pattern (non-
// negative doubles order as their bits) into a slot per sweep.
template <bool kSingleBlock>
__global__ __launch_bounds__(256) void k_pgs(PgsArgs P) {
    const int* cs = P.colour_start;
    const int slot0 = threadIdx.x * gridDim.x + blockIdx.x;
    const int stride = gridDim.x * blockDim.x;
    const int4* gdesc = reinterpret_cast<const int4*>(P.group_desc);
    const double2* grec = reinterpret_cast<const double2*>(P.rec);
    __shared__ int s_cols[kMaxColours];
    __shared__ int s_ncols;
    if (threadIdx.x == 0) {
        int nc = 0;
        for (int c = 0; c < kMaxColours; c++)
            if (cs[c + 1] > cs[c]) s_cols[nc++] = c;
        s_ncols = nc;
    }
    __syncthreads();
    const int ncols = s_ncols;
    const bool has_tail = cs[kMaxColours + 1] > cs[kMaxColours];
    const bool lane0 = blockIdx.x == 0 && threadIdx.x == 0;
    unsigned epoch = 0;
    // this lane's first group of the next colour: its descriptor does not
    // depend on other workgroups, so it is read before the barrier
    int4 head = make_int4(0, 0, 0, 0);
    bool have = false;
    auto prefetch = [&](int ci) {
        const int c = s_cols[ci];
        const int q = cs[c] + slot0;
        have = q < cs[c + 1];
        if (have) head = gdesc[q];
    };
    int it = 0;
    // one group; with per-scene exits a sweep skips the groups of a scene
    // whose previous sweep was under the tolerance (that sweep is final
    // after the grid barrier), and each group's max goes to its scene too
    auto run_group = [&](int q, int4 gd, int mode, bool last, double& mx) {
        if (!P.per_scene || mode == 0) {
            process_group(P, gd, grec + (size_t)gd.x * kRec2, mode, last, mx);
            return;
        }
        const int sc = P.g_scene[q];
        if (it > 0) {
            const double prev = __longlong_as_double((long long)__hip_atomic_load(
                &P.scene_max[(size_t)sc * P.iterations + it - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (prev < P.tolerance) return;
        }
        double gmx = 0.0;
        process_group(P, gd, grec + (size_t)gd.x * kRec2, mode, last, gmx);
        if (!(gmx <= 0))
            atomicMax(&P.scene_max[(size_t)sc * P.iterations + it], (unsigned long long)__double_as_longlong(gmx));
        mx = gmx != gmx ? gmx : fmax(mx, gmx);
    };
    auto run_colour = [&](int ci, int mode, bool last, double& mx) {
        const int c = s_cols[ci];
        if (have) run_group(cs[c] + slot0, head, mode, last, mx);
        for (int q = cs[c] + slot0 + stride; q < cs[c + 1]; q += stride) run_group(q, gdesc[q], mode, last, mx);
    };
    auto run_tail = [&](int mode, bool last, double& mx) {
        if (lane0)
            for (int q = cs[kMaxColours]; q < cs[kMaxColours + 1]; q++) run_group(q, gdesc[q], mode, last, mx);
    };
    // Ends a colour: wait for this lane's deltaV stores, read the next
    // colour's descriptor, meet the other workgroups.
    auto next = [&](int ci_next) -> bool {
        __builtin_amdgcn_s_waitcnt(0);
        if (ci_next >= 0) prefetch(ci_next);
        __syncthreads();
        return grid_barrier<kSingleBlock>(P, epoch);
    };

    if (ncols > 0) prefetch(0);
    const int first_after_tail = ncols > 0 ? 0 : -1;
    if (P.warm_start) {
        double dummy = 0;
        for (int ci = 0; ci < ncols; ci++) {
            run_colour(ci, 0, false, dummy);
            const int nx = ci + 1 < ncols ? ci + 1 : (has_tail ? -1 : first_after_tail);
            if (!next(nx)) return;
        }
        if (has_tail) {
            run_tail(0, false, dummy);
            if (!next(first_after_tail)) return;
        }
    }
    for (; it < P.iterations; it++) {
        const bool last = (it == P.iterations - 1);
        double mx = 0;
        for (int ci = 0; ci < ncols; ci++) {
            run_colour(ci, 1, last, mx);
            if (ci == ncols - 1 && !has_tail) {
                double wm = wave_max(mx);
                if ((threadIdx.x & 63) == 0 && !(wm <= 0))
                    atomicMax(&P.iter_max[it], (unsigned long long)__double_as_longlong(wm));
            }
            const int nx = ci + 1 < ncols ? ci + 1 : (has_tail ? -1 : first_after_tail);
            if (!next(nx)) return;
        }
        if (has_tail) {
            run_tail(1, last, mx);
            double wm = wave_max(mx);
            if ((threadIdx.x & 63) == 0 && !(wm <= 0))
                atomicMax(&P.iter_max[it], (unsigned long long)__double_as_longlong(wm));
            if (!next(first_after_tail)) return;
        }
        if (ncols == 0 && !has_tail) {
            if (!next(-1)) return;
        }
        double m = __longlong_as_double((long long)__hip_atomic_load(&P.iter_max[it], __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_AGENT));
        if (P.check_tolerance && m < P.tolerance) {
            it++;
            break;
        }
    }
    if (lane0) *P.iters_used = it;
}

int pgs_max_blocks(int block, bool* ok) {
    int dev = 0, n_cu = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pgs<false>, block, 0);
    *ok = per_cu > 0;
    return n_cu * (per_cu > 0 ? per_cu : 1);
}

hipError_t launch_pgs(hipStream_t s, const PgsArgs& P, int grid, int block, bool plain) {
    if (grid <= 1) {
        hipLaunchKernelGGL(k_pgs<true>, dim3(1), dim3(block), 0, s, P);
        return hipGetLastError();
    }
    if (plain) {
        // the caller clamps grid to the co-resident limit; the barrier's spin
        // limit turns a grid that is not co-resident into an error
        hipLaunchKernelGGL(k_pgs<false>, dim3(grid), dim3(block), 0, s, P);
        return hipGetLastError();
    }
    void* args[] = {(void*)&P};
    return hipLaunchCooperativeKernel((const void*)k_pgs<false>, dim3(grid), dim3(block), args, 0, s);
}

// ---------------------------------------------------------------------------
// Dataflow schedule of the same sweeps.
//
// Each dynamic body receives its group updates in sorted (colour-major) order,
// the order the coloured schedule and the exported sweep order define, and
// groups on disjoint bodies commute: the result is bit-identical to the
// coloured schedule, with no grid-wide barrier.  The order is kept by handing
// a body's deltaV from each update to the next through slots: sorted group q
// writes its updated deltaV of each dynamic body to its own slot (incidence
// 2 q + side, one set of slots per pass parity), and the body's next update
// (one rank higher, or for rank 0 the body's last group of the previous pass:
// pred[]) polls that slot until it no longer holds the empty pattern, takes
// the value and empties the slot for its next writer.  Pass 0 is
// confidentWarmStart, passes 1..iterations are sweeps.
//
// A wave runs steps of up to four groups of one colour, one group per 16-lane
// row (lanes hold one deltaV component each, quads A.v, A.w, B.v, B.w), so a
// row's J.dv is a product per lane and a fixed-shape DPP reduction in the
// reference's summation order.  Wave w runs steps w, w + W, ... pass by pass
// in order; the global minimum (pass, step) not yet done always has its
// inputs, so with every wave resident the schedule cannot deadlock, and
// bounded spins turn a wave that stops making progress into an error flag
// instead of a hang.
//
// Early exit (PGS.java:190) needs the max |dl| of sweep t before sweep t+1 is
// final.  Sweeps double-buffer lambda (odd sweeps in lam_alt) and deltaV (odd
// passes in dv_alt), and a wave starts sweep p only once sweep p-2 is known not
// to have converged, so at most one sweep runs speculatively and its buffers
// are simply not selected.  The last sweep (contact states) waits for the
// verdict on the one before it, so it never runs speculatively.

AMD_INLINE double* slot_at(const FlowArgs& F, int parity, int incidence) {
    return F.slot + ((size_t)parity * 2 * F.n_groups + incidence) * kSlotDoubles;
}
AMD_INLINE unsigned long long ld_cnt64(const unsigned long long* p) {
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
AMD_INLINE int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
// DPP move of a double (bound_ctrl: lanes without a source read 0)
template <int kCtrl>
AMD_INLINE double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, kCtrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), kCtrl, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
constexpr int kQuadLane0 = 0x00, kQuadLane1 = 0x55, kQuadLane2 = 0xAA;  // quad_perm broadcasts
constexpr int kRowHalfMirror = 0x141, kRowMirror = 0x140;

// Contact.getJdv (Contact.java:362-378) for the group held by a 16-lane row.
// Quads of the row hold A.v, A.w, B.v, B.w (lanes 0-2 of each quad, lane 3
// idle), p = J_i dv_i.  Every lane of the row gets the reference's sum
// ((p0 + p1) + p2) per quad, (v + w) per body, then A + B; the partner quads
// come through mirrors, which only swap the operands of an addition, and IEEE
// addition commutes.
AMD_INLINE double row_jdv(double p) {
    const double s = (dpp_d<kQuadLane0>(p) + dpp_d<kQuadLane1>(p)) + dpp_d<kQuadLane2>(p);
    const double h = s + dpp_d<kRowHalfMirror>(s);
    return h + dpp_d<kRowMirror>(h);
}

constexpr long kFlowSpinLimit = 1L << 22;  // polls of one wait (each a memory round trip)

// Per-contact operands of one row: this lane's J entries and update
// coefficients of the three rows, and the contact's uniform values.
struct WaveRec {
    double jc[6];  // as stored: the A.v lanes flip the sign on use (J)
    unsigned sign;
    double b[3], d[3], den[3], y[3], mu, clamp, lam[3];
};
// Where a lane's entries sit in a record: quads hold A.v, A.w, B.v, B.w;
// J entries ja.v = -jb.v, ja.w, jb.v, jb.w; coefficients -jb.v, jinv_a ja.w,
// jb.v, jinv_b jb.w (the v lanes' coefficient is their J entry, scaled by
// minv*dl at update time).
struct LaneMap {
    int jofs, cofs;
    unsigned sign;  // sign bit flipped on A.v lanes
};
AMD_INLINE LaneMap lane_map(int li) {
    const int q = li >> 2, j = (li & 3) < 3 ? (li & 3) : 0;
    LaneMap m;
    m.jofs = (q == 1 ? 0 : q == 3 ? 6 : 3) + j;
    m.cofs = q == 1 ? 9 + j : q == 3 ? 12 + j : m.jofs;
    m.sign = q == 0 ? 0x80000000u : 0u;
    return m;
}
AMD_INLINE double flip(double v, unsigned sign) {
    const long long b = __double_as_longlong(v) ^ ((long long)sign << 32);
    return __longlong_as_double(b);
}
AMD_INLINE const double* lam_src(const FlowArgs& F, int k, int p) {
    const bool from_alt = p >= 1 && ((p - 1) & 1);
    return from_alt ? &F.lam_alt[4 * k] : &F.P.rec[(size_t)k * kRecDoubles + kRecLambda];
}
// rec: the row's record (LDS or global); lam: its lambda for this pass
AMD_INLINE double J(const WaveRec& c, int i) { return flip(c.jc[i], c.sign); }
AMD_INLINE void read_wave_rec(const double* rec, const double* lam, const LaneMap& m, WaveRec& w) {
    w.sign = m.sign;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        w.jc[2 * r] = rec[r * kRecRow + m.jofs];
        w.jc[2 * r + 1] = rec[r * kRecRow + m.cofs];
        w.b[r] = rec[kRecB + r];
        w.d[r] = rec[kRecD + r];
        w.den[r] = rec[kRecDen + r];
        w.y[r] = rec[kRecY + r];
        w.lam[r] = lam[r];
    }
    w.mu = rec[kRecMu];
    w.clamp = rec[kRecClamp];
}

// A step's records are copied into LDS with DMA loads (16 B per lane, lane
// l's bytes at 16 l) before the wave waits for its dependencies: per contact
// slot, two blocks of two rows' 512-byte records, and a block with each row's
// lambda when the pass reads the odd-sweep buffer.
constexpr int kDmaSlots = 8;
typedef __attribute__((address_space(3))) void lds_void;
struct DmaSlot {
    double2 rec[2][64];
    double2 lam[64];
};

// A contact's new lambdas go to this pass's buffer; the last sweep also sets
// its state (Contact.updateContactState, Contact.java:385-398).
AMD_INLINE double* lam_dst(const FlowArgs& F, int k, int p) {
    return (p & 1) ? &F.lam_alt[4 * k] : &F.P.rec[(size_t)k * kRecDoubles + kRecLambda];
}
AMD_INLINE void contact_state(const FlowArgs& F, const WaveRec& c, int k, bool last, bool lead,
                              const double (&lam)[3], double dv) {
    const PgsArgs& P = F.P;
    if (last) {
        const double w1 = c.b[1] + row_jdv(J(c, 2) * dv);
        const double w2 = c.b[2] + row_jdv(J(c, 4) * dv);
        if (lead) {
            int st;
            if (fabs(lam[0]) <= 1e-14) st = AM_STATE_BROKEN;
            else if (fabs(w1) > P.sliding) st = AM_STATE_ONEDGE;
            else if (fabs(w2) > P.sliding) st = AM_STATE_ONEDGE;
            else st = AM_STATE_CLEAR;
            P.sstate[k] = st;
        }
    }
}
AMD_INLINE void contact_done(const FlowArgs& F, const WaveRec& c, int k, int p, bool last, bool lead,
                             const double (&lam)[3], double dv) {
    // every lane of the row stores the same values to the same place
    double* dst = lam_dst(F, k, p);
    dst[0] = lam[0];
    dst[1] = lam[1];
    dst[2] = lam[2];
    contact_state(F, c, k, last, lead, lam, dv);
}

// One contact of a row-held group in pass p (PGS.java:100-184; pass 0 is
// confidentWarmStart, PGS.java:250-256).  dv is this lane's deltaV component,
// f its body's minv on v lanes and 1 on w lanes (PGS.java:221-245 scales the
// v part by minv*dl and the w part by dl; 1*dl == dl exactly).  mx is the
// running max |dl| with NaN tracked apart (Math.max keeps a NaN).
AMD_INLINE void row_contact(const FlowArgs& F, const WaveRec& c, int k, int p, bool last, bool lead, double f,
                            double& dv, double& mx, bool& nan) {
    const PgsArgs& P = F.P;
    double lam[3] = {c.lam[0], c.lam[1], c.lam[2]};
    if (p == 0) {
#pragma unroll
        for (int r = 0; r < 3; r++) dv = (f * lam[r]) * J(c, 2 * r + 1) + dv;
        return;
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double jd = row_jdv(J(c, 2 * r) * dv);
        const double prev = lam[r];
        const double num = c.d[r] * lam[r] - P.omega * (c.b[r] + jd);
        double l = div_row(num, c.den[r], c.y[r]);
        // projection (PGS.java:111-113, 150-158), selected without a branch.
        // Math.max(0, l) is (l <= 0) ? +0 : l for every l (NaN and -0
        // included).  For a limit that is neither NaN nor -0, Math.max(l,
        // -limit) and Math.min(., limit) are the plain selects below (NaN l
        // falls through both, and the only zero tie, -0 against +0, resolves
        // the Java way); other limits take the full Java forms.
        double lc;
        if (r == 0) {
            lc = l <= 0.0 ? 0.0 : l;
        } else {
            const double limit = c.mu * lam[0];
            const bool plain = limit > 0.0 || __double_as_longlong(limit) == 0;
            const double a = l < -limit ? -limit : l;
            lc = a > limit ? limit : a;
            if (__builtin_expect(!__all(plain), 0)) {
                if (!plain) lc = jmin(jmax(l, -limit), limit);
            }
        }
        l = c.clamp != 0.0 ? lc : l;
        lam[r] = l;
        const double diff = l - prev;
        dv = (f * diff) * J(c, 2 * r + 1) + dv;
        mx = fmax(mx, fabs(diff));
        nan = nan || diff != diff;
    }
    contact_done(F, c, k, p, last, lead, lam, dv);
}

// The same update where the step has been found to need none of the general
// forms: every row's |numerator| in [2^-250, 2^250) with a nonzero stored
// reciprocal (the quotient is then q0 = n y, r = fma(-d, q0, n), q = fma(r, y,
// q0), exactly the division), every friction limit a finite value >= +0 (the
// plain selects are then Math.max/Math.min), and every contact clamped (its
// record's clamp flag 1, which also says every row's reciprocal is stored).  The
// conditions are gathered per lane into FastCheck without a branch; the step
// checks them once at its end and, if any failed, runs again from its start
// with row_contact.
struct FastCheck {
    unsigned num_hi_max = 0;            // high words of |numerator|
    unsigned num_hi_min = 0xffffffffu;
    unsigned limit_hi_max = 0;          // high words of the friction limits (sign included)
    bool plain = true;                  // every contact's clamp flag 1
    AMD_INLINE bool ok() const {
        return plain && num_hi_min >= 0x30500000u && num_hi_max < 0x4F900000u && limit_hi_max < 0x7FF00000u;
    }
};
AMD_INLINE unsigned hi_word(double v) { return (unsigned)((unsigned long long)__double_as_longlong(v) >> 32); }
AMD_INLINE void row_contact_fast(const FlowArgs& F, const WaveRec& c, int k, int p, bool last, bool lead, double f,
                                 double& dv, double& mx, FastCheck& chk, double (&lam)[3]) {
    const PgsArgs& P = F.P;
    lam[0] = c.lam[0];
    lam[1] = c.lam[1];
    lam[2] = c.lam[2];
    chk.plain = chk.plain && c.clamp == 1.0;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double jd = row_jdv(J(c, 2 * r) * dv);
        const double prev = lam[r];
        const double num = c.d[r] * lam[r] - P.omega * (c.b[r] + jd);
        const unsigned h = hi_word(num) & 0x7fffffffu;
        chk.num_hi_max = max(chk.num_hi_max, h);
        chk.num_hi_min = min(chk.num_hi_min, h);
        const double q0 = num * c.y[r];
        const double rr = __builtin_fma(-c.den[r], q0, num);
        double l = __builtin_fma(rr, c.y[r], q0);
        if (r == 0) {
            // l is finite and nonzero here (nonzero finite quotient), so
            // max(l, +0) is Math.max(0, l)
            l = __builtin_fmax(l, 0.0);
        } else {
            const double limit = c.mu * lam[0];
            chk.limit_hi_max = max(chk.limit_hi_max, hi_word(limit));
            const double a = l < -limit ? -limit : l;
            l = a > limit ? limit : a;
        }
        lam[r] = l;
        const double diff = l - prev;
        dv = (f * diff) * J(c, 2 * r + 1) + dv;
        // finite whenever the step's conditions hold (a NaN or infinite
        // numerator fails them), so no NaN is tracked here
        mx = fmax(mx, fabs(diff));
    }
    contact_state(F, c, k, last, lead, lam, dv);
}

// Steps: up to 4 consecutive sorted groups of one parallel colour (no shared
// dynamic body), one per 16-lane row of a wave.  A colour with fewer groups
// than there are waves, and the sequential tail colour, give one group per
// step: a step starts when all its groups' inputs are there, and the long
// chains of dependent updates run through the small late colours, where a
// group would otherwise wait for its step-mates' inputs.  Wave w runs steps
// w, w + W, ... of every pass in order.
constexpr int kRowsPerWave = 4;

// kScenes: a batched solve with per-scene exits (P.per_scene set); the
// single-scene kernel carries none of their code.
template <bool kScenes>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_pgs_flow(FlowArgs F) {
    const PgsArgs& P = F.P;
    const int* cs = P.colour_start;
    __shared__ int s_sp[kMaxColours + 2];  // first step of each colour, then the total
    __shared__ int s_per[kMaxColours + 1];  // groups per step of each colour
    const int W = gridDim.x;
    if (threadIdx.x == 0) {
        int n = 0;
        for (int c = 0; c <= kMaxColours; c++) {
            s_sp[c] = n;
            const int g = cs[c + 1] - cs[c];
            const int per = c < kMaxColours && g >= W ? kRowsPerWave : 1;
            s_per[c] = per;
            n += (g + per - 1) / per;
        }
        s_sp[kMaxColours + 1] = n;
    }
    __syncthreads();
    const int n_steps = s_sp[kMaxColours + 1];
    const int w = blockIdx.x;
    const int lane = threadIdx.x;
    const int row = lane >> 4, li = lane & 15, quad = li >> 2, jj = li & 3;
    const bool lead = li == 0;
    const bool mine = jj < 3;               // holds a deltaV component
    const bool side_b = quad >= 2;
    const bool vlane = (quad & 1) == 0;     // a linear-velocity component
    const int comp = (quad & 1) * 3 + (jj < 3 ? jj : 0);
    const int first = P.warm_start ? 0 : 1;
    const int lastp = P.iterations;
    const int4* gdesc = reinterpret_cast<const int4*>(P.group_desc);
    const int4* fdesc = reinterpret_cast<const int4*>(F.flow);
    __shared__ DmaSlot s_dma[kDmaSlots];
    __shared__ double s_lam[kRowsPerWave][kDmaSlots][4];  // fast-update lambdas of the step
    const int2* gpred = reinterpret_cast<const int2*>(F.pred);
    const LaneMap lmap = lane_map(li);
    int chk = 0;  // sweeps 1..chk are known not to have converged

    // per-scene exits: the scene counter this lane's row still has to add
    // its group to (added once the step after it has waited for its memory
    // operations, or at the end of the pass)
    unsigned long long* pend_done = nullptr;
    unsigned long long pend_add = 0;
    for (int p = first; p <= lastp; p++) {
        // pass gates: sweep p-2 must be known not converged (p-1 for the last sweep)
        if (P.check_tolerance) {
            const int need = (p == lastp && p >= 2) ? p - 1 : p - 2;
            long spins = 0;
            while (chk < need) {
                // the waves done with sweep chk + 1 and those over the
                // tolerance there (the sweep converged if none was), summed
                // over the sweep's words (read together)
                unsigned long long wd = 0;
                const unsigned long long* pw = &F.pass_done[(size_t)(chk + 1) * kPassWords];
#pragma unroll
                for (int q = 0; q < kPassWords; q++) wd += ld_cnt64(&pw[q]);
                const unsigned d = (unsigned)uniform((int)(unsigned)(wd >> 32));
                if (d == F.n_waves) {
                    if (uniform((int)(unsigned)wd) == 0) return;  // converged at sweep chk + 1
                    chk++;
                    spins = 0;
                    continue;
                }
                if (++spins > kFlowSpinLimit ||
                    ((spins & 255) == 0 && uniform(__hip_atomic_load(P.error, __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_AGENT)))) {
                    if (lane == 0) atomicOr(P.error, 1);
                    return;
                }
            }
        }
        double mx = 0.0;
        bool nan = false;
        double* cur = (p & 1) ? F.dv_alt : P.dv;
        double* prv = (p & 1) ? P.dv : F.dv_alt;
        int c = 0;  // colour of the current step (steps ascend)
        // this row's group of a step (the colour cursor advances with st)
        auto step_group = [&](int st, bool& act) {
            while (s_sp[c + 1] <= st) c++;
            const int per = s_per[c];
            const int q0 = cs[c] + (st - s_sp[c]) * per;
            const int count = min(per, cs[c + 1] - q0);
            act = row < count;
            return q0 + (act ? row : 0);
        };
        bool nx_active = false;
        int4 nx_gd = make_int4(0, 0, 0, 0), nx_fd = make_int4(0, 0, 0, 0);
        int2 nx_pr = make_int2(0, 0);
        int nx_sc = 0, nx_q = 0;
        if (w < n_steps) {
            const int q = step_group(w, nx_active);
            nx_q = q;
            nx_gd = gdesc[q];
            nx_fd = fdesc[q];
            nx_pr = gpred[q];
            if (kScenes) nx_sc = P.g_scene[q];
        }
        for (int st = w; st < n_steps; st += W) {
            // descriptors of this step were loaded during the previous one
            const bool has_group = nx_active;
            const int4 gd = nx_gd;
            const int4 fd = nx_fd;
            const int2 pr = nx_pr;
            const int sc = nx_sc, qg = nx_q;
            if (st + W < n_steps) {
                const int q = step_group(st + W, nx_active);
                nx_q = q;
                nx_gd = gdesc[q];
                nx_fd = fdesc[q];
                nx_pr = gpred[q];
                if (kScenes) nx_sc = P.g_scene[q];
            }
            // Per-scene exits: a row whose scene was under the tolerance at
            // sweep p-2 or p-3 (or p-1 before the last sweep) skips this
            // pass.  A skipped sweep leaves the scene's max at 0, so with
            // p-3 checked too a scene that skipped once skips every later
            // pass (its first sweep under the tolerance is final; a later
            // sweep over it must not bring back a scene whose groups
            // published nothing for the pass before).  Lane 0 of the row
            // learns it while the wave polls for its inputs below: it waits
            // until every group of the scene has finished those sweeps, then
            // reads the scene's max.  Every group it waits for comes earlier
            // in the (pass, step) order, so the earliest undone step can
            // always run.  Until then the row counts as active (its records
            // are copied and its inputs polled for).
            bool active = has_group;
            int n = active ? gd.y - gd.x : 0;
            int nmax = n;
            nmax = max(nmax, __shfl_xor(nmax, 16, 64));
            nmax = max(nmax, __shfl_xor(nmax, 32, 64));
            // the step's records do not depend on other waves: copy them into
            // LDS while the wave waits.  Lane l of a record block copies word
            // l & 31 of row 2i + (l >> 5)'s record.
            const int staged = min(nmax, kDmaSlots);
            const bool odd_lam = p >= 1 && ((p - 1) & 1);
            {
                int kb[kRowsPerWave], nl[kRowsPerWave];
#pragma unroll
                for (int r = 0; r < kRowsPerWave; r++) {
                    kb[r] = __builtin_amdgcn_readlane(gd.x, r * 16);
                    nl[r] = max(__builtin_amdgcn_readlane(n, r * 16) - 1, 0);
                }
                const bool upper = lane >= 32;
                for (int t = 0; t < staged; t++) {
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        const int r0 = 2 * i, r1 = 2 * i + 1;
                        const int k = upper ? kb[r1] + min(t, nl[r1]) : kb[r0] + min(t, nl[r0]);
                        const double* src = P.rec + (size_t)k * kRecDoubles + 2 * (lane & 31);
                        __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)&s_dma[t].rec[i][0], 16, 0, 0);
                    }
                    if (odd_lam) {
                        const int k = gd.x + min(t, max(n - 1, 0));  // this lane's row
                        const double* src = &F.lam_alt[4 * k + (li == 1 ? 2 : 0)];
                        __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)&s_dma[t].lam[0], 16, 0, 0);
                    }
                }
            }
            // Each lane of a dynamic body takes its deltaV component from the
            // slot of that body's previous update (this pass's, or for the
            // body's first group the previous pass's last): it polls until
            // the slot is filled, then empties it for the slot's next writer.
            // The body's first update of the launch, and pinned bodies, read
            // the launch buffers instead.
            const int e = side_b ? gd.w : gd.z;
            const int rank = side_b ? fd.z : fd.x;
            const int deg = side_b ? fd.w : fd.y;
            double dv = 0.0, f = 1.0;
            bool dyn = false;
            double* src = nullptr;
            if (active && mine) {
                if (vlane) f = P.minv[e];
                dyn = deg != 0;
                if (dyn && !(p == first && rank == 0))
                    src = slot_at(F, rank == 0 ? ((p - 1) & 1) : (p & 1), side_b ? pr.y : pr.x) + comp;
                else
                    dv = ld_coherent(&prv[e * 6 + comp]);
            }
            {
                bool got = src == nullptr;
                long spins = 0;
                auto spin = [&]() -> bool {  // false: give up (error flag set)
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > kFlowSpinLimit ||
                        ((spins & 255) == 0 && uniform(__hip_atomic_load(P.error, __ATOMIC_RELAXED,
                                                                          __HIP_MEMORY_SCOPE_AGENT)))) {
                        if (lane == 0) atomicOr(P.error, 1);
                        __builtin_amdgcn_s_waitcnt(0);  // this step's LDS copies land before the wave ends
                        return false;
                    }
                    return true;
                };
                auto poll = [&]() {
                    if (!got) {
                        const double v = ld_coherent(src);
                        if ((unsigned long long)__double_as_longlong(v) != kSlotEmpty) {
                            dv = v;
                            got = true;
                        }
                    }
                };
                const int t1 = p - 2, t2 = (p == lastp) ? p - 1 : 0;
                if (!(kScenes && p >= 1 && P.check_tolerance && (t1 >= 1 || t2 >= 1))) {
                    while (true) {
                        poll();
                        if (__all(got)) break;
                        if (!spin()) return;
                    }
                } else {
                    // lane 0 of each row resolves its scene's exit check
                    // while the row polls; a skipped row reads no inputs (its
                    // scene's other groups skip this pass too and publish none)
                    bool resolved = !(has_group && li == 0), skip = false;
                    const unsigned want = has_group ? P.scene_groups[sc] : 0u;
                    const unsigned long long* cnt = &P.scene_done[(size_t)sc * (P.iterations + 1)];
                    // sweeps checked: t1, t1 - 1 and t2 where they exist (a
                    // missing one repeats t1 or t2)
                    const int ta = t1 >= 1 ? t1 : t2, tb = t1 >= 2 ? t1 - 1 : ta, tc = t2 >= 1 ? t2 : ta;
                    while (true) {
                        // the slot and the scene's words are read together; a
                        // sweep's word is final once all the scene's groups
                        // have counted themselves in it, and the scene was
                        // under the tolerance there if none was over it
                        double v = 0.0;
                        if (!got) v = ld_coherent(src);
                        unsigned long long wa = 0, wb = 0, wc = 0;
                        if (!resolved) {
                            wa = ld_cnt64(&cnt[ta]);
                            wb = ld_cnt64(&cnt[tb]);
                            wc = ld_cnt64(&cnt[tc]);
                        }
                        if (!got && (unsigned long long)__double_as_longlong(v) != kSlotEmpty) {
                            dv = v;
                            got = true;
                        }
                        if (!resolved && (unsigned)(wa >> 32) == want && (unsigned)(wb >> 32) == want &&
                            (unsigned)(wc >> 32) == want) {
                            skip = (unsigned)wa == 0u || (unsigned)wb == 0u || (unsigned)wc == 0u;
                            resolved = true;
                        }
                        const bool row_resolved = __shfl(resolved, row * 16, 64);
                        const bool row_skip = __shfl(skip, row * 16, 64);
                        if (__all(row_resolved && (row_skip || got))) break;
                        if (!spin()) return;
                    }
                    if (__shfl(skip, row * 16, 64) && has_group) {
                        active = false;
                        n = 0;
                        src = nullptr;
                    }
                }
            }
            const bool last = p == lastp;
            __builtin_amdgcn_s_waitcnt(0);  // LDS copies landed
            if (kScenes && pend_done) {
                // the previous step's scene max has landed (the wait above)
                __hip_atomic_fetch_add(pend_done, pend_add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                pend_done = nullptr;
            }
            // empty the slot read above only now, so that this wait does not
            // include the store (the publish below waits for it)
            if (src) st_coherent(src, __longlong_as_double((long long)kSlotEmpty));
            // rows with fewer contacts idle; contacts past the LDS slots
            // (groups of more than kDmaSlots contacts) read global memory
            double smx = 0.0;  // this step's max |dl| and NaN flag (row-uniform)
            bool snan = false;
            // Contacts in the LDS slots and the rest (read from global
            // memory) run in separate loops, so that the first one waits on
            // LDS reads only and never on the lambda stores in flight.
            const int in_lds = min(nmax, kDmaSlots);
            auto lds_rec = [&](int t, WaveRec& rc) {
                const double* rl = reinterpret_cast<const double*>(&s_dma[t].rec[row >> 1][(row & 1) * 32]);
                const double* ll = odd_lam ? reinterpret_cast<const double*>(&s_dma[t].lam[row * 16]) : rl + kRecLambda;
                read_wave_rec(rl, ll, lmap, rc);
            };
            auto global_rec = [&](int t, WaveRec& rc) {
                const int k = gd.x + min(t, max(n - 1, 0));
                read_wave_rec(P.rec + (size_t)k * kRecDoubles, lam_src(F, k, p), lmap, rc);
            };
            // sweeps try the fast update first, which keeps the lambdas of the
            // contacts in the LDS slots in LDS; the general update stores its
            // lambdas at once (over the few the fast one stored directly, in
            // the same places from the same lanes, if the step runs again)
            bool general = p == 0;
            if (!general) {
                const double dv_start = dv;
                FastCheck fc;
                // the next contact's record is read while this one runs
                // (two buffers used in turn, so nothing is copied)
                WaveRec ra, rb;
                if (in_lds > 0) lds_rec(0, ra);
                double lam[3];
                for (int t = 0; t < in_lds; t += 2) {
                    if (t + 1 < in_lds) lds_rec(t + 1, rb);
                    if (t < n) {
                        row_contact_fast(F, ra, gd.x + t, p, last, lead, f, dv, smx, fc, lam);
                        // every lane of the row writes the same values to the same place
                        s_lam[row][t][0] = lam[0];
                        s_lam[row][t][1] = lam[1];
                        s_lam[row][t][2] = lam[2];
                    }
                    if (t + 1 >= in_lds) break;
                    if (t + 2 < in_lds) lds_rec(t + 2, ra);
                    if (t + 1 < n) {
                        row_contact_fast(F, rb, gd.x + t + 1, p, last, lead, f, dv, smx, fc, lam);
                        s_lam[row][t + 1][0] = lam[0];
                        s_lam[row][t + 1][1] = lam[1];
                        s_lam[row][t + 1][2] = lam[2];
                    }
                }
                for (int t = kDmaSlots; t < nmax; t++) {
                    global_rec(t, ra);
                    if (t < n) {
                        row_contact_fast(F, ra, gd.x + t, p, last, lead, f, dv, smx, fc, lam);
                        double* dst = lam_dst(F, gd.x + t, p);
                        dst[0] = lam[0];
                        dst[1] = lam[1];
                        dst[2] = lam[2];
                    }
                }
                if (__builtin_expect(!__all(fc.ok()), 0)) {
                    general = true;
                    dv = dv_start;
                    smx = 0.0;
                    snan = false;
                    __builtin_amdgcn_s_waitcnt(0);
                }
            }
            if (general) {
                for (int t = 0; t < in_lds; t++) {
                    WaveRec rc;
                    lds_rec(t, rc);
                    if (t < n) row_contact(F, rc, gd.x + t, p, last, lead, f, dv, smx, snan);
                }
                for (int t = kDmaSlots; t < nmax; t++) {
                    WaveRec rc;
                    global_rec(t, rc);
                    if (t < n) row_contact(F, rc, gd.x + t, p, last, lead, f, dv, smx, snan);
                }
            }
            mx = fmax(mx, smx);
            nan = nan || snan;
            // Publish deltaV to this group's slots.  The emptied slot must be
            // seen empty before these values can be seen (the chain of
            // updates they start leads to that slot's next writer), so the
            // stores wait for this wave's earlier memory operations; with the
            // fast update's lambdas still in LDS, only that store is left.
            // The body's last update of the pass also goes to the pass's
            // buffer, which holds the result when this is the last sweep.
            __builtin_amdgcn_s_waitcnt(0);
            if (active && mine && dyn) {
                st_coherent(slot_at(F, p & 1, 2 * qg + (side_b ? 1 : 0)) + comp, dv);
                if (rank == deg - 1) st_coherent(&cur[e * 6 + comp], dv);
            }
            if (!general && active && li < 3) {
                const int nl = min(n, kDmaSlots);
                for (int t = 0; t < nl; t++) {
                    lam_dst(F, gd.x + t, p)[li] = s_lam[row][t][li];
                }
            }
            // per-scene exits, after the publish (which is what later steps
            // wait for): the scene's max for this sweep now, the group's
            // count once a later wait has seen the max land
            if (kScenes && p >= 1 && has_group && lead) {
                if (active && !(smx <= 0))
                    atomicMax(&P.scene_max[(size_t)sc * P.iterations + p - 1],
                              (unsigned long long)__double_as_longlong(
                                  snan ? __longlong_as_double(0x7ff8000000000000ll) : smx));
                pend_done = &P.scene_done[(size_t)sc * (P.iterations + 1) + p];
                // over the tolerance (NaN included), as the scene's max would be
                const bool over = active && !(smx <= 0) && (snan || !(smx < P.tolerance));
                pend_add = (1ull << 32) | (over ? 1ull : 0ull);
            }
        }
        if (kScenes && pend_done) {
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_fetch_add(pend_done, pend_add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pend_done = nullptr;
        }
        if (p >= 1) {
            if (__any(nan)) mx = __longlong_as_double(0x7ff8000000000000ll);
            mx = wave_max(mx);
            if (lane == 0) {
                if (!(mx <= 0))  // NaN included
                    atomicMax(&P.iter_max[p - 1], (unsigned long long)__double_as_longlong(mx));
                // the wait also lands this wave's lambda stores before its next
                // pass reads them back
                __builtin_amdgcn_s_waitcnt(0);
                const bool over = !(mx <= 0) && !(mx < P.tolerance);  // NaN over
                __hip_atomic_fetch_add(&F.pass_done[(size_t)p * kPassWords + (w % kPassWords)],
                                       (1ull << 32) | (over ? 1ull : 0ull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Sweeps used (the first converged sweep, else all), then the buffers of the
// last pass become the results.
__global__ void k_flow_finish(PgsArgs P) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int used = P.iterations;
    if (P.check_tolerance)
        for (int t = 1; t <= P.iterations; t++)
            if (__longlong_as_double((long long)P.iter_max[t - 1]) < P.tolerance) {
                used = t;
                break;
            }
    *P.iters_used = used;
}
__global__ void k_flow_select_dv(double* dv, const double* dv_alt, int n, const int* iters_used,
                                 const int* scene_used, const am_body_desc* desc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int used = scene_used ? scene_used[desc[i / 6].scene] : *iters_used;
    if (used & 1) dv[i] = dv_alt[i];
}

// Sweeps used per scene: the first sweep whose max |dl| over the scene was
// under the tolerance, else all (a scene's skipped sweeps read 0).
__global__ void k_scene_finish(PgsArgs P) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n_scenes) return;
    int used = P.iterations;
    for (int t = 1; t <= P.iterations; t++)
        if (__longlong_as_double((long long)P.scene_max[(size_t)s * P.iterations + t - 1]) < P.tolerance) {
            used = t;
            break;
        }
    P.scene_used[s] = used;
}
void launch_scene_finish(hipStream_t s, const PgsArgs& P) {
    if (P.per_scene && P.n_scenes > 0)
        hipLaunchKernelGGL(k_scene_finish, dim3((P.n_scenes + 63) / 64), dim3(64), 0, s, P);
}

// Groups are counted per scene in an LDS histogram per workgroup (when the
// scenes fit), so each workgroup adds to a scene's global count once instead
// of every group contending for its scene's counter.
constexpr int kSceneHist = 4096;
__global__ __launch_bounds__(1024) void k_group_scene(const int* group_desc, int n_groups, const int* order,
                                                      const int* cb1, const am_body_desc* desc, int n_scenes,
                                                      int* g_scene, unsigned* scene_groups) {
    __shared__ unsigned hist[kSceneHist];
    const bool local = n_scenes <= kSceneHist;
    if (local)
        for (int i = threadIdx.x; i < n_scenes; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n_groups; q += gridDim.x * blockDim.x) {
        const int sc = desc[cb1[order[group_desc[q * 4]]]].scene;
        g_scene[q] = sc;
        if (local) atomicAdd(&hist[sc], 1u);
        else atomicAdd(&scene_groups[sc], 1u);
    }
    __syncthreads();
    if (local)
        for (int i = threadIdx.x; i < n_scenes; i += blockDim.x)
            if (hist[i]) atomicAdd(&scene_groups[i], hist[i]);
}
void launch_group_scene(hipStream_t s, const int* group_desc, int n_groups, const int* order, const int* cb1,
                        const am_body_desc* desc, int n_scenes, int* g_scene, unsigned* scene_groups) {
    if (n_groups > 0)
        hipLaunchKernelGGL(k_group_scene, dim3(std::min((n_groups + 4095) / 4096, 256)), dim3(1024), 0, s, group_desc,
                           n_groups, order, cb1, desc, n_scenes, g_scene, scene_groups);
}

int pgs_flow_max_blocks(bool* ok) {
    int dev = 0, n_cu = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    int per_cu_scenes = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pgs_flow<false>, 64, 0);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_scenes, k_pgs_flow<true>, 64, 0);
    per_cu = std::min(per_cu, per_cu_scenes);
    *ok = per_cu > 0;
    return n_cu * (per_cu > 0 ? per_cu : 1);
}

hipError_t launch_pgs_flow(hipStream_t s, const FlowArgs& F, int grid, bool plain, int n_dv, const BodiesView& B) {
    if (plain) {
        if (F.P.per_scene) hipLaunchKernelGGL(k_pgs_flow<true>, dim3(grid), dim3(64), 0, s, F);
        else hipLaunchKernelGGL(k_pgs_flow<false>, dim3(grid), dim3(64), 0, s, F);
    } else {
        const void* kernel = F.P.per_scene ? (const void*)k_pgs_flow<true> : (const void*)k_pgs_flow<false>;
        void* args[] = {(void*)&F};
        hipError_t e = hipLaunchCooperativeKernel(kernel, dim3(grid), dim3(64), args, 0, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_flow_finish, dim3(1), dim3(64), 0, s, F.P);
    launch_scene_finish(s, F.P);
    if (n_dv > 0)
        hipLaunchKernelGGL(k_flow_select_dv, dim3((n_dv + 255) / 256), dim3(256), 0, s, F.P.dv, F.dv_alt, n_dv,
                           F.P.iters_used, F.P.per_scene ? F.P.scene_used : nullptr, B.desc);
    return hipGetLastError();
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

// Whether a contact side (body, bv id) is a composite part that is not a box
// (Contact.csb1 / csb2 set and not RigidBodyGeomBox); AM_BV_SUB carries the part.
AMD_INLINE bool non_box_part(const BodiesView& B, int body, int bv) {
    const am_body_desc& d = B.desc[body];
    if (d.geom != AM_GEOM_COMPOSITE || bv > -3) return false;
    const int s = (-3 - bv) >> 16;
    return B.sub_desc[d.tree_first + s].geom != AM_GEOM_BOX;
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
        // box or composite on both sides takes the box-box path (CollisionProcessor.java:489-493)
        const int g1 = B.desc[b1].geom, g2 = B.desc[b2].geom;
        const bool boxish = (g1 == AM_GEOM_BOX || g1 == AM_GEOM_COMPOSITE) && (g2 == AM_GEOM_BOX || g2 == AM_GEOM_COMPOSITE);
        CKey key = make_key(b1, b2, C.bv1[c], C.bv2[c], C.info[c]);
        auto vanilla = [&]() {
            int o = old_lookup(O, key);
            if (o >= 0) take_old(C, c, O, o, post, false);
            else C.is_new[c] = 1;
        };
        // a part that is not a box: body 1's takes the vanilla start only
        // (:497-500), body 2's the vanilla start and then the box-box path (:501-503)
        if (!boxish || non_box_part(B, b1, C.bv1[c])) {
            vanilla();
            continue;
        }
        if (non_box_part(B, b2, C.bv2[c])) vanilla();
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

// RigidBody.applyForceW (RigidBody.java:335-340)
AMD_INLINE void apply_force_w(BodiesView B, int e, d3 p, d3 F) {
    st3(&B.f[e * 3], add(ld3(&B.f[e * 3]), F));
    const d3 r = sub(p, ld3(&B.x[e * 3]));
    st3(&B.t[e * 3], add(ld3(&B.t[e * 3]), cross(r, F)));
}
// RigidBodySystem.applySpringForces (RigidBodySystem.java:309-322) with
// Spring.applyOneBody / applyTwoBodies (Spring.java:165-214).  One thread walks
// the springs in list order, so every accumulator receives the reference's
// sequence of additions.
__global__ void k_springs(BodiesView B, const am_spring* S, int n, double ks, double ds) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i < n; i++) {
        const am_spring sp = S[i];
        const int b1 = sp.body1;
        const d3 p1w = xform(ldm(&B.R[b1 * 9]), ld3(&B.x[b1 * 3]), ld3(sp.pb1));
        if (sp.type != AM_SPRING_BODYBODY) {
            const d3 disp = sub(ld3(sp.pw), p1w);
            const double len = length(disp);
            if (len < 1e-3) continue;
            const d3 vel = spatial_velocity(B, b1, p1w);
            const double scl = -(sp.k * ks * (len - sp.l0 * sp.ls) - sp.d * ds * (dot(vel, disp) / len)) / len;
            const d3 F = scale(-scl, disp);
            apply_force_w(B, b1, p1w, F);
            if (B.parent[b1] >= 0) apply_force_w(B, B.parent[b1], p1w, F);
        } else {
            const int b2 = sp.body2;
            const d3 p2w = xform(ldm(&B.R[b2 * 9]), ld3(&B.x[b2 * 3]), ld3(sp.pb2));
            d3 disp = sub(p2w, p1w);
            const double dist = length(disp);
            if (dist < 1e-3) continue;
            disp = scale(1. / dist, disp);
            const d3 v2 = spatial_velocity(B, b2, p2w);
            const d3 v1 = spatial_velocity(B, b1, p1w);
            const d3 rel = sub(v2, v1);
            const double fi = sp.k * ks * (dist - sp.l0 * sp.ls) + sp.d * ds * dot(rel, disp);
            d3 F = scale(fi, disp);
            apply_force_w(B, b1, p1w, F);
            if (B.parent[b1] >= 0) apply_force_w(B, B.parent[b1], p1w, F);
            F = scale(-fi, disp);
            apply_force_w(B, b2, p2w, F);
            if (B.parent[b2] >= 0) apply_force_w(B, B.parent[b2], p2w, F);
        }
    }
}
void launch_springs(hipStream_t s, BodiesView B, const am_spring* springs, int n, double ks, double ds) {
    if (n > 0) hipLaunchKernelGGL(k_springs, dim3(1), dim3(64), 0, s, B, springs, n, ks, ds);
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

// Sine and cosine of the rotation angle (expRodrigues calls Math.sin and
// Math.cos).  Computed by fdlibm 5.3's algorithm (Sun Microsystems, 1993),
// which defines java.lang.StrictMath.sin/cos: Cody-Waite reduction by pi/2 in
// up to three steps (e_rem_pio2.c, |x| <= 2^19 pi/2) and the minimax kernels
// of k_sin.c / k_cos.c.  Larger arguments, which no physical step reaches,
// take the three-step reduction with a floating-point quotient.  Written out
// here rather than calling the device library so that every rounding is
// fixed by the source, whatever the library version.
namespace fdm {
__device__ __forceinline__ int hi(double x) { return (int)(__double_as_longlong(x) >> 32); }
__device__ __forceinline__ double from_hi(int h) { return __longlong_as_double((long long)(unsigned)h << 32); }
constexpr double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
constexpr double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
constexpr double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079632673412561417e+00,
                 pio2_1t = 6.07710050650619224932e-11, pio2_2 = 6.07710050630396597660e-11,
                 pio2_2t = 2.02226624879595063154e-21, pio2_3 = 2.02226624871116645580e-21,
                 pio2_3t = 8.47842766036889956997e-32;
// high words of the doubles nearest n pi/2, n = 1..32
__constant__ int npio2_hw[32] = {
    0x3ff921fb, 0x400921fb, 0x4012d97c, 0x401921fb, 0x401f6a7a, 0x4022d97c, 0x4025fdbb, 0x402921fb,
    0x402c463a, 0x402f6a7a, 0x4031475c, 0x4032d97c, 0x40346b9c, 0x4035fdbb, 0x40378fdb, 0x403921fb,
    0x403ab41b, 0x403c463a, 0x403dd85a, 0x403f6a7a, 0x40407e4c, 0x4041475c, 0x4042106c, 0x4042d97c,
    0x4043a28c, 0x40446b9c, 0x404534ac, 0x4045fdbb, 0x4046c6cb, 0x40478fdb, 0x404858eb, 0x404921fb};

__device__ double k_sin(double x, double y, bool tail) {
    int ix = hi(x) & 0x7fffffff;
    if (ix < 0x3e400000 && (int)x == 0) return x;
    double z = x * x, v = z * x;
    double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    if (!tail) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}
__device__ double k_cos(double x, double y) {
    int ix = hi(x) & 0x7fffffff;
    if (ix < 0x3e400000 && (int)x == 0) return 1.0;
    double z = x * x;
    double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    if (ix < 0x3fd33333) return 1.0 - (0.5 * z - (z * r - x * y));
    double qx = ix > 0x3fe90000 ? 0.28125 : from_hi(ix - 0x00200000);
    double hz = 0.5 * z - qx, a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}
// x = n pi/2 + y0 + y1 for |x| >= pi/4; returns n
__device__ int rem_pio2(double x, double& y0, double& y1) {
    int hx = hi(x), ix = hx & 0x7fffffff;
    if (ix < 0x4002d97c) {  // |x| < 3pi/4
        const double s = hx > 0 ? 1.0 : -1.0;
        double z = x - s * pio2_1;
        if (ix != 0x3ff921fb) {
            y0 = z - s * pio2_1t;
            y1 = (z - y0) - s * pio2_1t;
        } else {
            z = z - s * pio2_2;
            y0 = z - s * pio2_2t;
            y1 = (z - y0) - s * pio2_2t;
        }
        return hx > 0 ? 1 : -1;
    }
    double t = fabs(x);
    const bool medium = ix <= 0x413921fb;
    double fn;
    int n;
    if (medium) {
        n = (int)(t * invpio2 + 0.5);
        fn = (double)n;
    } else {
        fn = floor(t * invpio2 + 0.5);
        n = (int)(fn - 4.0 * floor(fn * 0.25));
    }
    double r = t - fn * pio2_1;
    double w = fn * pio2_1t;
    y0 = r - w;
    if (!(medium && n < 32 && ix != npio2_hw[n - 1])) {
        const int j = ix >> 20;
        int i = j - ((hi(y0) >> 20) & 0x7ff);
        if (i > 16 || !medium) {
            t = r;
            w = fn * pio2_2;
            r = t - w;
            w = fn * pio2_2t - ((t - r) - w);
            y0 = r - w;
            i = j - ((hi(y0) >> 20) & 0x7ff);
            if (i > 49 || !medium) {
                t = r;
                w = fn * pio2_3;
                r = t - w;
                w = fn * pio2_3t - ((t - r) - w);
                y0 = r - w;
            }
        }
    }
    y1 = (r - y0) - w;
    if (hx < 0) {
        y0 = -y0;
        y1 = -y1;
        return -n;
    }
    return n;
}
// both at once: the reduction is shared
__device__ void sincos(double x, double& s, double& c) {
    int ix = hi(x) & 0x7fffffff;
    if (ix <= 0x3fe921fb) {
        s = k_sin(x, 0.0, false);
        c = k_cos(x, 0.0);
        return;
    }
    if (ix >= 0x7ff00000) {
        s = c = x - x;
        return;
    }
    double y0, y1;
    const int q = rem_pio2(x, y0, y1) & 3;
    const double ks = k_sin(y0, y1, true), kc = k_cos(y0, y1);
    switch (q) {
        case 0: s = ks; c = kc; break;
        case 1: s = kc; c = -ks; break;
        case 2: s = -ks; c = -kc; break;
        default: s = -kc; c = ks; break;
    }
}
}  // namespace fdm

// Prescribed motion of the kinematic bodies at time t:
// x = x0 + (a sin(w t + phase)) dir, v = (a w cos(w t + phase)) dir, omega = 0.
__global__ void k_kinematic(BodiesView B, const am_kinematic* K, int n, double t) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const am_kinematic q = K[k];
    double s, c;
    fdm::sincos(q.omega * t + q.phase, s, c);
    const double ds = q.amplitude * s;
    const double vs = q.amplitude * q.omega * c;
    const int e = q.body;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        B.x[e * 3 + a] = q.x0[a] + ds * q.dir[a];
        B.v[e * 3 + a] = vs * q.dir[a];
        B.w[e * 3 + a] = 0.0;
    }
}
void launch_kinematic(hipStream_t s, BodiesView B, const am_kinematic* k, int n, double t) {
    if (n > 0) hipLaunchKernelGGL(k_kinematic, dim3((n + 63) / 64), dim3(64), 0, s, B, k, n, t);
}

// RigidBodySystem.applyImpulse (RigidBodySystem.java:262-267): each held
// force goes to its body's top-level entity through RigidBody.applyForceW.
// One thread, so forces on one entity add in the order they were held.
__global__ void k_impulses(BodiesView B, const ImpulseRec* I, int n) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int k = 0; k < n; k++) {
        const int b = I[k].body;
        const int e = B.parent[b] >= 0 ? B.parent[b] : b;
        apply_force_w(B, e, ld3(I[k].p), ld3(I[k].f));
    }
}
void launch_impulses(hipStream_t s, BodiesView B, const ImpulseRec* I, int n) {
    if (n > 0) hipLaunchKernelGGL(k_impulses, dim3(1), dim3(64), 0, s, B, I, n);
}

// RigidBodySystem.applyViscousDecay (RigidBodySystem.java:428-435): v and
// omega of the listed (top-level) entities scaled (Vector3d.scale).
__global__ void k_viscous(BodiesView B, const int* targets, int n, double a1, double a2) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int e = targets[k];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        B.v[e * 3 + a] = a1 * B.v[e * 3 + a];
        B.w[e * 3 + a] = a2 * B.w[e * 3 + a];
    }
}
void launch_viscous(hipStream_t s, BodiesView B, const int* targets, int n, double a1, double a2) {
    if (n > 0) hipLaunchKernelGGL(k_viscous, dim3((n + 255) / 256), dim3(256), 0, s, B, targets, n, a1, a2);
}

// The pose update of RigidBody.advancePositions (RigidBody.java:427-437):
// Rodrigues rotation (expRodrigues, RigidBody.java:382-401) and
// Matrix3d.normalizeCP.  Returns whether R changed.
__device__ __forceinline__ bool advance_pose(d3& x, m33& R, d3 v, d3 om, double dt) {
    x = scale_add(dt, v, x);
    double t = length(om) * dt;
    d3 d = normalized(om);
    if (t > 1e-8) {
        double s, c;
        fdm::sincos(t, s, c);
        const double c1 = 1 - c;
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
        return true;
    }
    return false;
}

// RigidBody.advancePositions (RigidBody.java:427-441) of the listed entities;
// with from_dv the velocities are deltaV's parts
// (advancePositionsPostStabilization, RigidBody.java:423-425).
__global__ void k_advance_positions(BodiesView B, const int* targets, int n, double dt, int from_dv) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = targets[k];
    d3 x = ld3(&B.x[i * 3]);
    m33 R = ldm(&B.R[i * 9]);
    const d3 v = from_dv ? ld3(&B.dv[i * 6]) : ld3(&B.v[i * 3]);
    const d3 om = from_dv ? ld3(&B.dv[i * 6 + 3]) : ld3(&B.w[i * 3]);
    const bool turned = advance_pose(x, R, v, om, dt);
    st3(&B.x[i * 3], x);
    if (turned) stm(&B.R[i * 9], R);
    stm(&B.jinv[i * 9], rm0rt(R, ldm(B.desc[i].jinv0)));
}

// Every entity's pose advanced by dt without changing the body state (the
// copies body1Advanced / body2Advanced of MotionMetricProcessor.java:79-83).
__global__ void k_advanced_poses(BodiesView B, double dt, double* xa, double* Ra) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    d3 x = ld3(&B.x[i * 3]);
    m33 R = ldm(&B.R[i * 9]);
    advance_pose(x, R, ld3(&B.v[i * 3]), ld3(&B.w[i * 3]), dt);
    st3(&xa[i * 3], x);
    stm(&Ra[i * 9], R);
}
void launch_advance_post(hipStream_t s, BodiesView B, const int* targets, int n, double dt) {
    if (n > 0) hipLaunchKernelGGL(k_advance_positions, dim3((n + 127) / 128), dim3(128), 0, s, B, targets, n, dt, 1);
}
void launch_advanced_poses(hipStream_t s, BodiesView B, double dt, double* xa, double* Ra) {
    if (B.n > 0) hipLaunchKernelGGL(k_advanced_poses, dim3((B.n + 127) / 128), dim3(128), 0, s, B, dt, xa, Ra);
}

void launch_advance(hipStream_t s, BodiesView B, const int* targets, int n, double dt, bool positions) {
    if (n <= 0) return;
    if (positions) hipLaunchKernelGGL(k_advance_positions, dim3((n + 127) / 128), dim3(128), 0, s, B, targets, n, dt, 0);
    else hipLaunchKernelGGL(k_advance_velocities, dim3((n + 255) / 256), dim3(256), 0, s, B, targets, n, dt);
}

// jinv = R jinv0 R^T for non-pinned entities, zero for pinned.
__global__ void k_update_jinv(BodiesView B) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    if (B.flags[i] & AM_BODY_PINNED) {
#pragma unroll
        for (int a = 0; a < 9; a++) B.jinv[i * 9 + a] = 0.0;
        return;
    }
    stm(&B.jinv[i * 9], rm0rt(ldm(&B.R[i * 9]), ldm(B.desc[i].jinv0)));
}
void launch_update_jinv(hipStream_t s, BodiesView B) {
    if (B.n > 0) hipLaunchKernelGGL(k_update_jinv, dim3((B.n + 255) / 256), dim3(256), 0, s, B);
}

// ---------------------------------------------------------------------------
// MotionMetricProcessor.getLargestVelocityNorm (MotionMetricProcessor.java:39-73)
__global__ void k_metric(BodiesView B, const int* pairs, int n, double* out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int a = pairs[k * 2], b = pairs[k * 2 + 1];
    out[k] = b < 0 ? metric_one(B, a) : metric_two(B, a, b);
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
                         int* col_out, int* remaining, const int* layer, int* blocker) {
    if (n_groups <= 0) return;
    hipLaunchKernelGGL(k_colour_round, dim3((n_groups + 255) / 256), dim3(256), 0, s, n_groups, ga, gb, bflags,
                       inc_start, inc_count, inc_items, col_in, col_out, remaining, layer, blocker);
}
void launch_colour_keys(hipStream_t s, int n_groups, const int* colour, const int* group_start, uint64_t* keys) {
    if (n_groups > 0)
        hipLaunchKernelGGL(k_colour_keys, dim3((n_groups + 255) / 256), dim3(256), 0, s, n_groups, colour,
                           group_start, keys);
}
void launch_colour_layout(hipStream_t s, int n_groups, const uint64_t* sorted, const int* group_start,
                          const int* group_a, const int* group_b, int* colour_start, int* sizes, int* pos_start,
                          int* order, int* group_desc, int* pos_a, int* spos, int* d_total, void* scan_temp) {
    hipLaunchKernelGGL(k_colour_starts, dim3(1), dim3(128), 0, s, n_groups, sorted, colour_start);
    if (n_groups <= 0) return;
    dim3 g((n_groups + 255) / 256), b(256);
    hipLaunchKernelGGL(k_sorted_sizes, dim3((n_groups + 256) / 256), b, 0, s, n_groups, sorted, group_start, sizes);
    exclusive_scan(s, sizes, pos_start, n_groups + 1, d_total, scan_temp);
    hipLaunchKernelGGL(k_solve_order, g, b, 0, s, n_groups, sorted, group_start, pos_start, group_a, group_b, order,
                       group_desc, pos_a, spos);
}

}  // namespace am
