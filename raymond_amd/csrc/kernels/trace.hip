// Traversal kernels for gfx950: closest hit (optionally fused with camera-ray
// generation) and any hit (fused with shadow-ray accumulation).
//
// Counterparts: generateRays (device/kernels/shading/generate.metal:72-192),
// raytrace / raytraceAny (device/kernels/shading/raytrace.metal:11-85) and
// handleShadowRays (device/kernels/shading/shadow.metal:6-29).
//
// Design:
//  * persistent waves over the device-resident queue length (no host
//    round-trip): most 64-ray batches are assigned statically, the tail is
//    claimed dynamically through a per-depth counter (BatchScheduler);
//  * compressed BVH8 (80-byte nodes, five 16-byte loads, 8-bit child boxes)
//    with ordered internal-child visits by ray octant;
//  * traversal stack: the bottom kLdsStack entries per lane in LDS (column
//    per lane), deeper entries spill to a per-thread global array.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "../device/common.h"
#include "../device/rng.h"
#include "../device/isect.h"
#include "../device/bvh.h"
#include "../device/queues.h"
#include "../device/camera.h"
#include "../device/scene.h"
#include "../device/rcache.h"
#include "../device/batch.h"
#include "trace_api.h"

namespace rm {

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
// bytes 2k, 2k+1 of w as the halves 1024 + byte (exact)
__device__ inline half2v byte_pair(uint32_t w, int k) {
    return __builtin_bit_cast(half2v, __builtin_amdgcn_perm(0x64646464u, w, k ? 0x04030402u : 0x04010400u));
}
__device__ inline float plane_q(const half2v (&h)[4], int c) { return (float)h[c >> 1][c & 1]; }
__device__ inline uint32_t byte_of(uint32_t lo, uint32_t hi, int c) {
    const uint32_t w = c < 4 ? lo : hi;
    return (w >> (8 * (c & 3))) & 0xffu;
}

// Stack entries as 64-bit words behind pointers typed by address space, so
// that the LDS rows and the spill rows stay ds_* and global_* accesses (a
// generic pointer chosen between the two would make every stack access a
// flat access).
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
typedef __attribute__((address_space(1))) unsigned long long global_u64;
typedef __attribute__((address_space(3))) float lds_f32;

__device__ inline unsigned long long pack_entry(uint2 v) { return ((unsigned long long)v.y << 32) | v.x; }
__device__ inline uint2 unpack_entry(unsigned long long w) { return make_uint2((uint32_t)w, (uint32_t)(w >> 32)); }

struct StackRef {
    lds_u64* lds;       // column base: lds[sp * 64]
    global_u64* spill;  // spill[(sp - kLdsStack) * stride]
    uint32_t stride;
    // Two-level trees: the world ray's origin and direction, world[k * 64].
    // It lives in LDS beside the stack rows rather than in six registers
    // that every step would carry, since only entering and leaving an
    // instance read it: the closest-hit kernel then fits 5 waves per SIMD
    // without the 36 bytes per lane of spills it had, which measured 2346-2351
    // against 2285 Mrays/s on the 1080p instanced field.
    lds_f32* world;
    __device__ void put_world(vec3 o, vec3 d) const {
        world[0] = o.x;
        world[kTraceBlock] = o.y;
        world[2 * kTraceBlock] = o.z;
        world[3 * kTraceBlock] = d.x;
        world[4 * kTraceBlock] = d.y;
        world[5 * kTraceBlock] = d.z;
    }
    __device__ vec3 world_o() const { return vec3(world[0], world[kTraceBlock], world[2 * kTraceBlock]); }
    __device__ vec3 world_d() const {
        return vec3(world[3 * kTraceBlock], world[4 * kTraceBlock], world[5 * kTraceBlock]);
    }
    __device__ void store(int sp, uint2 v) const {
        if (sp < kLdsStack) lds[sp * kTraceBlock] = pack_entry(v);
        else spill[(size_t)(sp - kLdsStack) * stride] = pack_entry(v);
    }
    __device__ uint2 load(int sp) const {
        if (sp < kLdsStack) return unpack_entry(lds[sp * kTraceBlock]);
        // a load of another kind, which the compiler cannot merge with the
        // LDS load into one flat load
        return unpack_entry(__builtin_nontemporal_load(&spill[(size_t)(sp - kLdsStack) * stride]));
    }
};

// Postponed leaf sets a lane may hold while it takes nodes (1 or 2), and the
// share of a wave's active lanes held up by their triangles (1/kHeldShare)
// that starts a triangle phase.  On the 5.2M-triangle sphere field
// (Mrays/s, against 2944-2976 for one slot at a quarter): two slots at 1/4
// 3035, 1/5 3049, 1/6 3037-3071, 1/8 3051, 1/12 3018, 1/16 2996; at 1/3
// 2935 and 1/2 2773.  Two slots visit more nodes per ray (16.1 against
// 15.3: closer hits arrive later) but keep more lanes busy.  Three slots
// measured 2875 (at 1/6 and 1/8): 16.9 nodes per ray, and the closest-hit
// kernel's 97 VGPRs fit 4 waves per SIMD.
#ifndef RM_LEAF_SLOTS
#define RM_LEAF_SLOTS 2
#endif
constexpr int kLeafSlots = RM_LEAF_SLOTS;
#ifndef RM_HELD_SHARE
#define RM_HELD_SHARE 6
#endif
constexpr int kHeldShare = RM_HELD_SHARE;
// The any-hit kernel, whose rays may end at their first hit, starts its
// triangle phases sooner: at 1/3 with 3 triangles per phase, its time on
// the sphere field measured 0.341-0.343 ms per spp against 0.353-0.354 at
// the closest-hit kernel's 1/6 and 2 (1/4 and 3: 0.342-0.344; 1/4 and 4:
// 0.342-0.344, slower on the instanced field; 1/10 and 2: 0.363).
#ifndef RM_HELD_SHARE_ANY
#define RM_HELD_SHARE_ANY 3
#endif
constexpr int kHeldShareAny = RM_HELD_SHARE_ANY;

// Traversal state of one lane's ray.  It persists across steps, so a lane
// whose ray has finished takes a new one while the rest of the wave goes on.
struct RayState {
    vec3 o, inv;
    RayShear sh;
    int oct;
    float tmin, tmax, best_t;
    uint32_t best_inst, best_prim;
    float bu, bv, bdet;  // closest hit's scaled barycentrics and their sum
    bool found;
    uint2 cur;  // node group: child base, (remaining child hits << 24) | internal mask
    uint2 tg;   // postponed leaf triangles: base, mask (top level of a two-level tree: instances)
    uint2 tg2;  // a second postponed leaf set (kLeafSlots == 2)
    int sp;
    uint32_t n_nodes, n_tris;
    // two-level trees: the instance entered, in the top level
    uint32_t inst;
    bool top;
    int marker_sp;  // stack index of the instance's marker entry (-1 in the top level)

    __device__ void set_ray(vec3 org, vec3 d) {
        o = org;
        inv = safe_inv_dir(d);
        sh = make_shear(d);
        oct = dir_octant(d);
    }

    __device__ void init(uint32_t root, vec3 org, vec3 d, float t0, float t1) {
        set_ray(org, d);
        inst = 0;
        top = true;
        marker_sp = -1;
        // a one-child pseudo node whose only child is the root
        cur = make_uint2(root, (1u << (24 + oct)) | 1u);
        sp = 0;
        tmin = t0;
        tmax = t1;
        best_t = t1;
        best_inst = 0xffffffffu;
        best_prim = 0xffffffffu;
        bu = 0.0f;
        bv = 0.0f;
        bdet = 1.0f;
        found = false;
        tg = make_uint2(0u, 0u);
        tg2 = make_uint2(0u, 0u);
        n_nodes = 0;
        n_tris = 0;
    }

    // Two-level trees: enters the next instance of the top-level leaf set in
    // tg.  The top-level group (if it has hits left) and a marker entry with
    // the set's other instances go on the stack; popping the marker returns
    // to the world ray.  The ray is mapped affinely into the instance's
    // object space without renormalising the direction, so t is the same
    // parameter on both sides, and the shape's tree is visited from its root.
    __device__ void enter(const TlasInstance* __restrict__ insts, const StackRef& stk) {
        const int b = __builtin_ctz(tg.y);
        tg.y &= tg.y - 1u;
        const uint4* ip = reinterpret_cast<const uint4*>(insts + tg.x + (uint32_t)b);
        const uint4 r0 = ip[0], r1 = ip[1], r2 = ip[2], r3 = ip[3];
        if (cur.y & 0xff000000u) {
            stk.store(sp, cur);
            sp++;
        }
        stk.store(sp, make_uint2(tg.x, kMarker2l | tg.y));
        marker_sp = sp;
        sp++;
        TlasInstance ti;
        const uint4 rows[3] = {r0, r1, r2};
        for (int j = 0; j < 3; j++) {
            ti.m[4 * j + 0] = __uint_as_float(rows[j].x);
            ti.m[4 * j + 1] = __uint_as_float(rows[j].y);
            ti.m[4 * j + 2] = __uint_as_float(rows[j].z);
            ti.m[4 * j + 3] = __uint_as_float(rows[j].w);
        }
        set_ray(tlas_point(ti, stk.world_o()), tlas_vector(ti, stk.world_d()));
        inst = r3.y;
        top = false;
        cur = make_uint2(r3.x, (1u << (24 + oct)) | 1u);
        tg = make_uint2(0u, 0u);
    }

    // One node (after a stack pop when the current group is used up, so
    // that no step is a bare pop), or the postponed triangles of the leaves
    // hit so far.  Returns true when the traversal has ended (ANY: also at
    // the first hit in (tmin, tmax)).
    template <bool ANY, bool STATS, bool TWO = false>
    __device__ bool step(const Bvh8Node* __restrict__ nodes, const BvhTri* __restrict__ tris,
                         const TlasInstance* __restrict__ insts, const StackRef& stk) {
        // Leaf triangles wait until 1/kHeldShare of the wave's active lanes
        // are held up by them (or no lane can take a node), then those lanes
        // test a few each together instead of each node step running a loop
        // most lanes sit out.  With a second leaf slot, a lane with one
        // pending leaf set goes on taking nodes until a second one arrives.
        // In a two-level tree a top-level leaf set holds instances: a lane
        // with one is held like a lane with triangles and enters the next
        // instance in the same phase as the others test triangles, so the
        // entry code runs for many lanes at once rather than in most steps
        // for a few.  A lane tests its pending triangles before it pops its
        // instance's marker: they belong to the instance's ray.
        constexpr bool kTwoSlots = kLeafSlots > 1;
        const bool to_enter = TWO && top && tg.y != 0u;
        const bool has_tris = !to_enter && tg.y != 0u;
        const bool node_work = (cur.y & 0xff000000u) || sp > 0;
        bool can_node = !to_enter && node_work && (kTwoSlots ? tg2.y == 0u : !has_tris);
        if (TWO && kTwoSlots && has_tris && !(cur.y & 0xff000000u) && sp - 1 == marker_sp) can_node = false;
        const unsigned long long held = __ballot((has_tris || to_enter) && !can_node);
        if (__popcll(held) * (ANY ? kHeldShareAny : kHeldShare) >= __popcll(__ballot(true)) ||
            __ballot(can_node) == 0ull) {
            if (to_enter) {
                enter(insts, stk);
                return false;
            }
            if (!has_tris) return !node_work;
            const bool ended = test_tris<ANY, STATS, TWO>(tris);
            if (kTwoSlots && tg.y == 0u) {
                tg = tg2;
                tg2 = make_uint2(0u, 0u);
            }
            return ended || (!node_work && tg.y == 0u);
        }
        if (!can_node) return false;
        if (!(cur.y & 0xff000000u)) {
            if (sp == 0) return true;
            sp--;
            const uint2 e = stk.load(sp);
            if (TWO && (e.y & kMarker2l)) {
                // back in the world: the set's other instances, then the
                // top-level groups below the marker (entering the next
                // instance maps the ray again, so the world ray is needed only
                // after the set's last one)
                tg = make_uint2(e.x, e.y & ~kMarker2l);
                if (tg.y == 0u) set_ray(stk.world_o(), stk.world_d());
                top = true;
                marker_sp = -1;
                cur = make_uint2(0u, 0u);
                return false;
            }
            cur = e;  // pushed groups always hold a hit child
        }
        const uint32_t hits = cur.y >> 24;
        const int k = __builtin_ctz(hits);
        cur.y &= ~(1u << (24 + k));
        const int slot = k ^ oct;
        const uint32_t node_idx = cur.x + __builtin_popcount(cur.y & 0xffu & ((1u << slot) - 1u));
        if (cur.y & 0xff000000u) {
            stk.store(sp, cur);
            sp++;
        }
        const uint4* np = reinterpret_cast<const uint4*>(nodes + node_idx);
        const uint4 q0 = np[0], q1 = np[1], q2 = np[2], q3 = np[3], q4 = np[4];
        if (STATS) n_nodes++;
        const float px = __uint_as_float(q0.x), py = __uint_as_float(q0.y), pz = __uint_as_float(q0.z);
        const uint32_t imask = q0.w >> 24;
        // Slab planes in ray space: t(q) = q * A + B with A = 2^(e-127) * inv
        // (exact) and B = (p - o) * inv, one FMA per plane.  The byte q
        // enters as the half 1024 + q (two per v_perm, exact), which
        // v_fma_mix_f32 multiplies in f32, so the offset becomes
        // B' = (B -+ E) - 1024 A.  With u = 2^-24 the computed plane is within
        // u (1534 |A| + 6 |B|) of its exact t, (q s + p - o) / d, to first
        // order, so moving B outward by E = 16u (|B| + 255 |A|) leaves every
        // near plane before and every far plane after the exact one, with a
        // margin of nearly 10u (|B| + 255 |A|) >= 10u |t| for the triangle
        // test's own rounding (best_t keeps the relative slack on top).  A
        // ray parallel to an axis has inv = +-1e30: the planes land at +-huge
        // t and cull exactly when the origin is outside the slab; where B
        // overflows, the NaN planes fall out of the max/min below (each drop
        // widens [tn, tf]).
        auto axis = [](float p, uint32_t ebits, float o, float inv, float& A, float& Bn, float& Bf) {
            A = exp_scale((uint8_t)ebits) * inv;
            const float B = (p - o) * inv;
            const float E = 9.53674316e-7f * (fabsf(B) + 255.0f * fabsf(A));
            Bn = __builtin_fmaf(-1024.0f, A, B - E);
            Bf = __builtin_fmaf(-1024.0f, A, B + E);
        };
        float ax, bnx, bfx, ay, bny, bfy, az, bnz, bfz;
        axis(px, q0.w & 0xffu, o.x, inv.x, ax, bnx, bfx);
        axis(py, (q0.w >> 8) & 0xffu, o.y, inv.y, ay, bny, bfy);
        axis(pz, (q0.w >> 16) & 0xffu, o.z, inv.z, az, bnz, bfz);
        // near and far planes per axis for the lane's direction, chosen once
        // per node (t decreases in q when inv < 0, so the high bytes are near)
        const bool fx = inv.x < 0.0f, fy = inv.y < 0.0f, fz = inv.z < 0.0f;
        const uint32_t nx0 = fx ? q3.z : q2.x, nx1 = fx ? q3.w : q2.y, xx0 = fx ? q2.x : q3.z, xx1 = fx ? q2.y : q3.w;
        const uint32_t ny0 = fy ? q4.x : q2.z, ny1 = fy ? q4.y : q2.w, xy0 = fy ? q2.z : q4.x, xy1 = fy ? q2.w : q4.y;
        const uint32_t nz0 = fz ? q4.z : q3.x, nz1 = fz ? q4.w : q3.y, xz0 = fz ? q3.x : q4.z, xz1 = fz ? q3.y : q4.w;
        const float t_far = best_t * kBoxSlack;
        const half2v hnx[4] = {byte_pair(nx0, 0), byte_pair(nx0, 1), byte_pair(nx1, 0), byte_pair(nx1, 1)};
        const half2v hxx[4] = {byte_pair(xx0, 0), byte_pair(xx0, 1), byte_pair(xx1, 0), byte_pair(xx1, 1)};
        const half2v hny[4] = {byte_pair(ny0, 0), byte_pair(ny0, 1), byte_pair(ny1, 0), byte_pair(ny1, 1)};
        const half2v hxy[4] = {byte_pair(xy0, 0), byte_pair(xy0, 1), byte_pair(xy1, 0), byte_pair(xy1, 1)};
        const half2v hnz[4] = {byte_pair(nz0, 0), byte_pair(nz0, 1), byte_pair(nz1, 0), byte_pair(nz1, 1)};
        const half2v hxz[4] = {byte_pair(xz0, 0), byte_pair(xz0, 1), byte_pair(xz1, 0), byte_pair(xz1, 1)};
        // Branch-free child loop.  A hit child ORs (meta >> 5) << (meta & 31)
        // into acc: a leaf's triangle span lands in bits 0-23, an internal
        // child in slot c sets bit 24 + c.  fmaxf/fminf (v_max3/v_min3) drop
        // NaN operands, which only widens [tn, tf].
        uint32_t acc = 0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const uint32_t meta = byte_of(q1.z, q1.w, c);
            const float tnx = __builtin_fmaf(plane_q(hnx, c), ax, bnx);
            const float tfx = __builtin_fmaf(plane_q(hxx, c), ax, bfx);
            const float tny = __builtin_fmaf(plane_q(hny, c), ay, bny);
            const float tfy = __builtin_fmaf(plane_q(hxy, c), ay, bfy);
            const float tnz = __builtin_fmaf(plane_q(hnz, c), az, bnz);
            const float tfz = __builtin_fmaf(plane_q(hxz, c), az, bfz);
            const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
            const float tf = fminf(fminf(tfx, tfy), fminf(tfz, t_far));
            const bool hit = tn <= tf && meta != 0u;
            acc |= hit ? (meta >> 5) << (meta & 31u) : 0u;
        }
        const uint32_t thits = acc & 0xffffffu;
        // internal hits from slot order (bit c) to visit order (bit c ^ oct)
        uint32_t ihits = acc >> 24;
        ihits = (oct & 1) ? ((ihits & 0x55u) << 1) | ((ihits >> 1) & 0x55u) : ihits;
        ihits = (oct & 2) ? ((ihits & 0x33u) << 2) | ((ihits >> 2) & 0x33u) : ihits;
        ihits = (oct & 4) ? ((ihits & 0x0fu) << 4) | (ihits >> 4) : ihits;
        // The hit children are visited in octant order, the group kept as one
        // stack entry.  Choosing the nearest child first visited fewer nodes
        // (14.1 against 15.3 per ray on the 5.2M-triangle sphere field) but
        // cost more per node step than it saved (2921 against 2948 Mrays/s).
        cur = make_uint2(q1.x, (ihits << 24) | imask);
        if (kTwoSlots && tg.y != 0u) tg2 = make_uint2(q1.y, thits);
        else tg = make_uint2(q1.y, thits);
        return false;
    }

    // Tests up to kTrisPerPhase postponed triangles; true when an any-hit
    // query found a hit.  Bounding the phase frees a lane with few triangles
    // left for node steps while others still have some.  On the 5.2M-triangle
    // sphere field (Mrays/s), one leaf slot: 1 per phase 2841, 2 per phase
    // 2909, 3 2891, 4 2870, emptying every lane's list per phase slower than
    // 1; two slots: 1 2899, 2 3033-3071, 3 3052.  Three per phase with the
    // phase started at 1/4, 1/5 or 1/8 of the lanes held measured 3050-3084
    // against 3104-3109 for two at 1/6 (the same box).
#ifndef RM_TRIS_PER_PHASE
#define RM_TRIS_PER_PHASE 2
#endif
    static constexpr int kTrisPerPhase = RM_TRIS_PER_PHASE;
// (any hit: 1 per phase 0.380 ms per spp against 0.354 for 2, 3 0.349-0.350)
#ifndef RM_TRIS_PER_PHASE_ANY
#define RM_TRIS_PER_PHASE_ANY 3
#endif
    static constexpr int kTrisPerPhaseAny = RM_TRIS_PER_PHASE_ANY;
    template <bool ANY, bool STATS, bool TWO>
    __device__ bool test_tris(const BvhTri* __restrict__ tris) {
        for (int k = 0; k < (ANY ? kTrisPerPhaseAny : kTrisPerPhase) && tg.y; k++) {
            const int b = __builtin_ctz(tg.y);
            tg.y &= tg.y - 1u;
            const uint4* tp = reinterpret_cast<const uint4*>(tris + tg.x + b);
            const uint4 a = tp[0], bq = tp[1], cq = tp[2];
            if (STATS) n_tris++;
            const vec3 v0(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
            const vec3 v1(__uint_as_float(bq.x), __uint_as_float(bq.y), __uint_as_float(bq.z));
            const vec3 v2(__uint_as_float(cq.x), __uint_as_float(cq.y), __uint_as_float(cq.z));
            float t, u, v, det;
            // a two-level tree's triangles are the shape's: the instance is the one entered
            const uint32_t ti = TWO ? inst : a.w;
            if (intersect_triangle_uvd(o, sh, v0, v1, v2, tmin, tmax, t, u, v, det) &&
                hit_before(t, ti, bq.w, best_t, best_inst, best_prim)) {
                best_t = t;
                best_inst = ti;
                best_prim = bq.w;
                bu = u;
                bv = v;
                bdet = det;
                found = true;
                if (ANY) return true;
            }
        }
        return false;
    }
};

// Waves per SIMD the trace kernels ask the register allocator for.  Left
// free they take 95 and 92 VGPRs (5 waves); asking for 6 gives 80 VGPRs
// with 48 and 36 bytes of spills per lane, which measured 2311 against
// 3079-3097 Mrays/s on the 5.2M-triangle sphere field.
#ifndef RM_TRACE_MIN_WAVES
#define RM_TRACE_MIN_WAVES 1
#endif
constexpr int kTraceMinWaves = RM_TRACE_MIN_WAVES;
// The two-level kernels take 98 and 95 VGPRs left free (4 and 5 waves per
// SIMD); held to 5 waves both fit without spills.  With the world ray still
// in registers they took 103 and 101 free, and 5 waves with spills measured
// 2331 against 2295 Mrays/s for 4 waves on the 1080p instanced field.
#ifndef RM_TWO_LEVEL_MIN_WAVES
#define RM_TWO_LEVEL_MIN_WAVES 5
#endif
constexpr int kTwoLevelMinWaves = RM_TWO_LEVEL_MIN_WAVES;

// Lanes are refilled once at least this many of a wave's 64 are idle
// (with two leaf slots, 4, 12 and 16 measured within 1% of 8).
#ifndef RM_REFILL_LANES
#define RM_REFILL_LANES 8
#endif
constexpr uint32_t kRefillLanes = RM_REFILL_LANES;

// Batch scheduler of the persistent waves.  The first 3/4 of the 64-ray
// batches are assigned statically (wave w takes w, w + nwaves, ...); the rest
// are claimed one at a time with an atomic on `counter`.  Waves without a
// batch of their own exit without touching the counter: device-scope atomics
// on one address serialise across the XCDs, so a queue of a few thousand
// rays must not cost an atomic per launched wave.
struct BatchScheduler {
    uint32_t nb, static_nb, nwaves, b;
    bool dynamic;
    __device__ BatchScheduler(uint32_t count) {
        nb = (count + (uint32_t)kTraceBlock - 1u) / (uint32_t)kTraceBlock;
        static_nb = nb - nb / 4u;
        nwaves = gridDim.x;
        b = blockIdx.x;
        dynamic = false;
    }
    // next batch index, or >= nb when done
    __device__ uint32_t next(uint32_t* counter) {
        if (!dynamic && b >= static_nb) dynamic = true;
        if (dynamic) {
            uint32_t v = 0;
            if (threadIdx.x == 0) v = atomicAdd(counter, 1u);
            b = static_nb + __shfl(v, 0);
        }
        const uint32_t cur = b;
        if (!dynamic) b += nwaves;
        return cur;
    }
};

// Rays of a wave: batches of 64 consecutive queue entries come from the
// scheduler, and idle lanes take the batch's next rays in lane order.  A
// refill happens once kRefillLanes lanes are idle (or all are), so the wave
// pays the ray set-up for several lanes at a time.  Wave-uniform control.
struct RayPool {
    BatchScheduler& sched;
    uint32_t* fetch;
    uint32_t count;
    uint32_t batch, taken;
    bool exhausted;
    uint32_t ray;  // the ray a lane was given by the last refill

    __device__ RayPool(BatchScheduler& s, uint32_t* f, uint32_t n) : sched(s), fetch(f), count(n), taken(0) {
        batch = sched.next(fetch);
        exhausted = batch >= sched.nb;
        ray = 0;
    }
    // Returns true for the lanes that received a new ray.
    __device__ bool refill(bool active) {
        if (exhausted) return false;
        unsigned long long idle = __ballot(!active);
        if ((uint32_t)__popcll(idle) < kRefillLanes && ~idle != 0ull) return false;
        const unsigned long long lt = (1ull << __lane_id()) - 1ull;
        bool got = false;
        while (idle != 0ull && !exhausted) {
            const uint32_t first = batch * (uint32_t)kTraceBlock + taken;
            const uint32_t room = (uint32_t)kTraceBlock - taken;
            const uint32_t left = count > first ? count - first : 0u;
            const uint32_t avail = room < left ? room : left;
            if (avail == 0u) {
                batch = sched.next(fetch);
                taken = 0;
                exhausted = batch >= sched.nb;
                continue;
            }
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            const uint32_t give = n_idle < avail ? n_idle : avail;
            const bool is_idle = (idle >> __lane_id()) & 1ull;
            const uint32_t rank = (uint32_t)__popcll(idle & lt);
            if (is_idle && rank < give) {
                ray = first + rank;
                got = true;
            }
            taken += give;
            idle = __ballot(!active && !got);
        }
        return got;
    }
};

// Closest hit over a path queue.  Camera paths (depth 0) use the camera's
// clip range, continuation rays (eps, inf) (intersection.metal:201-202);
// raw batches bring their own ranges.
template <bool STATS, bool TWO>
__global__ __launch_bounds__(kTraceBlock, TWO ? kTwoLevelMinWaves : kTraceMinWaves) void trace_closest_kernel(ClosestArgs a) {
    __shared__ uint2 lds[kLdsStack * kTraceBlock];
    __shared__ float world[TWO ? 6 * kTraceBlock : 1];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * kTraceBlock + lane;
    StackRef stk{(lds_u64*)(lds + lane), (global_u64*)(a.spill + gtid), gridDim.x * kTraceBlock,
                 (lds_f32*)(world + (TWO ? lane : 0u))};
    const uint32_t* count_ptr = a.count_ptr;
    uint32_t* fetch = a.fetch;
    if (a.iter_dev) {
        IterCounters& ic = a.counters->it[iteration_of(a.iter_dev, a.iter)];
        count_ptr = &ic.queue;
        fetch = &ic.fetch_closest;
    }
    const uint32_t count = count_ptr ? *count_ptr : a.count;
    unsigned long long sn = 0, st = 0;
    BatchScheduler sched(count);
    if (blockIdx.x >= sched.nb) return;
    const PathQueue& q = a.paths;
    RayState s;
    bool active = false;
    uint32_t i = 0;
    RayPool pool(sched, fetch, count);
    while (true) {
        // idle lanes take the next rays of the wave's batch
        if (pool.refill(active)) {
            i = pool.ray;
            const vec3 o(q.ox[i], q.oy[i], q.oz[i]);
            const vec3 d(q.dx[i], q.dy[i], q.dz[i]);
            uint32_t root = TWO ? a.root2l : 0u;
            float near_clip = a.camera.near_clip, far_clip = a.camera.far_clip;
            if (a.batch.jobs) {
                const JobInfo& job = a.batch.jobs[job_of_pixel(a.batch.jobs, a.batch.num_jobs, q.pixel[i])];
                root = job.root;
                near_clip = job.camera.near_clip;
                far_clip = job.camera.far_clip;
            }
            float tmin, tmax;
            if (a.tmin_arr) {
                tmin = a.tmin_arr[i];
                tmax = a.tmax_arr[i];
            } else {
                const uint32_t depth = meta_depth(q.meta[i]);
                // a void path's empty range ends its traversal at the root
                tmin = depth == 0 ? near_clip : (depth == kVoidDepth ? 1.0f : kEps);
                tmax = depth == 0 ? far_clip : (depth == kVoidDepth ? 0.0f : kInf);
            }
            s.init(root, o, d, tmin, tmax);
            if (TWO) stk.put_world(o, d);
            active = true;
        }
        if (!__any(active)) break;
        if (active && s.step<false, STATS, TWO>(a.nodes, a.tris, a.tlas, stk)) {
            a.hits.t[i] = s.found ? s.best_t : -1.0f;
            a.hits.prim[i] = s.best_prim;
            a.hits.inst[i] = s.best_inst;
            a.hits.b0[i] = s.bu / s.bdet;
            a.hits.b1[i] = s.bv / s.bdet;
            if (STATS) {
                sn += s.n_nodes;
                st += s.n_tris;
            }
            active = false;
        }
    }
    if (STATS) {
        atomicAdd(&a.counters->nodes_closest, sn);
        atomicAdd(&a.counters->tris_closest, st);
    }
}

// Path regeneration: appends new camera paths after the survivors of the
// previous iteration until the queue is full or every sample of the stream
// has been started.  Path g of the stream is sample s = g / (W H) of the
// pixel with block-linear index g mod (W H), keyed exactly as the
// reference's raygen keys (frame seed, rayIndex) (generate.metal:88-121).
__global__ void regenerate_kernel(RegenArgs a) {
    const uint32_t iter = iteration_of(a.iter_dev, a.iter);
    const IterCounters& cur = a.counters->it[iter];
    const uint32_t appended = cur.appended;
    const uint32_t generated = cur.generated;
    const uint32_t room = a.capacity - appended;
    const uint32_t left = a.total - generated;
    const uint32_t k = room < left ? room : left;
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 == 0) {
        a.counters->it[iter].queue = appended + k;
        a.counters->it[iter + 1].generated = generated + k;
    }
    const uint32_t wh = a.width * a.height;
    for (uint32_t t = t0; t < k; t += gridDim.x * blockDim.x) {
        const uint32_t g = generated + t;
        uint32_t s, ray_index, width = a.width, height = a.height, seed = a.random_seed, image_offset = a.image_offset;
        Camera cam = a.camera;
        if (a.slices) {
            // batched jobs: the slice (job, sample) holding path g
            const BatchSlice sl = a.slices[slice_of_path(a.slices, a.num_slices, g)];
            const JobInfo& job = a.jobs[sl.job];
            s = sl.sample;
            ray_index = g - sl.first;
            width = job.width;
            height = job.height;
            seed = job.seed;
            image_offset = job.image_offset;
            cam = job.camera;
        } else {
            s = g / wh;
            ray_index = g - s * wh;
        }
        uint32_t x, y;
        block_linear_to_pixel(ray_index, width, height, x, y);
        Prng prng(seed + s, ray_index);
        vec3 o, d, w(1.0f);
        bool ray = true;
        if (a.lens.num_surfaces > 0) ray = lens_camera_ray(cam, a.lens, x, y, width, height, prng, o, d, w);
        else camera_ray(cam, x, y, width, height, prng, o, d);
        if (a.lens.num_surfaces > 0) {
            // one count per wave of the paths that carry no ray
            const unsigned long long v = __ballot(!ray);
            if (v && (uint32_t)__lane_id() == (uint32_t)__builtin_ctzll(__ballot(true)))
                atomicAdd(&a.counters->it[iter].voids, (uint32_t)__popcll(v));
            if (!ray) {
                o = vec3(0.0f);
                d = vec3(0.0f, 0.0f, 1.0f);
                w = vec3(0.0f);
            }
        }
        const uint32_t j = appended + t;
        PathQueue& q = a.paths;
        q.ox[j] = o.x; q.oy[j] = o.y; q.oz[j] = o.z;
        q.dx[j] = d.x; q.dy[j] = d.y; q.dz[j] = d.z;
        q.wx[j] = w.x; q.wy[j] = w.y; q.wz[j] = w.z;
        q.pdf[j] = kInf;
        q.seed[j] = prng.seed;
        q.meta[j] = pack_meta(prng.index, RayCamera, ray ? 0u : kVoidDepth);
        q.pixel[j] = image_offset + y * width + x;
        if (a.slots) {
            const uint32_t slot = rc_training_slot(g, a.random_seed, a.train_shift);
            a.slots[j] = slot;
            if (slot != kRcNoSlot) {
                a.slot_radiance[3 * (size_t)slot + 0] = 0.0f;
                a.slot_radiance[3 * (size_t)slot + 1] = 0.0f;
                a.slot_radiance[3 * (size_t)slot + 2] = 0.0f;
            }
        }
    }
}

// Any-hit over the shadow queue.  With accumulate, an unoccluded ray adds
// (weight, 1) to its pixel (shadow.metal:6-29); otherwise the occlusion flag
// is written to `occluded`.
template <bool STATS, bool TWO>
__global__ __launch_bounds__(kTraceBlock, TWO ? kTwoLevelMinWaves : kTraceMinWaves) void trace_any_kernel(AnyArgs a) {
    __shared__ uint2 lds[kLdsStack * kTraceBlock];
    __shared__ float world[TWO ? 6 * kTraceBlock : 1];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * kTraceBlock + lane;
    StackRef stk{(lds_u64*)(lds + lane), (global_u64*)(a.spill + gtid), gridDim.x * kTraceBlock,
                 (lds_f32*)(world + (TWO ? lane : 0u))};
    const uint32_t* count_ptr = a.count_ptr;
    uint32_t* fetch = a.fetch;
    if (a.iter_dev) {
        IterCounters& ic = a.counters->it[iteration_of(a.iter_dev, a.iter)];
        count_ptr = &ic.shadows;
        fetch = &ic.fetch_any;
    }
    const uint32_t count = count_ptr ? *count_ptr : a.count;
    unsigned long long sn = 0, st = 0;
    BatchScheduler sched(count);
    if (blockIdx.x >= sched.nb) return;
    const ShadowQueue& q = a.shadows;
    RayState s;
    bool active = false;
    uint32_t i = 0;
    RayPool pool(sched, fetch, count);
    while (true) {
        if (pool.refill(active)) {
            i = pool.ray;
            const vec3 o(q.ox[i], q.oy[i], q.oz[i]);
            const vec3 d(q.dx[i], q.dy[i], q.dz[i]);
            const float tmin = a.tmin_arr ? a.tmin_arr[i] : kEps;
            const float tmax = q.tmax[i];
            uint32_t root = TWO ? a.root2l : 0u;
            if (a.batch.jobs) root = a.batch.jobs[job_of_pixel(a.batch.jobs, a.batch.num_jobs, q.pixel[i])].root;
            s.init(root, o, d, tmin, tmax);
            if (TWO) stk.put_world(o, d);
            active = true;
        }
        if (!__any(active)) break;
        if (active && s.step<true, STATS, TWO>(a.nodes, a.tris, a.tlas, stk)) {
            const bool hit = s.found;
            if (a.image) {
                if (!hit) {
                    float* px = a.image + 4u * (size_t)q.pixel[i];
                    atomicAdd(px + 0, q.wx[i]);
                    atomicAdd(px + 1, q.wy[i]);
                    atomicAdd(px + 2, q.wz[i]);
                    atomicAdd(px + 3, 1.0f);
                    // a training path has one shadow ray in flight, so this is its only writer
                    const uint32_t slot = q.slot ? q.slot[i] : kRcNoSlot;
                    if (slot != kRcNoSlot && a.slot_radiance) {
                        float* sr = a.slot_radiance + 3 * (size_t)(slot & ~kRcRecorded);
                        sr[0] += q.wx[i];
                        sr[1] += q.wy[i];
                        sr[2] += q.wz[i];
                    }
                }
            } else {
                a.occluded[i] = hit ? 1u : 0u;
            }
            if (STATS) {
                sn += s.n_nodes;
                st += s.n_tris;
            }
            active = false;
        }
    }
    if (STATS) {
        atomicAdd(&a.counters->nodes_any, sn);
        atomicAdd(&a.counters->tris_any, st);
    }
}

}  // namespace rm

// ---------------------------------------------------------------------------
// Host launchers

namespace rm {

namespace {
// The two-level kernels hold a few more registers than the flat ones; their
// persistent grid shrinks with their occupancy, so that every wave of it is
// resident at once.  Callers size grids (and the spill rows) for the flat
// kernels.
int two_level_blocks(int blocks, bool any) {
    static const int ratio[2][2] = {
        {trace_blocks_per_cu(false), [] {
             int n = 0;
             (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_closest_kernel<false, true>, kTraceBlock, 0);
             return n;
         }()},
        {trace_blocks_per_cu(true), [] {
             int n = 0;
             (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_any_kernel<false, true>, kTraceBlock, 0);
             return n;
         }()}};
    const int flat = ratio[any][0], two = ratio[any][1];
    if (flat <= 0 || two <= 0 || two >= flat) return blocks;
    return std::max(1, (int)((long long)blocks * two / flat));
}
}  // namespace

hipError_t launch_trace_closest(const ClosestArgs& a, int blocks, bool stats, hipStream_t stream) {
    if (a.tlas) {
        blocks = two_level_blocks(blocks, false);
        if (stats) hipLaunchKernelGGL(HIP_KERNEL_NAME(trace_closest_kernel<true, true>), dim3(blocks), dim3(kTraceBlock), 0, stream, a);
        else hipLaunchKernelGGL(HIP_KERNEL_NAME(trace_closest_kernel<false, true>), dim3(blocks), dim3(kTraceBlock), 0, stream, a);
        return hipGetLastError();
    }
    if (stats) hipLaunchKernelGGL(HIP_KERNEL_NAME(trace_closest_kernel<true, false>), dim3(blocks), dim3(kTraceBlock), 0, stream, a);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(trace_closest_kernel<false, false>), dim3(blocks), dim3(kTraceBlock), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_trace_any(const AnyArgs& a, int blocks, bool stats, hipStream_t stream) {
    if (a.tlas) {
        blocks = two_level_blocks(blocks, true);
        if (stats) hipLaunchKernelGGL(HIP_KERNEL_NAME(trace_any_kernel<true, true>), dim3(blocks), dim3(kTraceBlock), 0, stream, a);
        else hipLaunchKernelGGL(HIP_KERNEL_NAME(trace_any_kernel<false, true>), dim3(blocks), dim3(kTraceBlock), 0, stream, a);
        return hipGetLastError();
    }
    if (stats) hipLaunchKernelGGL(HIP_KERNEL_NAME(trace_any_kernel<true, false>), dim3(blocks), dim3(kTraceBlock), 0, stream, a);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(trace_any_kernel<false, false>), dim3(blocks), dim3(kTraceBlock), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_regenerate(const RegenArgs& a, int blocks, hipStream_t stream) {
    hipLaunchKernelGGL(regenerate_kernel, dim3(blocks), dim3(256), 0, stream, a);
    return hipGetLastError();
}

int trace_blocks_per_cu(bool any) {
    int n = 0;
    if (any) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_any_kernel<false, false>, kTraceBlock, 0);
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_closest_kernel<false, false>, kTraceBlock, 0);
    return n;
}

}  // namespace rm
