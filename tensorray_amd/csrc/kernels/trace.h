// Ray queries against the 4-wide BVH: watertight FP32 ray/triangle test and
// closest-hit / any-hit traversal with a short stack in LDS that spills to a
// per-thread HBM slice only when a path is deeper than the LDS part.
//
// Replaces RayIntersectOptix (LibTensorRay/Renderer/Optix.cpp:252-305) and
// the OptiX programs __raygen__/__miss__/__closesthit__ (ptx.cu:61-113).
// Hit record semantics follow ptx.cu:92-113: a miss has tri = -1, u = v = 0,
// t = -1; barycentrics (u, v) weight vertices 1 and 2.
#pragma once
#include "../common/tr_layout.h"

namespace trc {

// Ray precomputation of the watertight test (Woop, Benthin, Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013): kz is the dominant axis of
// the direction, kx / ky the other two (swapped when d[kz] < 0 to keep the
// winding), S the shear constants.
struct RayPre {
  F3 o;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

TR_HD float comp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

TR_HD RayPre ray_pre(F3 o, F3 d) {
  RayPre r;
  r.o = o;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  float dz = comp(d, kz);
  if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
  r.kx = kx; r.ky = ky; r.kz = kz;
  r.Sx = comp(d, kx) / dz;
  r.Sy = comp(d, ky) / dz;
  r.Sz = 1.0f / dz;
  return r;
}

// Watertight test.  Returns true and (t, u, v) when the ray's line crosses
// the triangle (edges included, degenerate triangles excluded); the caller
// applies the (tmin, tmax) interval.
TR_HD bool tri_intersect(const RayPre& r, F3 p0, F3 p1, F3 p2, float* t_out, float* u_out,
                         float* v_out) {
  F3 A = p0 - r.o, B = p1 - r.o, C = p2 - r.o;
  float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  float Ax = comp(A, r.kx) - r.Sx * Akz;
  float Ay = comp(A, r.ky) - r.Sy * Akz;
  float Bx = comp(B, r.kx) - r.Sx * Bkz;
  float By = comp(B, r.ky) - r.Sy * Bkz;
  float Cx = comp(C, r.kx) - r.Sx * Ckz;
  float Cy = comp(C, r.ky) - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    // edge case: redo the 2-D edge functions in double precision
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  float inv = 1.0f / det;
  *t_out = T * inv;
  *u_out = V * inv;
  *v_out = W * inv;
  return true;
}

struct Hit {
  int32_t tri;
  float u, v, t;
};

// Conservative slab test factor 1 + 2*gamma(3) (gamma(n) = n eps / (1 - n eps),
// eps = 2^-24): the far distance is inflated so rounding never culls a box
// the exact ray enters.
constexpr float kSlabScale = 1.0000003576278687f;

TR_HD float safe_inv(float d) {
  return fabsf(d) > 1e-30f ? 1.0f / d : (d >= 0.0f ? 1e30f : -1e30f);
}

// Per-thread traversal context: LDS part of the stack (stride = block size,
// lane-interleaved so consecutive lanes use consecutive banks) and an HBM
// spill slice for deeper entries.
struct TraceCtx {
  uint32_t* lds;        // &lds_stack[threadIdx.x]
  uint32_t* leafq;      // postponed-leaf queue of this lane (kLeafQueue entries, stride lds_stride)
  int lds_stride;       // blockDim.x
  int lds_depth;        // entries held in LDS
  uint32_t* spill;      // the grid's spill buffer [depth][grid threads] (wave-uniform)
  int spill_stride;     // total threads of the grid
  int cap;              // total stack capacity (lds_depth + spill depth)
  // this lane's result slots in LDS for the wave-wide leaf tests of
  // trace_run<true> (best hit key, its barycentrics); null where unused
  unsigned long long* slot_key;
  float2* slot_uv;
  // work counters (node visits, triangle tests, queries); read only by the
  // statistics kernel, otherwise dead code
  uint32_t n_nodes, n_tris, n_queries;
};

// This thread's column of the spill buffer, computed where it is used so
// that no per-lane pointer stays live through the traversal.
__device__ inline uint32_t spill_lane() { return blockIdx.x * blockDim.x + threadIdx.x; }

__device__ inline void stack_push(TraceCtx& c, int& sp, uint32_t v) {
  if (sp < c.lds_depth) c.lds[sp * c.lds_stride] = v;
  else c.spill[(size_t)(sp - c.lds_depth) * c.spill_stride + spill_lane()] = v;
  ++sp;
}
__device__ inline uint32_t stack_pop(TraceCtx& c, int& sp) {
  --sp;
  return sp < c.lds_depth ? c.lds[sp * c.lds_stride]
                          : c.spill[(size_t)(sp - c.lds_depth) * c.spill_stride + spill_lane()];
}

__device__ inline F3 tri_vertex(const float* v) { return F3{v[0], v[1], v[2]}; }

// State of one ray query between calls of trace_run.  The stack and the
// postponed-leaf queue stay in LDS (and the spill slice) meanwhile, so a
// wave can leave the traversal loop with some lanes unfinished, give its
// finished lanes new rays, and continue all of them together.
struct TraceState {
  float best_t, u, v;  // closest accepted hit so far (t, barycentrics)
  uint32_t best_id;    // its triangle, kNoHit if none
  uint32_t node;       // next inner node to visit
  uint32_t bits;       // stack depth (bits 0-7), queued leaves (8-11), inner_done (12), found (13),
                       // chain (14), second (15), first query hit (16)
};
constexpr uint32_t kNoHit = 0xFFFFFFFFu;
constexpr uint32_t kTsInnerDone = 1u << 12;
constexpr uint32_t kTsFound = 1u << 13;
// A chained pair of queries from one origin: an any-hit query (a shadow
// ray) and then a closest-hit query in a second direction.  kTsChain: the
// second query is still to start; kTsSecond: the current query is the
// second; kTsFirstHit: the first query found a hit.
constexpr uint32_t kTsChain = 1u << 14;
constexpr uint32_t kTsSecond = 1u << 15;
constexpr uint32_t kTsFirstHit = 1u << 16;
constexpr uint32_t kTsKeep = kTsChain | kTsSecond | kTsFirstHit;  // bits carried across trace_run calls

// A new query with far bound tmax.
__device__ inline void trace_begin(TraceState& s, float tmax) {
  s.best_t = tmax;
  s.u = 0.0f;
  s.v = 0.0f;
  s.best_id = kNoHit;
  s.node = 0u;
  s.bits = 0u;
}
// A chained pair: the any-hit query with far bound tmax, then the
// closest-hit query (trace_run's d and d2, tmax2).
__device__ inline void trace_begin_chain(TraceState& s, float tmax) {
  trace_begin(s, tmax);
  s.bits = kTsChain;
}
// No query (a lane without a ray): finished and missed.
__device__ inline void trace_none(TraceState& s) {
  s.best_id = kNoHit;
  s.bits = kTsInnerDone;
}
// The query (or both of a chained pair) has finished: no inner node left,
// no leaf queued, no chained query to start.
__device__ inline bool trace_finished(const TraceState& s) {
  return (s.bits & (kTsInnerDone | 0xF00u | kTsChain)) == kTsInnerDone;
}
// Of a finished chained pair: whether the first (any-hit) query hit.
__device__ inline bool trace_first_hit(const TraceState& s) { return (s.bits & kTsFirstHit) != 0u; }
// Hit record of a finished query (ptx.cu:92-113 semantics).
__device__ inline Hit trace_hit(const TraceState& s) {
  Hit h;
  if (s.best_id != kNoHit) {
    h.tri = (int32_t)s.best_id;
    h.u = s.u;
    h.v = s.v;
    h.t = s.best_t;
  } else {
    h.tri = -1;
    h.u = 0.0f;
    h.v = 0.0f;
    h.t = -1.0f;
  }
  return h;
}

// Stalled lanes (queue full, or no inner node left with leaves queued), in
// 64ths of the live lanes, at which trace_run<true> tests the queued leaves.
constexpr int kWaveStall64 = 16;
// Lanes with leaves queued, in 64ths of the live lanes, at which it does.
constexpr int kWavePend64 = 48;

// Leaf tests of the whole wave at once.  The queued leaves of all 64 lanes
// form one list (lane by lane, in queue order), dealt round-robin over the
// lanes: lane j tests items j, j + 64, ... against the ray of the item's
// owner, read from the owner's registers, and offers its best hit to the
// owner's slot in LDS with an atomic minimum on the key (t bits, triangle
// id), which orders positive t first and ties by the smaller id, as the
// lane-serial loop does.  The winner of each round also leaves its
// barycentrics in the owner's slot.  So a flush takes ceil(total / 64)
// rounds instead of the longest queue's length, and the hit of every query
// is the one the lane-serial loop finds (closest hit) or a hit inside the
// query's interval (any hit).  All 64 lanes of the wave must be active.
__device__ inline void flush_wave(const SceneDev& S, TraceCtx& ctx, const RayPre& rp, float tmin, bool any,
                                  int nq, float& best_t, float& hu, float& hv, uint32_t& best_id, bool& found) {
  const int lane = (int)(threadIdx.x & 63u);
  // exclusive prefix sum of the queue lengths over the wave (nq < 16: one
  // ballot per bit)
  const uint64_t below = (1ull << lane) - 1ull;
  int pre = 0, total = 0;
#pragma unroll
  for (int bit = 0; bit < 4; ++bit) {
    const uint64_t m = __ballot((nq >> bit) & 1);
    pre += __popcll(m & below) << bit;
    total += __popcll(m) << bit;
  }
  const int incl = pre + nq;
  unsigned long long* keys = ctx.slot_key - lane;
  float2* uvs = ctx.slot_uv - lane;
  const unsigned long long key0 = ((unsigned long long)__float_as_uint(best_t) << 32) | best_id;
  keys[lane] = key0;
  // LDS accesses of the lanes of one wave execute in program order; the
  // wavefront fences keep the compiler from moving the slot accesses across
  // the rounds
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const int kxyz = rp.kx | (rp.ky << 2) | (rp.kz << 4) | (any ? 64 : 0);
  const uint32_t* leafq_w = ctx.leafq - lane;
  const int qs = ctx.lds_stride;
  for (int base = 0; base < total; base += 64) {
    const int k = base + lane;
    // owner: the number of lanes whose items all lie below k
    int owner = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
      if (__shfl(incl, owner + step - 1) <= k) owner += step;
    }
    const int qi = k - __shfl(pre, owner);
    RayPre r;
    r.o = F3{__shfl(rp.o.x, owner), __shfl(rp.o.y, owner), __shfl(rp.o.z, owner)};
    r.Sx = __shfl(rp.Sx, owner);
    r.Sy = __shfl(rp.Sy, owner);
    r.Sz = __shfl(rp.Sz, owner);
    const int kk = __shfl(kxyz, owner);
    r.kx = kk & 3;
    r.ky = (kk >> 2) & 3;
    r.kz = (kk >> 4) & 3;
    const bool any_o = (kk & 64) != 0;
    const float bt_o = __shfl(best_t, owner);
    const uint32_t bid_o = (uint32_t)__shfl((int)best_id, owner);
    unsigned long long bkey = ~0ull;
    float bu = 0.0f, bv = 0.0f;
    if (k < total) {
      const uint32_t c = leafq_w[owner + qi * qs];
      uint32_t cur = leaf_first(c);
      for (uint32_t left = leaf_count(c); left > 0u; --left, ++cur) {
        ctx.n_tris++;
        const TriRecord& tr = S.tris[cur];
        float4 a = *reinterpret_cast<const float4*>(tr.v0);
        float4 b = *reinterpret_cast<const float4*>(tr.v1);
        float4 e = *reinterpret_cast<const float4*>(tr.v2);
        float tt, uu, vv;
        if (!tri_intersect(r, F3{a.x, a.y, a.z}, F3{b.x, b.y, b.z}, F3{e.x, e.y, e.z}, &tt, &uu, &vv)) continue;
        const uint32_t id = __float_as_uint(a.w);
        // inside (tmin, tmax) and better than the owner's hit so far
        if (!(tt > tmin)) continue;
        const bool ok = any_o ? tt < bt_o : (tt < bt_o || (tt == bt_o && bid_o != kNoHit && id < bid_o));
        if (!ok) continue;
        const unsigned long long key = ((unsigned long long)__float_as_uint(tt) << 32) | id;
        if (key < bkey) {
          bkey = key;
          bu = uu;
          bv = vv;
        }
      }
    }
    if (bkey != ~0ull) {
      atomicMin(&keys[owner], bkey);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (keys[owner] == bkey) uvs[owner] = make_float2(bu, bv);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  const unsigned long long kf = keys[lane];
  if (kf != key0) {
    best_t = __uint_as_float((uint32_t)(kf >> 32));
    best_id = (uint32_t)kf;
    const float2 uv = uvs[lane];
    hu = uv.x;
    hv = uv.y;
    if (any) found = true;
  }
}

// Closest hit (any = false) or first found hit (any = true) in (tmin, tmax),
// tmax being the far bound the query's state was begun with.
// Ties in t resolve to the smallest triangle id, so the result does not
// depend on traversal order.  A ray with tmax < tmin visits the root only
// and misses (used for lanes without a query).
//
// Per node visit: slab tests of the four children; hit inner children are
// sorted near to far and pushed, hit leaves are appended to a per-lane queue
// in LDS.  Leaves are tested later, in one flat loop over (leaf, triangle)
// pairs, when the wave decides to: at least three quarters of its live lanes
// have queued leaves, or some lane has no inner node left or a nearly full
// queue (the threshold was picked by timing 1/4, 1/2 and 3/4).
// Testing a leaf late only loosens culling (best_t shrinks later); it never
// changes the result, since the closest hit does not depend on test order.
// Lanes run this loop in step with ballots, so the triangle test executes
// with most lanes active instead of once per child slot that any lane hit.
//
// Runs the wave's queries until all have finished or until at most
// keep64/64 of the queries live on entry are still unfinished (keep64 = 0:
// to the end).  The state of unfinished queries is in `st`; calling again
// with the same ray continues them exactly where they stopped, so the
// result is the same however the traversal is split.
//
// kWave: the queued leaves are tested by the whole wave (flush_wave), which
// needs all 64 lanes active and the LDS slots of the context; otherwise
// each lane tests its own queue.
//
// A chained pair (trace_begin_chain) runs its any-hit query along d, then,
// within the same loop, its closest-hit query along d2 from the same origin
// with far bound tmax2.
template <bool kWave>
__device__ inline void trace_run(const SceneDev& S, TraceCtx& ctx, F3 o, F3 d, F3 d2, float tmax2, float tmin,
                                 bool any, TraceState& st, int keep64) {
  uint32_t keep = st.bits & kTsKeep;
  const F3 dq = (keep & kTsSecond) ? d2 : d;
  bool anyq = any && !(keep & kTsSecond);
  RayPre rp = ray_pre(o, dq);
  float ix = safe_inv(dq.x), iy = safe_inv(dq.y), iz = safe_inv(dq.z);
  float best_t = st.best_t, hu = st.u, hv = st.v;
  uint32_t best_id = st.best_id;
  uint32_t node = st.node;
  int sp = (int)(st.bits & 0xFFu);
  int nq = (int)((st.bits >> 8) & 0xFu);
  bool inner_done = (st.bits & kTsInnerDone) != 0u;
  bool found = (st.bits & kTsFound) != 0u;
  int n_keep = 0;
  if (keep64 > 0) n_keep = (__popcll(__ballot(!inner_done || nq > 0 || (keep & kTsChain))) * keep64) >> 6;
  const int qs = ctx.lds_stride;
  for (;;) {
    // ---- the first query of a chained pair has ended: start the second ----
    if ((keep & kTsChain) && inner_done && nq == 0) {
      keep = kTsSecond | (best_id != kNoHit ? kTsFirstHit : 0u);
      anyq = false;
      rp = ray_pre(o, d2);
      ix = safe_inv(d2.x);
      iy = safe_inv(d2.y);
      iz = safe_inv(d2.z);
      best_t = tmax2;
      hu = 0.0f;
      hv = 0.0f;
      best_id = kNoHit;
      node = 0u;
      sp = 0;
      inner_done = false;
      found = false;
    }
    // ---- visit one inner node (lanes with room in their leaf queue) ----
    if (!inner_done && nq <= kLeafQueue - 4) {
      ctx.n_nodes++;
      const float4* np = reinterpret_cast<const float4*>(S.nodes + node);
      float4 lox = np[0], loy = np[1], loz = np[2], hix = np[3], hiy = np[4], hiz = np[5];
      uint4 ch = reinterpret_cast<const uint4*>(np)[6];
      float lo_x[4] = {lox.x, lox.y, lox.z, lox.w}, lo_y[4] = {loy.x, loy.y, loy.z, loy.w};
      float lo_z[4] = {loz.x, loz.y, loz.z, loz.w}, hi_x[4] = {hix.x, hix.y, hix.z, hix.w};
      float hi_y[4] = {hiy.x, hiy.y, hiy.z, hiy.w}, hi_z[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
      uint32_t cid[4] = {ch.x, ch.y, ch.z, ch.w};
      float kt[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
      uint32_t kc[4] = {kEmpty, kEmpty, kEmpty, kEmpty};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        float t0x = (lo_x[s] - o.x) * ix, t1x = (hi_x[s] - o.x) * ix;
        float t0y = (lo_y[s] - o.y) * iy, t1y = (hi_y[s] - o.y) * iy;
        float t0z = (lo_z[s] - o.z) * iz, t1z = (hi_z[s] - o.z) * iz;
        float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
        float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), best_t));
        uint32_t c = cid[s];
        bool hit = c != kEmpty && tn <= tf * kSlabScale;
        if (hit && (c & kLeafBit)) {
          ctx.leafq[nq * qs] = c;
          ++nq;
        }
        if (hit && !(c & kLeafBit)) {
          kt[s] = tn;
          kc[s] = c;
        }
      }
      // 4-element sorting network, near to far
#define TR_CSWAP(i, j)                                                   \
      if (kt[j] < kt[i]) {                                               \
        float _t = kt[i]; kt[i] = kt[j]; kt[j] = _t;                     \
        uint32_t _c = kc[i]; kc[i] = kc[j]; kc[j] = _c;                  \
      }
      TR_CSWAP(0, 1) TR_CSWAP(2, 3) TR_CSWAP(0, 2) TR_CSWAP(1, 3) TR_CSWAP(1, 2)
#undef TR_CSWAP
      if (kc[0] != kEmpty) {
        int npush = (kc[1] != kEmpty) + (kc[2] != kEmpty) + (kc[3] != kEmpty);
        if (sp + npush > ctx.cap) {
          // unreachable: the host refuses a launch whose BVH needs more than
          // `cap` entries (stack_need, computed at build time); this branch
          // only keeps the stack in bounds
          inner_done = true;
        } else {
          if (kc[3] != kEmpty) stack_push(ctx, sp, kc[3]);
          if (kc[2] != kEmpty) stack_push(ctx, sp, kc[2]);
          if (kc[1] != kEmpty) stack_push(ctx, sp, kc[1]);
          node = kc[0];
        }
      } else if (sp > 0) {
        node = stack_pop(ctx, sp);
      } else {
        inner_done = true;
      }
    }
    // ---- decide, wave-wide, whether queued leaves are tested now ----
    const bool pend = nq > 0;
    const bool live = !inner_done || pend || (keep & kTsChain);
    const uint64_t m_live = __ballot(live);
    if (m_live == 0ull) break;
    if (__popcll(m_live) <= n_keep) break;
    const uint64_t m_pend = __ballot(pend);
    const uint64_t m_urgent = __ballot(pend && (inner_done || nq > kLeafQueue - 4));
    bool flush;
    if constexpr (kWave) {
      // a wave-wide flush costs one round for up to 64 leaves: wait until
      // kWaveStall64/64 of the live lanes are stalled on their queue,
      // kWavePend64/64 have leaves queued, or no lane can visit a node
      const uint64_t m_visit = __ballot(!inner_done && nq <= kLeafQueue - 4);
      flush = m_visit == 0ull || __popcll(m_urgent) * 64 >= kWaveStall64 * __popcll(m_live) ||
              __popcll(m_pend) * 64 >= kWavePend64 * __popcll(m_live);
    } else {
      flush = m_urgent != 0ull || 4 * __popcll(m_pend) >= 3 * __popcll(m_live);
    }
    if constexpr (kWave) {
      if (flush) {
        flush_wave(S, ctx, rp, tmin, anyq, nq, best_t, hu, hv, best_id, found);
        nq = 0;
        if (found) inner_done = true;
      }
    } else if (flush && pend) {
      int qi = 0;
      uint32_t cur = 0u, left = 0u;
      for (;;) {
        if (left == 0u) {
          if (qi == nq) break;
          const uint32_t c = ctx.leafq[qi * qs];
          ++qi;
          cur = leaf_first(c);
          left = leaf_count(c);
        }
        ctx.n_tris++;
        const TriRecord& tr = S.tris[cur];
        ++cur;
        --left;
        float4 a = *reinterpret_cast<const float4*>(tr.v0);
        float4 b = *reinterpret_cast<const float4*>(tr.v1);
        float4 e = *reinterpret_cast<const float4*>(tr.v2);
        float tt, uu, vv;
        if (!tri_intersect(rp, F3{a.x, a.y, a.z}, F3{b.x, b.y, b.z}, F3{e.x, e.y, e.z}, &tt, &uu, &vv)) continue;
        uint32_t id = __float_as_uint(a.w);
        // best_t is the far bound tmax until a hit is accepted, and below
        // it afterwards, so these tests keep every hit inside (tmin, tmax)
        if (!(tt > tmin)) continue;
        if (anyq) {
          if (!(tt < best_t)) continue;
          best_id = id; best_t = tt; hu = uu; hv = vv;
          found = true;
          break;
        }
        if (tt < best_t || (tt == best_t && best_id != kNoHit && id < best_id)) {
          best_t = tt; best_id = id; hu = uu; hv = vv;
        }
      }
      nq = 0;
      if (found) inner_done = true;
    }
  }
  st.best_t = best_t;
  st.u = hu;
  st.v = hv;
  st.best_id = best_id;
  st.node = node;
  st.bits = (uint32_t)sp | ((uint32_t)nq << 8) | (inner_done ? kTsInnerDone : 0u) | (found ? kTsFound : 0u) | keep;
}

// One query to the end.
__device__ inline Hit trace_ray_rt(const SceneDev& S, TraceCtx& ctx, F3 o, F3 d, float tmin, float tmax, bool any) {
  ctx.n_queries++;
  TraceState st;
  trace_begin(st, tmax);
  trace_run<false>(S, ctx, o, d, d, tmax, tmin, any, st, 0);
  return trace_hit(st);
}

template <bool kAnyHit>
__device__ inline Hit trace_ray(const SceneDev& S, TraceCtx& ctx, F3 o, F3 d, float tmin, float tmax) {
  return trace_ray_rt(S, ctx, o, d, tmin, tmax, kAnyHit);
}

}  // namespace trc
