// pt_kernels.hip — the wavefront path tracer for gfx950.
//
// One pass of the wavefront: raygen -> { extend -> shade -> shadow -> advance } x max_depth -> film.
//   raygen  : camera rays and per-path RNG seeding          (ScalarCpuPathTracer.Render.cpp:55-119, :228-262)
//   extend  : closest-hit BVH traversal, persistent waves  (CpuBvhAccelerator.cpp:360-458)
//   shade   : one trace() iteration per live path          (ScalarCpuPathTracer.cpp:1148-1289)
//   shadow  : NEE visibility, any-hit traversal            (ScalarCpuPathTracer.cpp:1066-1087)
//   film    : per-pixel accumulation in tile order          (FilmBuffer.cpp:80-91)
//
// Every float expression is written in the operand order of the reference and the
// library is built with -ffp-contract=off and correctly rounded fp32 div/sqrt, so a
// path makes exactly the decisions the CPU restatement makes.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "pt_device.h"
#include "pt_math.h"

// Pointer qualifiers for the LDS and global address spaces, where an access must be issued as that
// memory's instruction (a pointer that may address either compiles to a generic access).
#define PT_LDS __attribute__((address_space(3)))
#define PT_GLOBAL __attribute__((address_space(1)))

namespace ptd {
namespace {

constexpr float kEps = 1e-4f;                       // ScalarCpuPathTracer.cpp:16
constexpr float kPi = 3.14159265358979323846f;     // :14
constexpr float kInvPi = 0.31830988618f;           // :15
constexpr float kInf = __builtin_huge_valf();

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 muls(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 mulv(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ float sclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
__device__ __forceinline__ bool finite(float v) { return __builtin_isfinite(v); }
__device__ __forceinline__ float clamp01(float v) { return finite(v) ? smin(1.0f, smax(0.0f, v)) : 0.0f; }
__device__ __forceinline__ float vmax3(V3 a) { return smax(a.x, smax(a.y, a.z)); }
__device__ __forceinline__ V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

// normalize(), ScalarCpuPathTracer.cpp:104-110
__device__ __forceinline__ V3 normalize(V3 a) {
  const float l = __builtin_sqrtf(dot(a, a));
  if (l <= kEps) return mk(0.0f, 1.0f, 0.0f);
  return divs(a, l);
}

// splitmix64 / Rng, ScalarCpuPathTracer.cpp:36-41, :383-393
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t rng_seed(uint64_t seed, uint64_t frame, uint64_t sample, uint64_t path_id,
                                             uint64_t pixel) {
  // dimension = 0, path_depth = 0 for both streams the path tracer creates
  const uint64_t z = splitmix64(0);
  const uint64_t mix = splitmix64(seed ^ splitmix64(frame) ^ splitmix64(sample) ^ splitmix64(path_id) ^ z ^ z);
  return splitmix64(mix + splitmix64(pixel));
}
__device__ __forceinline__ float next01(uint64_t& s) {
  s = splitmix64(s);
  return static_cast<float>(static_cast<uint32_t>((s >> 8) & 0x00ffffffu)) * (1.0f / 16777215.0f);
}

// ------------------------------------------------------------------ traversal
struct RayQ {
  float o[3], inv[3];
  bool par[3];
  bool general;  // some axis parallel, or a non-finite origin/direction: take the select form of slab
};

__device__ __forceinline__ RayQ make_query(V3 o, V3 d) {
  RayQ q;
  q.o[0] = o.x; q.o[1] = o.y; q.o[2] = o.z;
  const float dd[3] = {d.x, d.y, d.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    q.par[k] = __builtin_fabsf(dd[k]) <= kEps;
    q.inv[k] = q.par[k] ? 0.0f : 1.0f / dd[k];
  }
  q.general = q.par[0] || q.par[1] || q.par[2] || !(finite(o.x) && finite(o.y) && finite(o.z) && finite(d.x) &&
                                                   finite(d.y) && finite(d.z));
  return q;
}

// intersect_aabb, CpuBvhAccelerator.cpp:125-154 (the early returns do not change the result).
__device__ __forceinline__ bool slab(const RayQ& q, float mnx, float mny, float mnz, float mxx, float mxy,
                                     float mxz, float max_t, float& t_near) {
  const float mn[3] = {mnx, mny, mnz}, mx[3] = {mxx, mxy, mxz};
  float tmin = kEps, tmax = max_t;
  bool outside = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (q.par[k]) {
      outside = outside || q.o[k] < mn[k] || q.o[k] > mx[k];
    } else {
      float t0 = (mn[k] - q.o[k]) * q.inv[k];
      float t1 = (mx[k] - q.o[k]) * q.inv[k];
      const float lo = (t0 > t1) ? t1 : t0;
      const float hi = (t0 > t1) ? t0 : t1;
      tmin = smax(tmin, lo);
      tmax = smin(tmax, hi);
    }
  }
  t_near = tmin;
  return !outside && !(tmin > tmax) && tmax > kEps;
}

// The same test for a ray with no parallel axis and finite origin, direction and box: every t
// is then a non-NaN float, so the std::min/std::max selects equal v_min/v_max except for the sign
// of a zero, which cannot change tmin (>= kEps) nor the outcome (a zero tmax fails tmax > kEps).
// Both children of a node in one pass.
__device__ __forceinline__ void slab2_fast(const RayQ& q, const float4 a, const float4 b, const float4 c, float max_t,
                                           bool& h0, float& t0, bool& h1, float& t1) {
  const float ox = q.o[0], oy = q.o[1], oz = q.o[2], ix = q.inv[0], iy = q.inv[1], iz = q.inv[2];
  {
    const float x0 = (a.x - ox) * ix, x1 = (a.w - ox) * ix;
    const float y0 = (a.y - oy) * iy, y1 = (b.x - oy) * iy;
    const float z0 = (a.z - oz) * iz, z1 = (b.y - oz) * iz;
    const float tmin = fmaxf(fmaxf(fmaxf(kEps, fminf(x0, x1)), fminf(y0, y1)), fminf(z0, z1));
    const float tmax = fminf(fminf(fminf(max_t, fmaxf(x0, x1)), fmaxf(y0, y1)), fmaxf(z0, z1));
    t0 = tmin;
    h0 = tmin <= tmax && tmax > kEps;
  }
  {
    const float x0 = (b.z - ox) * ix, x1 = (c.y - ox) * ix;
    const float y0 = (b.w - oy) * iy, y1 = (c.z - oy) * iy;
    const float z0 = (c.x - oz) * iz, z1 = (c.w - oz) * iz;
    const float tmin = fmaxf(fmaxf(fmaxf(kEps, fminf(x0, x1)), fminf(y0, y1)), fminf(z0, z1));
    const float tmax = fminf(fminf(fminf(max_t, fmaxf(x0, x1)), fmaxf(y0, y1)), fmaxf(z0, z1));
    t1 = tmin;
    h1 = tmin <= tmax && tmax > kEps;
  }
}

// Per-lane counts (visits, tests, hits) and per-wave counts (inner, leaves: the lanes that took an
// inner-node or a leaf step, added once per step from the step's lane count).
struct TravStats {
  uint32_t visits, tests, hits;
  uint32_t inner, leaves;
};

// Moller-Trumbore against precomputed edges, intersect_scalar (BvhTriangleIntersector.cpp:35-74).
// True when eps < t < best_t.
// The record is v0 + primitive id, e1, e2 in three 16-byte slots; the unused fourth floats of e1 and e2
// are not read (40 bytes per test).
__device__ __forceinline__ bool tri_test(const float4* tp, V3 o, V3 d, float best_t, float& t_out, float& u_out,
                                         float& v_out) {
  const float4 v0 = tp[0];
  const float3 e1 = *reinterpret_cast<const float3*>(tp + 1), e2 = *reinterpret_cast<const float3*>(tp + 2);
  const float hx = d.y * e2.z - d.z * e2.y;
  const float hy = d.z * e2.x - d.x * e2.z;
  const float hz = d.x * e2.y - d.y * e2.x;
  const float det = e1.x * hx + e1.y * hy + e1.z * hz;
  if (__builtin_fabsf(det) < kEps) return false;
  const float inv_det = 1.0f / det;
  const float sx = o.x - v0.x, sy = o.y - v0.y, sz = o.z - v0.z;
  const float u = (sx * hx + sy * hy + sz * hz) * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  const float qx = sy * e1.z - sz * e1.y;
  const float qy = sz * e1.x - sx * e1.z;
  const float qz = sx * e1.y - sy * e1.x;
  const float v = (d.x * qx + d.y * qy + d.z * qz) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return false;
  const float t = (e2.x * qx + e2.y * qy + e2.z * qz) * inv_det;
  t_out = t;
  u_out = u;
  v_out = v;
  return !(t <= kEps || t >= best_t);
}

// Traversal stack of one lane: the newest kStackLds entries live in an LDS ring; older entries are
// spilled to a per-lane global area (entry i at ovf[i * ovf_stride + gid], 32-bit offsets: the host
// sizes the area below 2^32 entries) when the ring is full and read back from there when the pops
// reach them.
// The ring pointers carry the LDS address space and the spill pointers the global one, so that no
// load or store can be formed that addresses either memory (a generic access and its wait on both
// counters on every pop).
struct Stack {
  PT_LDS int32_t* lds_ref;
  PT_LDS float* lds_t;
  PT_GLOBAL int32_t* ovf_ref;
  PT_GLOBAL float* ovf_t;
  uint32_t ovf_depth, ovf_stride, lane_slot, gid;
  uint32_t* error;
  const PT_LDS float* hot;  // the workgroup's LDS copy of the hot nodes (node indices hot_base + h), 16 floats each
  uint32_t hot_base, hot_count;
};

// One BVH query in flight on a lane.
struct Trav {
  RayQ q;
  V3 d;
  float best_t;
  int32_t best_prim, cur;
  uint32_t sp;   // entries on the stack
  uint32_t lo;   // entries [lo, sp) are in the LDS ring, [0, lo) in the global area
  uint32_t top;  // ring slot of the next push (entry sp's slot); the oldest ring entry's when it is full
};

// Starts a query below t_max. False when it ends at once (no primitives, or the root box is
// missed); best_prim is then -1.
__device__ __forceinline__ bool trav_begin(const SceneRec& R, Trav& s, V3 o, V3 d, float tmax, TravStats& st) {
  s.best_prim = -1;
  s.best_t = tmax;
  s.d = d;
  s.sp = 0;
  s.lo = 0;
  s.top = 0;
  s.cur = R.root_ref;
  if (R.prim_count == 0) return false;
  s.q = make_query(o, d);
  ++st.visits;
  float tn;
  return slab(s.q, R.root_box[0], R.root_box[1], R.root_box[2], R.root_box[3], R.root_box[4], R.root_box[5],
              s.best_t, tn);
}

// Traversal in three pieces, following CpuBvhAccelerator::intersect's visit order for closest hit
// (kAny = false) or the first leaf batch holding an accepted hit below best_t (kAny = true):
// children tested at the parent, the nearer pushed last (left on ties), every pop re-tested against
// the current best t. A popped entry carries the t_near of its box; t_near <= best_t is the same
// test as the reference's re-run of intersect_aabb with max_t = best_t.
enum StepResult : int { kStepGo = 0, kStepPop = 1, kStepDone = 2 };

// Inner node s.cur: descend into the nearer hit child (pushing the other when both are hit), or
// ask for a pop when neither is hit. A hot copy's record comes from the workgroup's LDS copy.
__device__ __forceinline__ int trav_inner(const DevScene& sc, const Stack& stk, Trav& s, TravStats& st) {
  float4 a, b, c;
  int32_t rx, ry;
  const uint32_t h = static_cast<uint32_t>(s.cur) - stk.hot_base;
  if (h < stk.hot_count) {
    const PT_LDS float* hp = stk.hot + 16u * h;
    a = make_float4(hp[0], hp[1], hp[2], hp[3]);
    b = make_float4(hp[4], hp[5], hp[6], hp[7]);
    c = make_float4(hp[8], hp[9], hp[10], hp[11]);
    rx = __float_as_int(hp[12]);
    ry = __float_as_int(hp[13]);
  } else {
    const float4* np = sc.nodes + 4u * static_cast<uint32_t>(s.cur);
    a = np[0];
    b = np[1];
    c = np[2];
    const int2 r = *reinterpret_cast<const int2*>(np + 3);  // the two child references (8 of the slot's 16 bytes)
    rx = r.x;
    ry = r.y;
  }
  float t0, t1;
  bool h0, h1;
  if (!s.q.general && !sc.general_boxes) {
    slab2_fast(s.q, a, b, c, s.best_t, h0, t0, h1, t1);
  } else {
    h0 = slab(s.q, a.x, a.y, a.z, a.w, b.x, b.y, s.best_t, t0);
    h1 = slab(s.q, b.z, b.w, c.x, c.y, c.z, c.w, s.best_t, t1);
  }
  if (h0 && h1) {
    const bool left_first = t0 <= t1;
    const int32_t near_ref = left_first ? rx : ry;
    const int32_t far_ref = left_first ? ry : rx;
    const float far_t = left_first ? t1 : t0;
    const uint32_t slot = s.top * kBlock + stk.lane_slot;
    if (s.sp - s.lo == static_cast<uint32_t>(kStackLds)) {  // ring full: spill its oldest entry (slot top)
      if (s.lo >= stk.ovf_depth) {
        atomicOr(stk.error, 1u);  // overflow area smaller than the tree depth: refuse, never write out of bounds
        return kStepDone;
      }
      const uint32_t o2 = s.lo * stk.ovf_stride + stk.gid;
      stk.ovf_ref[o2] = stk.lds_ref[slot];
      stk.ovf_t[o2] = stk.lds_t[slot];
      ++s.lo;
    }
    stk.lds_ref[slot] = far_ref;
    stk.lds_t[slot] = far_t;
    s.top = s.top + 1u == static_cast<uint32_t>(kStackLds) ? 0u : s.top + 1u;
    ++s.sp;
    s.cur = near_ref;
    ++st.visits;
    return kStepGo;
  }
  if (h0 || h1) {
    s.cur = h0 ? rx : ry;
    ++st.visits;
    return kStepGo;
  }
  return kStepPop;
}

// Leaf s.cur: its triangles in batches of 8 (kBvhTriangleBatchWidth). kStepDone when an any-hit
// query accepted a hit, else kStepPop.
template <bool kAny>
__device__ __forceinline__ int trav_leaf(const DevScene& sc, Trav& s, TravStats& st) {
  const uint32_t code = ~static_cast<uint32_t>(s.cur);
  const uint32_t first = code >> 5, count = code & 31u;
  const V3 o = mk(s.q.o[0], s.q.o[1], s.q.o[2]);
  uint32_t accepted = 0;
  for (uint32_t k = 0; k < count; ++k) {
    // IntersectBvhTriangleBatch skips a batch of 8 whose entry best t is not above kEps (only a
    // NaN t accepted from a NaN ray gets here) but the triangles still count as tested.
    if ((k & 7u) == 0u && !(s.best_t > kEps)) {
      st.tests += count - k;
      break;
    }
    ++st.tests;
    float t, bu, bv;
    const float4* tp = sc.tris + 3u * (first + k);
    if (tri_test(tp, o, s.d, s.best_t, t, bu, bv)) {  // commit_ordered_lane_hit, BvhTriangleIntersector.cpp:25-33
      ++accepted;
      s.best_t = t;
      s.best_prim = __float_as_int(tp[0].w);
    }
    // an any-hit query stops at the end of the first batch that accepted a hit
    if (kAny && accepted && ((k & 7u) == 7u || k + 1u == count)) {
      st.hits += accepted;
      return kStepDone;
    }
  }
  st.hits += accepted;
  return kStepPop;
}

// Pops until an entry passes t_near <= best_t (kStepGo, s.cur set) or the stack is empty (kStepDone).
// The ring slot below top is read unconditionally (an LDS read; stale when the ring is empty) and
// replaced by the spilled entry in that rare case, so the common pop stays an LDS access instead of
// one generic load that could address either memory.
__device__ __forceinline__ int trav_pop(const Stack& stk, Trav& s, TravStats& st) {
  while (s.sp > 0) {
    --s.sp;
    const uint32_t below = (s.top == 0u ? static_cast<uint32_t>(kStackLds) : s.top) - 1u;
    const uint32_t slot = below * kBlock + stk.lane_slot;
    int32_t ref = stk.lds_ref[slot];
    float t = stk.lds_t[slot];
    if (s.sp >= s.lo) {
      s.top = below;
    } else {  // below the ring: the ring is empty from here on (top stays where the next push goes)
      const uint32_t o2 = s.sp * stk.ovf_stride + stk.gid;
      ref = stk.ovf_ref[o2];
      t = stk.ovf_t[o2];
      s.lo = s.sp;
    }
    ++st.visits;
    if (t <= s.best_t) {
      s.cur = ref;
      return kStepGo;
    }
  }
  return kStepDone;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// The wave's index in the grid, read from lane 0 so that the compiler keeps it (and the chunk
// bookkeeping derived from it) in scalar registers.
__device__ __forceinline__ uint32_t global_wave() {
  return __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
}

// The XCD (accelerator complex die) the wave runs on, read from the XCC_ID hardware register.
__device__ __forceinline__ uint32_t xcc_id() {
  uint32_t v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return __builtin_amdgcn_readfirstlane(v);
}

// Ballot of a lane predicate straight from its compare mask (no select to a VGPR and back).
__device__ __forceinline__ uint64_t wave_ballot(bool pred) { return __builtin_amdgcn_ballot_w64(pred); }

// Lanes whose predicate holds, as a 32-bit scalar (compares of two counts then stay scalar instead of
// being widened to a 64-bit vector compare).
__device__ __forceinline__ uint32_t wave_count(bool pred) {
  return __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(__popcll(wave_ballot(pred))));
}

// The first active lane takes the next chunk index from the shared cursor; every lane gets it.
// Called with the wave converged.
__device__ __forceinline__ uint32_t wave_grab(uint32_t* cursor) {
  uint32_t v = 0;
  if (lanes_below(wave_ballot(true)) == 0u) v = atomicAdd(cursor, 1u);
  return __builtin_amdgcn_readfirstlane(v);
}

// A wave leaves drive_queries. It counts itself in its stripe; a stripe's last wave counts the stripe,
// and the last stripe's resets the chunk cursor and every count for the next launch on the stream:
// every other wave's last grab returned before it counted itself here.
__device__ __forceinline__ void wave_finish(uint32_t* queue, uint32_t waves) {
  if (lanes_below(wave_ballot(true)) == 0u) {
    constexpr uint32_t D = static_cast<uint32_t>(kDoneStripes);
    const uint32_t st = global_wave() % D;
    const uint32_t in_stripe = waves / D + (st < waves % D ? 1u : 0u);  // waves w < waves with w % D == st
    const uint32_t stripes = waves < D ? waves : D;                      // stripes holding a wave
    if (atomicAdd(queue + kQueueDoneStripe + st * kSegLine, 1u) == in_stripe - 1u &&
        atomicAdd(queue + kQueueWavesDone, 1u) == stripes - 1u) {
      for (int r = 0; r < kRegions; ++r) atomicExch(queue + kQueueChunkCursor + r * kSegLine, 0u);
      for (uint32_t k = 0; k < D; ++k) atomicExch(queue + kQueueDoneStripe + k * kSegLine, 0u);
      atomicExch(queue + kQueueWavesDone, 0u);
    }
  }
}

// Persistent query driver with per-lane refill. The queries [0, n) are cut into 64-entry chunks, and
// the chunks into one contiguous region per XCD; a wave takes its chunks one at a time from a cursor
// of its own XCD's region (then of the other regions, once its own is handed out), so waves that drew
// cheap rays take more chunks and the launch ends evenly. A lane that finishes takes the next entry of
// its wave's chunk as soon as `refill_min` lanes are idle, so a wave is not held by its slowest ray.
// Every wave of the grid must call it exactly once (the last one resets the cursors).
//   begin(j, s): start query j on this lane; false if it ended at once.
//   end(j, s)  : the query finished; true if the lane started a follow-up query on s.
template <bool kAny, typename Begin, typename End>
__device__ __forceinline__ void drive_queries(const DevScene& sc, const Stack& stk, uint32_t* queue, uint32_t n,
                                              uint32_t refill_min, TravStats& st, Begin begin, End end) {
  const uint32_t waves = gridDim.x * kWavesPerBlock;
  // The queue's chunks form `regions` contiguous regions of R chunks (kRegions, or one when the grid
  // has a wave for every chunk). A wave takes chunks of its own XCD's region from that region's
  // cursor, and once that region is exhausted, of the following regions in turn; it leaves when every
  // region is exhausted.
  const uint32_t nchunks = (n + 63u) / 64u;
  const uint32_t regions = nchunks > waves ? static_cast<uint32_t>(kRegions) : 1u;
  const uint32_t R = (nchunks + regions - 1u) / regions;
  uint32_t region = xcc_id() % regions;
  uint32_t exhausted = 0;  // regions whose cursor has passed their end
  auto grab = [&]() -> uint32_t {
    while (true) {
      const uint32_t c = wave_grab(queue + kQueueChunkCursor + region * kSegLine);
      if (c < R && region * R + c < nchunks) return region * R + c;
      exhausted |= 1u << region;
      if (exhausted == (1u << regions) - 1u) return nchunks;  // every chunk handed out
      while (exhausted & (1u << region)) region = (region + 1u) % regions;
    }
  };
  uint32_t chunk = grab();
  uint32_t next = 0;  // entries of `chunk` already handed out
  bool active = false;
  uint32_t j = 0;
  Trav s;
  while (true) {
    const uint32_t base = chunk * 64u;
    if (base < n) {
      const uint64_t idle = wave_ballot(!active);
      const uint32_t k = static_cast<uint32_t>(__popcll(idle));
      if (k >= refill_min || k == 64u) {
        const uint32_t left = n - base - next;
        const uint32_t avail = (64u - next) < left ? (64u - next) : left;
        const uint32_t rank = lanes_below(idle);
        if (!active && rank < avail) {
          j = base + next + rank;
          active = begin(j, s) || end(j, s);
        }
        next += k < avail ? k : avail;
        if (next >= 64u || base + next >= n) {
          chunk = grab();
          next = 0;
        }
      }
    }
    // One step per iteration for the larger of the two groups of lanes, those at inner nodes or
    // those at leaves; the others wait, so each step runs with most of its lanes busy.
    bool at_inner = active && s.cur >= 0;
    uint32_t n_inner = wave_count(at_inner);
    const uint32_t n_leaf = wave_count(active && !at_inner);
    if (n_inner + n_leaf == 0u) {
      if (chunk * 64u >= n) break;
      continue;
    }
    if (static_cast<int32_t>(n_inner - n_leaf) >= 0) {  // n_inner >= n_leaf (both <= 64)
      // Inner steps back to back for as long as the schedule above would choose them again: no
      // refill due (fewer idle lanes than refill_min, or no queries left to hand out) and inner
      // lanes at least as many as leaf lanes.
      const bool can_refill = chunk * 64u < n;
      while (true) {
        st.inner += n_inner;
        if (at_inner) {
          int r = trav_inner(sc, stk, s, st);
          if (r == kStepPop) r = trav_pop(stk, s, st);
          if (r == kStepDone) active = end(j, s);
        }
        at_inner = active && s.cur >= 0;
        n_inner = wave_count(at_inner);
        const uint32_t n_left = wave_count(active && !at_inner);
        if (n_inner == 0u || static_cast<int32_t>(n_inner - n_left) < 0 ||
            (can_refill && 64u - n_inner - n_left >= refill_min))
          break;
      }
      continue;
    }
    st.leaves += n_leaf;
    if (active && !at_inner) {
      int r = trav_leaf<kAny>(sc, s, st);
      if (r == kStepPop) r = trav_pop(stk, s, st);
      if (r == kStepDone) active = end(j, s);
    }
  }
  wave_finish(queue, waves);
}

__device__ __forceinline__ void wave_add(unsigned long long* counters, int slot, uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(static_cast<unsigned long long>(v), off);
  if ((threadIdx.x & 63u) == 0u && v != 0u)
    atomicAdd(counters + (global_wave() & (kStripes - 1u)) * kCounterSlots + slot, static_cast<unsigned long long>(v));
}

// Adds a value every lane of the wave holds (the same on all of them) once.
__device__ __forceinline__ void wave_add_uniform(unsigned long long* counters, int slot, uint32_t v) {
  if ((threadIdx.x & 63u) == 0u && v != 0u)
    atomicAdd(counters + (global_wave() & (kStripes - 1u)) * kCounterSlots + slot, static_cast<unsigned long long>(v));
}

// Shadow traversals add to the same SampleCounters fields as extension rays, as the
// reference's shadow query does (ScalarCpuPathTracer.cpp:1071); the shadow-only slots
// keep their share for the byte accounting.
__device__ __forceinline__ void flush_stats(unsigned long long* counters, const TravStats& s, bool shadow) {
  if (shadow) {
    wave_add(counters, kShadowNodeVisits, s.visits);
    wave_add(counters, kShadowTriTests, s.tests);
    wave_add_uniform(counters, kShadowInnerFetches, s.inner);
    wave_add_uniform(counters, kShadowLeafVisits, s.leaves);
  }
  wave_add(counters, kNodeVisits, s.visits);
  wave_add_uniform(counters, kLeafVisits, s.leaves);
  wave_add(counters, kTriTests, s.tests);
  wave_add(counters, kTriHits, s.hits);
  wave_add_uniform(counters, kInnerFetches, s.inner);
}

// Wave-wide append: returns this lane's slot (valid when pred is true).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  const uint64_t m = wave_ballot(pred);
  uint32_t base = 0;
  if (m != 0u) {
    const uint32_t leader = static_cast<uint32_t>(__builtin_ctzll(m));
    if ((threadIdx.x & 63u) == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(m)));
    base = __shfl(base, static_cast<int>(leader));
  }
  return base + lanes_below(m);
}

// prefix[0..kSegments] of the kSegments counts counts[s * stride], into LDS. Every thread of the
// block must call it (it ends with a barrier).
__device__ __forceinline__ void load_prefix(const uint32_t* counts, int stride, uint32_t* prefix) {
  if (threadIdx.x < 64u) {
    uint32_t v = counts[threadIdx.x * static_cast<uint32_t>(stride)];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u = __shfl_up(v, off);
      if (static_cast<int>(threadIdx.x) >= off) v += u;
    }
    prefix[threadIdx.x + 1u] = v;
    if (threadIdx.x == 0u) prefix[0] = 0u;
  }
  __syncthreads();
}

// Position of queue entry j (< prefix[kSegments]).
__device__ __forceinline__ uint32_t queue_pos(const uint32_t* prefix, uint32_t j, uint32_t seg_cap) {
  uint32_t lo = 0, hi = kSegments;  // prefix[lo] <= j < prefix[hi]
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= j) lo = mid;
    else hi = mid;
  }
  return lo * seg_cap + (j - prefix[lo]);
}

// Appends pred lanes to segment `seg` of a segmented queue; returns the position.
__device__ __forceinline__ uint32_t segment_append(uint32_t* cursors, uint32_t seg, uint32_t seg_cap, bool pred) {
  return seg * seg_cap + wave_append(cursors + seg * kSegLine, pred);
}

// ------------------------------------------------------------------ kernels
// Initial queue of a pass: entries 0..live-1 at positions 0..live-1 (segments filled in order).
__global__ void k_init_queue(uint32_t* queue, uint32_t live, uint32_t seg_cap) {
  const uint32_t s = threadIdx.x;
  if (s < static_cast<uint32_t>(kSegments)) {
    const uint64_t start = static_cast<uint64_t>(s) * seg_cap;
    const uint64_t left = live > start ? live - start : 0u;
    queue[kQueueInCount + s] = static_cast<uint32_t>(left < seg_cap ? left : seg_cap);
    queue[kQueueOutCursor + s * kSegLine] = 0u;
    queue[kQueueShadowCursor + s * kSegLine] = 0u;
  }
  if (s < static_cast<uint32_t>(kDoneStripes)) queue[kQueueDoneStripe + s * kSegLine] = 0u;
  if (s == 0u) {
    for (int r = 0; r < kRegions; ++r) queue[kQueueChunkCursor + r * kSegLine] = 0u;
    queue[kQueueWavesDone] = 0u;
  }
}

// Between bounces: the appended entries become the next input.
__global__ void k_advance(uint32_t* queue) {
  const uint32_t s = threadIdx.x;
  if (s < static_cast<uint32_t>(kSegments)) {
    queue[kQueueInCount + s] = queue[kQueueOutCursor + s * kSegLine];
    queue[kQueueOutCursor + s * kSegLine] = 0u;
    queue[kQueueShadowCursor + s * kSegLine] = 0u;
  }
}

__global__ void k_zero_counters(unsigned long long* c) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < static_cast<uint32_t>(kCounterWords)) c[i] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_clear_film(DevFilm film, uint32_t pixels) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < pixels) {
    film.sum[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    film.count[i] = 0u;
  }
}

// ------------------------------------------------------------------ dynamic-instance refit
// CpuBvhAccelerator::apply_instance_transform_update (CpuBvhAccelerator.cpp:310-358): the updated
// instances' triangles are recomputed from local geometry (refit_instance_primitives, :527-584), then the
// node boxes bottom-up, one launch per tree height (refit_node, :586-603). The tree's topology and leaf
// order stay; every box is the union of padded primitive boxes, so it equals a build's box.

// rotate_quat / transform_instance_vertex / normalize_vec3 (CpuBvhAccelerator.cpp:48-82), evaluated as
// the host builder evaluates them.
__device__ inline V3 refit_vertex(V3 local, const RefitInst& r) {
  const V3 v = mk(local.x * r.s[0], local.y * r.s[1], local.z * r.s[2]);
  V3 o = v;
  const float len_sq = r.q[0] * r.q[0] + r.q[1] * r.q[1] + r.q[2] * r.q[2] + r.q[3] * r.q[3];
  if (!(len_sq <= kEps * kEps)) {
    const float inv_len = 1.0f / __builtin_sqrtf(len_sq);
    const V3 qv = mk(r.q[0] * inv_len, r.q[1] * inv_len, r.q[2] * inv_len);
    const float qw = r.q[3] * inv_len;
    const V3 t = muls(cross(qv, v), 2.0f);
    o = add(add(v, muls(t, qw)), cross(qv, t));
  }
  return add(o, mk(r.t[0], r.t[1], r.t[2]));
}

__device__ inline V3 refit_normal(V3 v) {
  const float len_sq = dot(v, v);
  if (len_sq <= kEps * kEps) return mk(0.0f, 1.0f, 0.0f);
  return muls(v, 1.0f / __builtin_sqrtf(len_sq));
}

__device__ inline float first_min3(float a, float b, float c) {  // std::min({a, b, c})
  float m = a;
  if (b < m) m = b;
  if (c < m) m = c;
  return m;
}
__device__ inline float first_max3(float a, float b, float c) {  // std::max({a, b, c})
  float m = a;
  if (m < b) m = b;
  if (m < c) m = c;
  return m;
}

__global__ __launch_bounds__(kBlock) void k_refit_tris(const RefitInst* insts, uint32_t inst_count,
                                                       uint32_t tri_count, const float* lv, const uint32_t* li,
                                                       const uint32_t* leaf_pos, float4* tris, float4* prims,
                                                       float* tri_box) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= tri_count) return;
  uint32_t lo = 0, hi = inst_count;  // last instance with tri_begin <= g
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (insts[mid].tri_begin <= g) lo = mid;
    else hi = mid;
  }
  const RefitInst& r = insts[lo];
  const uint32_t tri = g - r.tri_begin;
  const uint32_t base = r.local_first_index + 3u * tri;
  V3 v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = refit_vertex(ld3(lv + 3u * (r.local_first_vertex + li[base + k])), r);
  const V3 e1 = sub(v[1], v[0]), e2 = sub(v[2], v[0]);
  const V3 n = refit_normal(cross(e1, e2));
  const uint32_t pid = r.prim_first + tri;
  const uint32_t lp = leaf_pos[pid];
  tris[3u * lp] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(pid));
  tris[3u * lp + 1u] = make_float4(e1.x, e1.y, e1.z, 0.0f);
  tris[3u * lp + 2u] = make_float4(e2.x, e2.y, e2.z, 0.0f);
  const float4 old = prims[pid];
  prims[pid] = make_float4(n.x, n.y, n.z, old.w);
  float* b = tri_box + 6u * lp;
  b[0] = first_min3(v[0].x, v[1].x, v[2].x) - kEps;
  b[1] = first_min3(v[0].y, v[1].y, v[2].y) - kEps;
  b[2] = first_min3(v[0].z, v[1].z, v[2].z) - kEps;
  b[3] = first_max3(v[0].x, v[1].x, v[2].x) + kEps;
  b[4] = first_max3(v[0].y, v[1].y, v[2].y) + kEps;
  b[5] = first_max3(v[0].z, v[1].z, v[2].z) + kEps;
}

// The box of child reference `ref`: a leaf's is the union of its triangles' padded boxes (in leaf order,
// starting from the first as refit_node does), an inner node's the union of its two child boxes.
__device__ inline void refit_child_box(const float4* nodes, int32_t ref, const float* tri_box, float* mn,
                                       float* mx) {
  if (ref < 0) {
    const uint32_t code = ~static_cast<uint32_t>(ref);
    const uint32_t first = code >> 5, count = code & 31u;
    const float* b = tri_box + 6u * first;
    for (int k = 0; k < 3; ++k) {
      mn[k] = b[k];
      mx[k] = b[3 + k];
    }
    for (uint32_t i = 1; i < count; ++i) {
      b += 6;
      for (int k = 0; k < 3; ++k) {
        mn[k] = smin(mn[k], b[k]);
        mx[k] = smax(mx[k], b[3 + k]);
      }
    }
    return;
  }
  const float4 f0 = nodes[4u * ref], f1 = nodes[4u * ref + 1u], f2 = nodes[4u * ref + 2u];
  const float c0mn[3] = {f0.x, f0.y, f0.z}, c0mx[3] = {f0.w, f1.x, f1.y};
  const float c1mn[3] = {f1.z, f1.w, f2.x}, c1mx[3] = {f2.y, f2.z, f2.w};
  for (int k = 0; k < 3; ++k) {
    mn[k] = smin(c0mn[k], c1mn[k]);
    mx[k] = smax(c0mx[k], c1mx[k]);
  }
}

__global__ __launch_bounds__(kBlock) void k_refit_nodes(float4* nodes, const uint32_t* order, uint32_t begin,
                                                        uint32_t count, const float* tri_box, uint32_t* nonfinite) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  const uint32_t ni = order[begin + i];
  const int4 refs = reinterpret_cast<const int4*>(nodes)[4u * ni + 3u];
  float a0[3], b0[3], a1[3], b1[3];
  refit_child_box(nodes, refs.x, tri_box, a0, b0);
  refit_child_box(nodes, refs.y, tri_box, a1, b1);
  nodes[4u * ni] = make_float4(a0[0], a0[1], a0[2], b0[0]);
  nodes[4u * ni + 1u] = make_float4(b0[1], b0[2], a1[0], a1[1]);
  nodes[4u * ni + 2u] = make_float4(a1[2], b1[0], b1[1], b1[2]);
  bool fin = true;
  for (int k = 0; k < 3; ++k) fin = fin && finite(a0[k]) && finite(b0[k]) && finite(a1[k]) && finite(b1[k]);
  if (!fin) atomicOr(nonfinite, 1u);
}

__global__ __launch_bounds__(kBlock) void k_hot_refresh(float4* nodes, const uint32_t* src, uint32_t hot_base,
                                                        uint32_t count) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= 3u * count) return;
  const uint32_t h = i / 3u, k = i % 3u;
  nodes[4u * (hot_base + h) + k] = nodes[4u * src[h] + k];
}

__global__ void k_refit_root(const float4* nodes, int32_t root_ref, const float* tri_box, SceneRec* rec) {
  if (threadIdx.x != 0) return;
  float mn[3], mx[3];
  refit_child_box(nodes, root_ref, tri_box, mn, mx);
  for (int k = 0; k < 3; ++k) {
    rec->root_box[k] = mn[k];
    rec->root_box[3 + k] = mx[k];
  }
}

// FilmBuffer::resolve_hdr (FilmBuffer.cpp:160-214) and ApplyFilmResolve (FilmBuffer.cpp:255-332), one
// thread per pixel: the mean of the pixel's samples, then exposure, white balance, tone map, gamma,
// clamp and 8-bit quantisation in the reference's operation order (HBM-bound: 20 bytes read, 4 or 12
// written per pixel).
__device__ inline float resolve_tone(float x, uint32_t mode) {
  switch (mode) {
    case 1:
      return x / (1.0f + x);
    case 2: {
      const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f, W = 11.2f;
      const float fx = ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
      const float fw = ((W * (A * W + C * B) + D * E) / (W * (A * W + B) + D * F)) - E / F;
      return fx / fw;
    }
    case 3:
      return (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
    default:
      return x;
  }
}

// static_cast<uint8_t>(float) as x86-64 code evaluates it (truncation to int32, INT_MIN for NaN and
// out-of-range values, then the low byte), so unclamped output quantises as on the reference's host.
__device__ inline uint32_t resolve_byte(float v) {
  const int32_t i = (v > -2147483648.0f && v < 2147483648.0f) ? static_cast<int32_t>(v) : INT32_MIN;
  return static_cast<uint32_t>(i) & 0xFFu;
}

template <bool kLdr>
__global__ __launch_bounds__(kBlock) void k_resolve(DevFilm film, uint32_t pixels, DevResolve rs, float* hdr,
                                                   uint32_t* ldr) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= pixels) return;
  const float4 s = film.sum[i];
  const uint32_t c = film.count[i];
  const float inv = 1.0f / static_cast<float>(c > 1u ? c : 1u);
  const float m[3] = {s.x * inv, s.y * inv, s.z * inv};
  if constexpr (!kLdr) {
    hdr[3u * i] = m[0];
    hdr[3u * i + 1u] = m[1];
    hdr[3u * i + 2u] = m[2];
  } else {
    uint32_t packed = 0xFF000000u;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float v = m[k] * rs.exposure * rs.wb[k];
      if (!isfinite(v)) v = 0.0f;
      v = resolve_tone(v, rs.tone_map);
      if (rs.gamma_out) v = ptm::powf_p(smax(0.0f, v), rs.inv_gamma);
      if (rs.clamp_out) v = smin(1.0f, smax(0.0f, v));
      packed |= resolve_byte(v * 255.0f + 0.5f) << (8 * k);
    }
    ldr[i] = packed;
  }
}

// camera_rays + shade_pixel seeding, ScalarCpuPathTracer.Render.cpp:55-119, :243-262.
// The adjoint variant also loads the path's loss adjoint from the film gradient.
template <bool kAdj>
__global__ __launch_bounds__(kBlock) void k_raygen(DevPaths P, DevRays out, DevScene sc, DevSettings st,
                                                  const DevTile* tiles, uint32_t tile_count, uint32_t path_count,
                                                  DevAdjoint A) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= path_count) return;
  uint32_t lo = 0, hi = tile_count;  // last tile with path_begin <= p: the last tile of p's group
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tiles[mid].path_begin <= p) lo = mid; else hi = mid;
  }
  const uint32_t G = tiles[lo].group >> 16;
  const uint32_t off = p - tiles[lo].path_begin;
  const DevTile tile = tiles[lo - (G - 1u) + off % G];
  const SceneRec& rec = sc.recs[tile.scene];
  const DevCamera& cam = rec.cam;
  const uint32_t width = rec.width, height = rec.height;
  const uint32_t local = off / G;
  const uint32_t x = tile.x + local % tile.w;
  const uint32_t y = tile.y + local / tile.w;
  const uint32_t sample = tile.sample;

  const uint64_t frame = st.frame_dev ? *st.frame_dev : st.frame;
  const uint64_t sample_seed = (static_cast<uint64_t>(y) << 32) | x;
  const uint64_t path_id = sample_seed + static_cast<uint64_t>(sample) + frame;
  uint64_t crng = rng_seed(st.seed, frame, sample, static_cast<uint32_t>(path_id),
                           static_cast<uint64_t>(y) * width + x);
  const float fx = (static_cast<float>(x) + next01(crng)) / static_cast<float>(width > 1u ? width : 1u);
  const float fy = (static_cast<float>(y) + next01(crng)) / static_cast<float>(height > 1u ? height : 1u);
  const uint64_t seed_mix = static_cast<uint64_t>(next01(crng) * static_cast<float>(0xffffu));
  const float nx = (2.0f * fx - 1.0f) * cam.aspect * cam.tan_half_fov;
  const float ny = (1.0f - 2.0f * fy) * cam.tan_half_fov;
  const V3 F = ld3(cam.fwd), R = ld3(cam.right), U = ld3(cam.up), C = ld3(cam.pos);
  const V3 dir = normalize(add(add(F, muls(R, nx)), muls(U, ny)));
  V3 ro = C, rd = dir;
  if (cam.aperture > 0.0f && cam.focus > kEps) {
    const float lens_radius_sample = __builtin_sqrtf(next01(crng));
    const float lens_phi = 2.0f * kPi * next01(crng);
    float boundary = 1.0f;
    if (cam.iris_blades >= 3u && cam.iris_roundness < 0.999f) {
      const float sector = 2.0f * kPi / static_cast<float>(cam.iris_blades);
      const float local_phi = lens_phi - cam.iris_rotation;
      const float wrapped = local_phi - sector * __builtin_floorf(local_phi / sector);
      const float centered = wrapped > sector * 0.5f ? wrapped - sector : wrapped;
      const float polygon = ptm::cosf_p(sector * 0.5f) / smax(0.1f, ptm::cosf_p(centered));
      boundary = polygon * (1.0f - cam.iris_roundness) + cam.iris_roundness;
    }
    const float lens_r = cam.aperture * lens_radius_sample * boundary;
    float sp, cp;
    ptm::sincosf_p(lens_phi, sp, cp);
    const V3 offset = add(muls(R, lens_r * cp * cam.anamorphic), muls(U, lens_r * sp));
    const V3 focus_point = add(C, muls(dir, cam.focus));
    ro = add(C, offset);
    rd = normalize(sub(focus_point, ro));
  }
  const uint64_t prng = rng_seed(st.seed + sample_seed + seed_mix, frame, sample, path_id, sample_seed);
  out.c[0][p] = ro.x; out.c[1][p] = ro.y; out.c[2][p] = ro.z;
  out.c[3][p] = rd.x; out.c[4][p] = rd.y; out.c[5][p] = rd.z;
  out.qscene[p] = tile.scene;
  // The initial queue is the identity (entry p serves path p, so its path slot is not stored).
  // Throughput 1 and radiance 0 are not stored either: the first bounce's shade starts from them. Settings keep max_depth >= 1, so every forward
  // path ends in some bounce's shade, which writes its radiance to its slot.
  out.rng[p] = prng;
  if constexpr (kAdj) {
    // the replay's items (buffer 0, beside the initial queue): the forward pass's radiance of the path,
    // which it left in the path's slot, and the loss adjoint of its pixel
    const size_t pix = rec.film_base + static_cast<size_t>(y) * rec.width + x;
    float* f = A.items[0].f;
    const size_t S = A.stride;
    for (int c = 0; c < 3; ++c) {
      f[(kItemLtot + c) * S + p] = P.rad[c][p];
      f[(kItemDelta + c) * S + p] = A.film_grad[3 * pix + c];
    }
    A.items[0].u[kItemFlags * S + p] = 0u;
  }
}

// ------------------------------------------------------------------ adjoint gradient accumulation
// Gradient slots: albedo 3*nmat, emission 3*nmat, light colour 3*nlight, roughness nmat. Staged in
// LDS per block when they fit, else added straight to the global fp32 buffers.
enum GradKind { kGradAlbedo = 0, kGradEmission = 1, kGradLight = 2, kGradRough = 3 };

// LDS staging keeps lds_copies copies of the slots, lds_stride floats apart (an odd stride puts the
// copies of a slot in different banks); lane l adds into copy l mod lds_copies, so the lanes of a wave
// that hit the same material (neighbouring pixels usually do) spread over the copies instead of
// serialising on one address.
__device__ __forceinline__ float* lds_copy(const DevAdjoint& A, float* lds) {
  return lds + (threadIdx.x & (A.lds_copies - 1u)) * A.lds_stride;
}

// The staged add is an LDS atomic and the direct one a global atomic: each is issued through a pointer
// of its own address space, so the two branches cannot fold into one generic (flat) atomic.
__device__ __forceinline__ void lds_fadd(float* p, float v) {
  __hip_atomic_fetch_add((PT_LDS float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void global_fadd(float* p, float v) {
  __hip_atomic_fetch_add((PT_GLOBAL float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void gadd(const DevAdjoint& A, float* lds, GradKind kind, uint32_t idx, int c, float v) {
  if (v == 0.0f) return;
  if (kind == kGradRough) {
    if (!A.g_rough) return;
    if (A.use_lds) lds_fadd(lds_copy(A, lds) + 6u * A.nmat + 3u * A.nlight + idx, v);
    else global_fadd(A.g_rough + idx, v);
    return;
  }
  const uint32_t slot = 3u * idx + static_cast<uint32_t>(c);
  if (A.use_lds) {
    const uint32_t off = kind == kGradAlbedo ? 0u : (kind == kGradEmission ? 3u * A.nmat : 6u * A.nmat);
    lds_fadd(lds_copy(A, lds) + off + slot, v);
  } else {
    float* g = kind == kGradAlbedo ? A.g_albedo : (kind == kGradEmission ? A.g_emission : A.g_light);
    global_fadd(g + slot, v);
  }
}

__device__ __forceinline__ void grad_lds_clear(const DevAdjoint& A, float* lds) {
  if (!A.use_lds) return;
  const uint32_t n = A.lds_copies * A.lds_stride;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = 0.0f;
  __syncthreads();
}

// Texel gradients through a per-wave LDS table. A wave's lanes usually hold samples of one pixel, so
// their bilinear taps add into the same few texels; adding them straight to global memory serialises
// those atomics in L2. Each wave owns kTexSlots entries {texel, r, g, b} after the gradient slots of
// the block's LDS array: a tap claims the entry its texel hashes to (or finds it already holding that
// texel) and adds there with LDS atomics; a tap whose entry holds another texel adds to global memory
// directly. tex_grad_flush, run by the converged wave, moves every claimed entry to global memory
// (one atomic per channel) and frees it. The sums are those of the direct adds, in another order.
constexpr uint32_t kTexSlots = 64;
constexpr uint32_t kTexEmpty = 0xFFFFFFFFu;
constexpr int kTexTableFloats = kWavesPerBlock * static_cast<int>(kTexSlots) * 4;

__device__ __forceinline__ float* tex_table(float* lds) {
  return lds + kAdjLdsFloats + (threadIdx.x >> 6) * (kTexSlots * 4u);
}

__device__ __forceinline__ void tex_table_init(float* lds) {
  float* e = tex_table(lds) + 4u * (threadIdx.x & 63u);
  e[0] = __uint_as_float(kTexEmpty);
  e[1] = e[2] = e[3] = 0.0f;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
}

__device__ __forceinline__ void tex_grad_add(const DevAdjoint& A, float* lds, uint32_t texel, const float v[3]) {
  float* e = tex_table(lds) + 4u * ((texel * 2654435761u) >> 26);  // 64 entries
  uint32_t held = kTexEmpty;
  const bool claimed = __hip_atomic_compare_exchange_strong((PT_LDS uint32_t*)e, &held, texel, __ATOMIC_RELAXED,
                                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (claimed || held == texel) {
    for (int c = 0; c < 3; ++c)
      if (v[c] != 0.0f) lds_fadd(e + 1 + c, v[c]);
  } else {
    for (int c = 0; c < 3; ++c)
      if (v[c] != 0.0f) global_fadd(A.g_tex + 3u * texel + static_cast<uint32_t>(c), v[c]);
  }
}

// Called by every lane of the wave at a point where the wave is converged.
__device__ __forceinline__ void tex_grad_flush(const DevAdjoint& A, float* lds) {
  if (!A.g_tex) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");  // the wave's table adds are complete
  float* e = tex_table(lds) + 4u * (threadIdx.x & 63u);
  const uint32_t texel = __float_as_uint(e[0]);
  if (texel != kTexEmpty) {
    for (int c = 0; c < 3; ++c) {
      const float v = e[1 + c];
      if (v != 0.0f) global_fadd(A.g_tex + 3u * texel + static_cast<uint32_t>(c), v);
      e[1 + c] = 0.0f;
    }
    e[0] = __uint_as_float(kTexEmpty);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
}

__device__ __forceinline__ void grad_lds_flush(const DevAdjoint& A, const float* lds) {
  if (!A.use_lds) return;
  __syncthreads();
  const uint32_t n = 7u * A.nmat + 3u * A.nlight;
  float* row = A.partial + static_cast<size_t>(blockIdx.x) * n;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    float v = 0.0f;
    for (uint32_t k = 0; k < A.lds_copies; ++k) v += lds[k * A.lds_stride + i];
    if (v != 0.0f) row[i] += v;
  }
}

__global__ __launch_bounds__(kBlock) void k_grad_clear(DevAdjoint A, uint32_t count) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < count) A.partial[i] = 0.0f;
}

// Thread (slot i, group y) sums rows y, y + 16, y + 32, ... of slot i and adds the sum to the
// caller's buffer (16 atomic adds per slot).
constexpr uint32_t kReduceGroups = 16;
__global__ __launch_bounds__(kBlock) void k_grad_reduce(DevAdjoint A) {
  const uint32_t n = 7u * A.nmat + 3u * A.nlight, na = 3u * A.nmat, nl = 3u * A.nlight;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float v = 0.0f;
  for (uint32_t r = blockIdx.y; r < A.partial_rows; r += kReduceGroups) v += A.partial[static_cast<size_t>(r) * n + i];
  if (v == 0.0f) return;
  if (i < na) atomicAdd(A.g_albedo + i, v);
  else if (i < 2u * na) atomicAdd(A.g_emission + (i - na), v);
  else if (i < 2u * na + nl) atomicAdd(A.g_light + (i - 2u * na), v);
  else if (A.g_rough) atomicAdd(A.g_rough + (i - 2u * na - nl), v);
}

// Copies the hot nodes into the workgroup's LDS array (kHotNodes records); the caller's next barrier
// (load_prefix's) publishes them.
__device__ __forceinline__ void load_hot(const DevScene& sc, float4* s_hot) {
  for (uint32_t i = threadIdx.x; i < 4u * sc.hot_count; i += kBlock) s_hot[i] = sc.nodes[4u * sc.hot_base + i];
}

__device__ __forceinline__ Stack make_stack(const DevPaths& P, const DevScene& sc, int32_t* s_ref, float* s_t,
                                            const float4* s_hot) {
  Stack s;
  s.hot = (const PT_LDS float*)s_hot;
  s.hot_base = sc.hot_base;
  s.hot_count = sc.hot_count;
  s.lds_ref = (PT_LDS int32_t*)s_ref;
  s.lds_t = (PT_LDS float*)s_t;
  s.ovf_ref = (PT_GLOBAL int32_t*)P.stack_ref;
  s.ovf_t = (PT_GLOBAL float*)P.stack_t;
  s.ovf_depth = P.stack_overflow_depth;
  s.ovf_stride = P.stack_threads;
  s.lane_slot = threadIdx.x;
  s.gid = blockIdx.x * kBlock + threadIdx.x;
  s.error = P.error;
  return s;
}

// Persistent closest-hit kernel over the bounce's segmented ray queue.
// ------------------------------------------------------------------ SDF primitives
// ScalarCpuPathTracer's analytic / sphere-traced SDF intersectors (ScalarCpuPathTracer.cpp:809-962) and
// the SDF loop of intersect_scene (ScalarCpuPathTracer.Intersect.cpp:91-129), in the reference's float
// expressions. Euler rotations use the primitive's precomputed sines and cosines.
constexpr float kMinMarchStep = 1.0e-3f;     // ScalarCpuPathTracer.cpp:17-19
constexpr uint32_t kMaxMarchSteps = 192u;
constexpr float kMaxMarchDistance = 10000.0f;

__device__ __forceinline__ V3 rot_x(V3 v, float s, float c) { return mk(v.x, c * v.y - s * v.z, s * v.y + c * v.z); }
__device__ __forceinline__ V3 rot_y(V3 v, float s, float c) { return mk(c * v.x + s * v.z, v.y, -s * v.x + c * v.z); }
__device__ __forceinline__ V3 rot_z(V3 v, float s, float c) { return mk(c * v.x - s * v.y, s * v.x + c * v.y, v.z); }
__device__ __forceinline__ V3 euler_inv(V3 v, const DevSdf& p) {  // rotate_euler_inv, :131-133
  return rot_z(rot_y(rot_x(v, p.inv_s[0], p.inv_c[0]), p.inv_s[1], p.inv_c[1]), p.inv_s[2], p.inv_c[2]);
}
__device__ __forceinline__ V3 euler_fwd(V3 v, const DevSdf& p) {  // rotate_euler, :135-137
  return rot_x(rot_y(rot_z(v, p.fwd_s[2], p.fwd_c[2]), p.fwd_s[1], p.fwd_c[1]), p.fwd_s[0], p.fwd_c[0]);
}
__device__ __forceinline__ V3 div_scale(V3 v, const float* s) {  // divide_by_scale, :154-159
  return mk(v.x / (s[0] != 0.0f ? s[0] : 1.0f), v.y / (s[1] != 0.0f ? s[1] : 1.0f), v.z / (s[2] != 0.0f ? s[2] : 1.0f));
}
__device__ __forceinline__ float vlen(V3 v) { return __builtin_sqrtf(dot(v, v)); }

__device__ float sdf_distance(const DevSdf& p, V3 w) {  // sdf_distance, :809-846
  const V3 l = div_scale(euler_inv(sub(w, mk(p.pos[0], p.pos[1], p.pos[2])), p), p.scale);
  switch (p.shape) {
    case 1u: {  // box
      const V3 q = mk(__builtin_fabsf(l.x) - p.scale[0], __builtin_fabsf(l.y) - p.scale[1], __builtin_fabsf(l.z) - p.scale[2]);
      const V3 mq = mk(smax(0.0f, q.x), smax(0.0f, q.y), smax(0.0f, q.z));
      return vlen(mq) + smin(smax(q.x, smax(q.y, q.z)), 0.0f);
    }
    case 2u: {  // rounded box
      const float r = p.pa;
      const V3 q = mk(__builtin_fabsf(l.x) - p.scale[0] + r, __builtin_fabsf(l.y) - p.scale[1] + r,
                      __builtin_fabsf(l.z) - p.scale[2] + r);
      const V3 mq = mk(smax(0.0f, q.x), smax(0.0f, q.y), smax(0.0f, q.z));
      return vlen(mq) + smin(smax(q.x, smax(q.y, q.z)), 0.0f) - p.pa;
    }
    case 3u:  // plane
      return l.y + p.pb;
    case 4u: {  // torus
      const float a = smax(0.05f, p.pa);
      const float b = smax(0.02f, p.radius);
      const float qx = vlen(mk(l.x, 0.0f, l.z)) - a;
      return vlen(mk(qx, l.y, 0.0f)) - b;
    }
    case 5u: {  // capsule
      const V3 a = mk(0.0f, -p.pa, 0.0f);
      const V3 b = mk(0.0f, p.pa, 0.0f);
      const V3 ab = sub(b, a);
      const float t = dot(sub(l, a), ab) / smax(kEps, dot(ab, ab));
      const float c = smax(0.0f, smin(1.0f, t));
      const V3 closest = add(a, muls(ab, c));
      return vlen(sub(l, closest)) - p.radius;
    }
    default:  // sphere
      return vlen(l) - p.radius;
  }
}

__device__ V3 sdf_normal(const DevSdf& p, V3 w) {  // sdf_normal, :848-854
  const float e = 0.0005f;
  const float dx = sdf_distance(p, add(w, mk(e, 0.0f, 0.0f))) - sdf_distance(p, sub(w, mk(e, 0.0f, 0.0f)));
  const float dy = sdf_distance(p, add(w, mk(0.0f, e, 0.0f))) - sdf_distance(p, sub(w, mk(0.0f, e, 0.0f)));
  const float dz = sdf_distance(p, add(w, mk(0.0f, 0.0f, e))) - sdf_distance(p, sub(w, mk(0.0f, 0.0f, e)));
  return normalize(muls(mk(dx, dy, dz), 0.5f));
}

struct SdfStats {
  uint32_t tests, steps, hits;
};

// One primitive's test (intersect_sphere :855-903, intersect_box and its aliases :905-962): true with
// the hit's t.
__device__ bool sdf_test(const DevSdf& p, V3 o, V3 d, float& t_out, SdfStats& st) {
  ++st.tests;
  if (p.shape == 0u) {
    const float radius = smax(0.01f, p.radius);
    const V3 lo = div_scale(euler_inv(sub(o, mk(p.pos[0], p.pos[1], p.pos[2])), p), p.scale);
    const V3 ld = div_scale(euler_inv(d, p), p.scale);
    const float a = dot(ld, ld);
    ++st.steps;
    if (a <= kEps * kEps) return false;
    const float hb = dot(lo, ld);
    const float c = dot(lo, lo) - radius * radius;
    const float disc = hb * hb - a * c;
    if (disc < 0.0f) return false;
    const float root = __builtin_sqrtf(disc);
    float cand = (-hb - root) / a;
    if (cand <= kEps) cand = (-hb + root) / a;
    if (cand <= kEps || !finite(cand)) return false;
    t_out = cand;
    ++st.hits;
    return true;
  }
  float t = 0.0f;
  for (uint32_t i = 0; i < kMaxMarchSteps; ++i) {
    ++st.steps;
    const float dd = sdf_distance(p, add(o, muls(d, t)));
    if (dd <= kEps) {
      t_out = t;
      ++st.hits;
      return true;
    }
    t += smax(kMinMarchStep, dd);
    if (t > kMaxMarchDistance) return false;
  }
  return false;
}

// The surface normal intersect_scene reports for an SDF hit at t (the sphere's analytic normal, the
// others' central differences at the hit point), normalised once more as intersect_scene does.
__device__ V3 sdf_hit_normal(const DevSdf& p, V3 o, V3 d, float t) {
  V3 n;
  if (p.shape == 0u) {
    const V3 lo = div_scale(euler_inv(sub(o, mk(p.pos[0], p.pos[1], p.pos[2])), p), p.scale);
    const V3 ld = div_scale(euler_inv(d, p), p.scale);
    const V3 lhit = add(lo, muls(ld, t));
    n = normalize(euler_fwd(div_scale(normalize(lhit), p.scale), p));
  } else {
    n = sdf_normal(p, add(o, muls(d, t)));
  }
  return normalize(n);
}

// The SDF loop of intersect_scene: the nearest SDF hit below best_t (index into the pool, or -1).
__device__ int32_t sdf_closest(const DevScene& sc, const SceneRec& R, V3 o, V3 d, float& best_t, SdfStats& st) {
  int32_t best = -1;
  for (uint32_t k = 0; k < R.sdf_count; ++k) {
    float t = 0.0f;
    if (sdf_test(sc.sdfs[R.sdf_base + k], o, d, t, st) && !(t >= best_t)) {
      best_t = t;
      best = static_cast<int32_t>(R.sdf_base + k);
    }
  }
  return best;
}

__device__ __forceinline__ void flush_sdf(unsigned long long* counters, const SdfStats& st) {
  wave_add(counters, kSdfTests, st.tests);
  wave_add(counters, kSdfSteps, st.steps);
  wave_add(counters, kSdfHits, st.hits);
}

// After k_extend, when some scene has SDF primitives: each extension ray's triangle hit is replaced
// by a nearer SDF hit, encoded as hit_prim = -2 - (pool SDF index).
__global__ __launch_bounds__(kBlock) void k_sdf_extend(DevPaths P, DevRays in, DevScene sc) {
  __shared__ uint32_t prefix[kSegments + 1];
  load_prefix(P.queue + kQueueInCount, 1, prefix);
  const uint32_t n = prefix[kSegments];
  SdfStats st = {0, 0, 0};
  for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
    const uint32_t pos = queue_pos(prefix, j, P.seg_cap);
    const SceneRec& R = sc.recs[in.qscene[pos]];
    if (R.sdf_count == 0u) continue;
    const V3 o = mk(in.c[0][pos], in.c[1][pos], in.c[2][pos]);
    const V3 d = mk(in.c[3][pos], in.c[4][pos], in.c[5][pos]);
    float bt = P.hit_prim[pos] >= 0 ? P.hit_t[pos] : kInf;
    const int32_t k = sdf_closest(sc, R, o, d, bt, st);
    if (k >= 0) {
      P.hit_t[pos] = bt;
      P.hit_prim[pos] = -2 - k;
    }
  }
  flush_sdf(P.counters, st);
}

__global__ __launch_bounds__(kBlock) void k_extend(DevPaths P, DevRays in, DevScene sc) {
  __shared__ int32_t s_ref[kStackLds * kBlock];
  __shared__ float s_t[kStackLds * kBlock];
  __shared__ float4 s_hot[4 * (kHotNodes > 0 ? kHotNodes : 1)];
  load_hot(sc, s_hot);
  __shared__ uint32_t prefix[kSegments + 1];
  load_prefix(P.queue + kQueueInCount, 1, prefix);
  const Stack stk = make_stack(P, sc, s_ref, s_t, s_hot);
  const uint32_t n = prefix[kSegments];
  TravStats st = {0, 0, 0, 0, 0};
  uint32_t rays = 0;
  uint32_t pos = 0;
  auto begin = [&](uint32_t j, Trav& s) {
    pos = queue_pos(prefix, j, P.seg_cap);
    const V3 o = mk(in.c[0][pos], in.c[1][pos], in.c[2][pos]);
    const V3 d = mk(in.c[3][pos], in.c[4][pos], in.c[5][pos]);
    return trav_begin(sc.recs[in.qscene[pos]], s, o, d, kInf, st);
  };
  auto end = [&](uint32_t, Trav& s) {
    P.hit_t[pos] = s.best_t;
    P.hit_prim[pos] = s.best_prim;
    ++rays;
    return false;
  };
  drive_queries<false>(sc, stk, P.queue, n, P.refill_min, st, begin, end);
  flush_stats(P.counters, st, false);
  wave_add(P.counters, kRays, rays);
}

// The pool as seen by one path of scene sid: that scene's lights and environment.
__device__ __forceinline__ DevScene scene_view(const DevScene& pool, uint32_t sid) {
  const SceneRec& R = pool.recs[sid];
  DevScene v = pool;
  v.lights = pool.lights + R.light_base;
  v.light_count = R.light_count;
  v.has_transmissive = R.has_transmissive;
  for (int k = 0; k < 3; ++k) {
    v.env[k] = R.env[k];
    v.env_scale[k] = R.env_scale[k];
  }
  v.env_map = (R.env_w && R.env_h) ? pool.env_map + R.env_base : nullptr;
  v.env_w = R.env_w;
  v.env_h = R.env_h;
  return v;
}

// ------------------------------------------------------------------ textures

// Texture coordinates at barycentrics (u, v) of a primitive with vertex coordinates uvp[0..5].
__device__ __forceinline__ float2 uv_at(const float* uvp, float u, float v) {
  const float w = 1.0f - u - v;
  return make_float2(uvp[0] * w + uvp[2] * u + uvp[4] * v, uvp[1] * w + uvp[3] * u + uvp[5] * v);
}

// A bilinear tap (repeat addressing, texel centres at half-integers): its four texels (texel
// indices into the pool, float offset / 3), their weights and the filtered colour.
struct TexTap {
  uint32_t tex;
  uint32_t idx[4];
  float w[4];
  V3 color;
};

__device__ __forceinline__ TexTap tex_tap(const DevScene& sc, uint32_t ti, float2 uv) {
  TexTap r;
  r.tex = ti;
  const TexRec T = sc.textures[ti];
  float uf = uv.x - __builtin_floorf(uv.x), vf = uv.y - __builtin_floorf(uv.y);
  if (!(uf >= 0.0f && uf < 1.0f)) uf = 0.0f;  // NaN, or a huge coordinate rounding to 1
  if (!(vf >= 0.0f && vf < 1.0f)) vf = 0.0f;
  const float x = uf * static_cast<float>(T.width) - 0.5f, y = vf * static_cast<float>(T.height) - 0.5f;
  const float x0 = __builtin_floorf(x), y0 = __builtin_floorf(y);
  const float fx = x - x0, fy = y - y0;
  int ix0 = static_cast<int>(x0), iy0 = static_cast<int>(y0);
  if (ix0 < 0) ix0 += static_cast<int>(T.width);
  if (iy0 < 0) iy0 += static_cast<int>(T.height);
  const int ix1 = ix0 + 1 == static_cast<int>(T.width) ? 0 : ix0 + 1;
  const int iy1 = iy0 + 1 == static_cast<int>(T.height) ? 0 : iy0 + 1;
  const uint32_t base = T.offset;
  r.idx[0] = base + static_cast<uint32_t>(iy0) * T.width + static_cast<uint32_t>(ix0);
  r.idx[1] = base + static_cast<uint32_t>(iy0) * T.width + static_cast<uint32_t>(ix1);
  r.idx[2] = base + static_cast<uint32_t>(iy1) * T.width + static_cast<uint32_t>(ix0);
  r.idx[3] = base + static_cast<uint32_t>(iy1) * T.width + static_cast<uint32_t>(ix1);
  r.w[0] = (1.0f - fx) * (1.0f - fy);
  r.w[1] = fx * (1.0f - fy);
  r.w[2] = (1.0f - fx) * fy;
  r.w[3] = fx * fy;
  V3 c = mk(0.0f, 0.0f, 0.0f);
  for (int k = 0; k < 4; ++k) c = add(c, muls(ld3(sc.texels + 3u * r.idx[k]), r.w[k]));
  r.color = c;
  return r;
}

// Adds G = df/d(base colour) of a vertex on pool material m whose texture coordinate is uv: base =
// albedo * texture tap (tap colour 1 when untextured), so df/dalbedo_c = G_c * tap_c and
// df/dtexel_{j,c} = G_c * albedo_c * w_j for the tap's four texels (global fp32 atomics).
__device__ __forceinline__ void gadd_base(const DevAdjoint& A, const DevScene& sc, float* lds, uint32_t m, float2 uv,
                                          V3 G) {
  const PtMaterial& mat = sc.materials[m];
  const float g[3] = {G.x, G.y, G.z};
  if (mat.base_color_texture_index >= sc.texture_count) {
    for (int c = 0; c < 3; ++c) gadd(A, lds, kGradAlbedo, m, c, g[c]);
    return;
  }
  const TexTap tap = tex_tap(sc, mat.base_color_texture_index, uv);
  const float tc[3] = {tap.color.x, tap.color.y, tap.color.z};
  for (int c = 0; c < 3; ++c) gadd(A, lds, kGradAlbedo, m, c, g[c] * tc[c]);
  if (!A.g_tex) return;
  const float ga[3] = {g[0] * mat.albedo[0], g[1] * mat.albedo[1], g[2] * mat.albedo[2]};
  if (ga[0] == 0.0f && ga[1] == 0.0f && ga[2] == 0.0f) return;
  for (int j = 0; j < 4; ++j) {
    const float v[3] = {ga[0] * tap.w[j], ga[1] * tap.w[j], ga[2] * tap.w[j]};
    tex_grad_add(A, lds, tap.idx[j], v);
  }
}

// History slot h of the adjoint state. The arrays of all slots are laid out `stride` entries apart,
// so a per-lane slot index is address arithmetic, not a per-lane choice among kernel-argument pointers.
__device__ __forceinline__ float* hist_dx(const DevAdjoint& A, uint32_t h, int c) {
  return A.h_dx[0][0] + static_cast<size_t>(3u * h + static_cast<uint32_t>(c)) * A.stride;
}
__device__ __forceinline__ uint32_t* hist_mat(const DevAdjoint& A, uint32_t h) {
  return A.h_mat[0] + static_cast<size_t>(h) * A.stride;
}
__device__ __forceinline__ float* hist_uv(const DevAdjoint& A, uint32_t h, int k) {
  return A.h_uv[0][0] + static_cast<size_t>(2u * h + static_cast<uint32_t>(k)) * A.stride;
}

__device__ __forceinline__ V3 channel(int c, float v) {
  return mk(c == 0 ? v : 0.0f, c == 1 ? v : 0.0f, c == 2 ? v : 0.0f);
}

// A path's adjoint item, loaded from its queue entry (AdjItems) for one shade.
struct AdjItem {
  float ltot[3], delta[3], pdx[3], score, uv[2];
  uint32_t mat, flags;
};

__device__ __forceinline__ AdjItem load_item(const AdjItems& it, size_t stride, uint32_t q) {
  AdjItem a;
  for (int c = 0; c < 3; ++c) {
    a.ltot[c] = it.f[(kItemLtot + c) * stride + q];
    a.delta[c] = it.f[(kItemDelta + c) * stride + q];
    a.pdx[c] = it.f[(kItemPdx + c) * stride + q];
  }
  a.score = it.f[kItemScore * stride + q];
  a.uv[0] = it.f[kItemUv * stride + q];
  a.uv[1] = it.f[(kItemUv + 1) * stride + q];
  a.mat = it.u[kItemMat * stride + q];
  a.flags = it.u[kItemFlags * stride + q];
  return a;
}

__device__ __forceinline__ void store_item(const AdjItems& it, size_t stride, uint32_t q, const AdjItem& a) {
  for (int c = 0; c < 3; ++c) {
    it.f[(kItemLtot + c) * stride + q] = a.ltot[c];
    it.f[(kItemDelta + c) * stride + q] = a.delta[c];
    it.f[(kItemPdx + c) * stride + q] = a.pdx[c];
  }
  it.f[kItemScore * stride + q] = a.score;
  it.f[kItemUv * stride + q] = a.uv[0];
  it.f[(kItemUv + 1) * stride + q] = a.uv[1];
  it.u[kItemMat * stride + q] = a.mat;
  it.u[kItemFlags * stride + q] = a.flags;
}

// Applies the bounce of item `a` (path slot p) that waited for its remainder radiance R = L_total -
// L_so_far: df/dx_j = delta * R_j for its albedo / texels and roughness, plus the Russian-roulette terms
// of the history vertices through rr's channel while the roulette derivative is active.
__device__ __forceinline__ void adj_flush(const DevAdjoint& A, const DevScene& sc, float* lds, uint32_t p, V3 L,
                                          AdjItem& a) {
  uint32_t& flags = a.flags;
  if (!(flags & 1u)) return;
  const float* dl = a.delta;
  const float R[3] = {a.ltot[0] - L.x, a.ltot[1] - L.y, a.ltot[2] - L.z};
  const float S = dl[0] * R[0] + dl[1] * R[1] + dl[2] * R[2];
  const uint32_t m = a.mat;
  gadd_base(A, sc, lds, m, make_float2(a.uv[0], a.uv[1]),
            mk(dl[0] * R[0] * a.pdx[0], dl[1] * R[1] * a.pdx[1], dl[2] * R[2] * a.pdx[2]));
  gadd(A, lds, kGradRough, m, 0, S * a.score);
  if (flags & 2u) {
    const int cs = static_cast<int>((flags >> 2) & 3u);
    const uint32_t nh = (flags >> 4) & 7u;
    for (uint32_t h = 0; h < nh; ++h)
      gadd_base(A, sc, lds, hist_mat(A, h)[p], make_float2(hist_uv(A, h, 0)[p], hist_uv(A, h, 1)[p]),
                channel(cs, -S * hist_dx(A, h, cs)[p]));
    if (flags & 0x80u) {
      const int pc = static_cast<int>((flags >> 8) & 3u);
      const uint32_t pn = (flags >> 10) & 7u;
      for (uint32_t h = 0; h < pn; ++h)
        gadd_base(A, sc, lds, hist_mat(A, h)[p], make_float2(hist_uv(A, h, 0)[p], hist_uv(A, h, 1)[p]),
                  channel(pc, S * hist_dx(A, h, pc)[p]));
    }
    flags = (flags & ~0x1F80u) | 0x80u | (static_cast<uint32_t>(cs) << 8) | (nh << 10);
  }
  flags &= ~1u;
}

struct MatFlags {
  bool emissive_model, mirror, metallic, transmissive, diffuse, sheen, clearcoat, toon;
};

__device__ __forceinline__ const PtMaterial& material_at(const DevScene& sc, uint32_t idx) {
  const uint32_t last = sc.material_count - 1u;
  return sc.materials[idx < last ? idx : last];
}

// std::fmod(x, 2) for the integer-valued sums the effects pass (exact).
__device__ __forceinline__ float fmod2(float x) { return x - 2.0f * __builtin_truncf(x * 0.5f); }

// hash01 / smoothstep, ScalarCpuPathTracer.cpp:187-196
__device__ __forceinline__ float hash01(float x, float y, float z, float seed) {
  const float v = ptm::sinf_p(x * 12.9898f + y * 78.233f + z * 37.719f + seed * 19.19f) * 43758.5453f;
  return v - __builtin_floorf(v);
}
__device__ __forceinline__ float smoothstep_f(float e0, float e1, float x) {
  const float t = clamp01((x - e0) / smax(1.0e-6f, e1 - e0));
  return t * t * (3.0f - 2.0f * t);
}

// procedural_wood_albedo, ScalarCpuPathTracer.cpp:198-262 (linear in base: k receives the factor).
__device__ V3 wood_albedo(V3 base, V3 p, float roughness, float clearcoat, uint32_t style, float& k) {
  const float rough = clamp01(roughness), coat = clamp01(clearcoat);
  const float polish = clamp01(coat * (1.25f - rough * 0.75f));
  const float parquet = style == 17u ? 1.0f : 0.0f;
  const float walnut = style == 18u ? 1.0f : 0.0f;
  const float sandal = style == 19u ? 1.0f : 0.0f;
  const float plank_width = 0.40f + 0.16f * parquet + 0.08f * polish - 0.18f * walnut;
  const float plank_length = 1.15f + 0.85f * parquet + 0.55f * polish - 0.55f * walnut;
  const float px = (p.x + sandal * p.y * 0.35f) / smax(0.12f, plank_width);
  const float pz = (p.z + walnut * p.y * 0.70f + sandal * p.x * 0.22f) / smax(0.30f, plank_length);
  const float cell_x = __builtin_floorf(px), cell_z = __builtin_floorf(pz);
  const float local_x = px - cell_x, local_z = pz - cell_z;
  const float seed = hash01(cell_x, cell_z, 0.0f, 16.0f);
  const float macro_noise = hash01(__builtin_floorf(p.x * 3.7f), __builtin_floorf(p.y * 3.7f),
                                   __builtin_floorf(p.z * 3.7f), 31.0f + static_cast<float>(style));
  const float micro_noise = hash01(__builtin_floorf(p.x * 34.0f), __builtin_floorf(p.y * 34.0f),
                                   __builtin_floorf(p.z * 34.0f), 47.0f + static_cast<float>(style));
  const bool rotate = parquet > 0.5f && fmod2(cell_x + cell_z) >= 1.0f;
  const float along = rotate ? local_z : local_x, across = rotate ? local_x : local_z;
  const float grain_wobble = (macro_noise - 0.5f) * (0.06f + rough * 0.12f + walnut * 0.08f + sandal * 0.04f);
  const float warped_along = along + grain_wobble + ptm::sinf_p(across * 6.0f + seed * 5.0f) * 0.018f;
  const float grain_freq = 12.0f + rough * 8.0f + walnut * 22.0f + sandal * 36.0f;
  const float wave = ptm::sinf_p((warped_along * grain_freq + seed * 6.2831853f) +
                                 ptm::sinf_p(across * (6.0f + rough * 8.0f) + seed * 4.0f) *
                                     (0.25f + rough * 0.45f + walnut * 0.45f));
  const float fine = ptm::sinf_p(warped_along * (55.0f + rough * 55.0f + sandal * 110.0f) +
                                 across * (9.0f + walnut * 25.0f) + seed * 12.0f);
  const float fine_weight = 0.10f + rough * 0.18f + walnut * 0.18f + sandal * 0.22f;
  const float ring = smoothstep_f(-0.35f + polish * 0.18f, 0.82f, wave) * (1.0f - fine_weight) +
                     (0.5f + 0.5f * fine) * fine_weight;
  const float seam_x = smin(local_x, 1.0f - local_x), seam_z = smin(local_z, 1.0f - local_z);
  const float seam = (1.0f - sandal) * (1.0f - smoothstep_f(0.0f, 0.028f + polish * 0.018f, smin(seam_x, seam_z)));
  const float kd = 0.46f - rough * 0.10f - walnut * 0.18f;
  const float ka = 1.14f + coat * 0.22f + sandal * 0.18f;
  const V3 dark = muls(base, kd);
  const V3 amber = add(muls(base, ka), mk(0.08f + sandal * 0.18f, 0.055f + sandal * 0.06f, 0.025f));
  V3 wood = add(muls(dark, 1.0f - ring), muls(amber, ring));
  const float board_tint = 0.92f + (0.08f + walnut * 0.12f + sandal * 0.06f) * seed +
                           (macro_noise - 0.5f) * (0.08f + rough * 0.06f);
  const float pores = smoothstep_f(0.74f - walnut * 0.08f, 1.0f, micro_noise) *
                      (0.025f + rough * 0.055f + walnut * 0.055f + sandal * 0.018f);
  const float seam_f = 1.0f - seam * (0.10f + rough * 0.35f + parquet * 0.18f) * (1.0f - coat * 0.45f);
  wood = muls(wood, board_tint);
  wood = muls(wood, 1.0f - pores);
  wood = muls(wood, seam_f);
  const float wood_mix = 0.30f + rough * 0.34f - polish * 0.14f + walnut * 0.26f + sandal * 0.18f;
  k = (1.0f - wood_mix) + (kd * (1.0f - ring) + ka * ring) * board_tint * (1.0f - pores) * seam_f * wood_mix;
  return add(muls(base, 1.0f - wood_mix), muls(wood, wood_mix));
}

// apply_material_effect, ScalarCpuPathTracer.cpp:264-355. Affine in the albedo before the final
// [0, 1.5] clamp: k is d(colour)/d(albedo) there and pre the unclamped colour (for the adjoint).
__device__ V3 effect_albedo(const PtMaterial& m, V3 base, V3 pos, V3 n, V3 incoming, float& k, V3& pre) {
  V3 c = base;
  k = 1.0f;
  const uint32_t e = m.material_effect;
  if (e != 0u) {
    const float h = hash01(pos.x, pos.y, pos.z, static_cast<float>(e));
    const float ndi = dot(n, neg(incoming));
    switch (e) {
      case 1u: {
        const float r0 = smax(0.0f, 1.0f - __builtin_fabsf(ndi));
        const float rim = r0 * r0;
        k = 0.65f + 0.25f * h;
        c = add(muls(c, k), muls(mk(0.25f, 0.22f, 0.28f), rim * (0.4f + m.sheen)));
        break;
      }
      case 2u: {
        const float stripes = 0.5f + 0.5f * ptm::sinf_p(pos.x * 7.0f + pos.z * 5.0f + h * 6.0f);
        k = 0.45f + 0.55f * stripes;
        c = muls(c, k);
        break;
      }
      case 3u:
        k = h > 0.72f ? 0.18f : 1.0f;
        c = muls(c, k);
        break;
      case 4u: {
        const float vein = 0.5f + 0.5f * ptm::sinf_p((pos.x + pos.y * 0.4f + pos.z * 0.7f) * 9.0f + h * 3.0f);
        k = 0.55f + 0.45f * vein;
        c = add(muls(c, k), muls(mk(0.18f, 0.20f, 0.23f), 1.0f - vein));
        break;
      }
      case 5u:
        k = 0.55f + 0.35f * h;
        c = add(muls(c, k), muls(mk(0.45f, 0.16f, 0.04f), 0.25f + 0.25f * h));
        break;
      case 6u:
        k = 0.65f + 0.35f * h;
        c = muls(c, k);
        break;
      case 7u:
        k = 0.78f;
        c = add(muls(c, k), muls(mk(1.0f, 0.62f, 0.42f), 0.16f));
        break;
      case 8u:
        k = 0.55f;
        c = add(muls(c, k), muls(mk(0.35f + 0.45f * h, 0.25f + 0.35f * (1.0f - h), 0.85f), 0.45f));
        break;
      case 9u: {
        const float check = fmod2(__builtin_floorf(pos.x * 4.0f) + __builtin_floorf(pos.z * 4.0f));
        k = check < 0.5f ? 1.0f : 0.28f;
        c = muls(c, k);
        break;
      }
      case 10u:
        k = 0.45f;
        c = add(muls(c, k), muls(mk(1.0f, 0.35f + 0.3f * h, 0.08f), 0.55f));
        break;
      case 11u: {
        const float streak = 0.5f + 0.5f * ptm::sinf_p(pos.y * 18.0f + h * 5.0f);
        k = 0.55f + 0.45f * streak;
        c = muls(c, k);
        break;
      }
      case 12u: {
        const float v = smax(0.0f, ndi);
        const float v2 = v * v;
        c = add(c, muls(mk(0.55f, 0.65f, 0.95f), v2 * v2 * v2));
        break;
      }
      case 13u:
        k = 0.65f + 0.35f * __builtin_fabsf(ptm::sinf_p(pos.x * 10.0f) * ptm::cosf_p(pos.z * 10.0f));
        c = muls(c, k);
        break;
      case 14u: {
        const float rim = ptm::powf_p(smax(0.0f, 1.0f - __builtin_fabsf(ndi)), 0.8f);
        k = 0.0f;
        c = muls(mk(0.15f, 0.75f, 1.0f), 0.2f + 0.8f * rim);
        break;
      }
      case 15u: {
        const float bands = 0.5f + 0.5f * ptm::sinf_p(pos.y * 11.0f + pos.x * 3.0f + h * 7.0f);
        const float rim = ptm::powf_p(smax(0.0f, 1.0f - __builtin_fabsf(ndi)), 1.5f);
        k = 0.25f + 0.45f * bands;
        c = add(muls(c, k), muls(mk(0.48f, 0.56f, 0.68f), 0.18f + 0.32f * rim));
        break;
      }
      case 16u:
      case 17u:
      case 18u:
      case 19u:
        c = wood_albedo(c, pos, m.roughness, m.clearcoat, e, k);
        break;
      default:
        break;
    }
  }
  pre = c;
  return mk(smin(1.5f, smax(0.0f, c.x)), smin(1.5f, smax(0.0f, c.y)), smin(1.5f, smax(0.0f, c.z)));
}

// SampleSceneEnvironment, PathTracer.cpp:213-272
__device__ V3 environment(const DevScene& sc, V3 d) {
  const uint32_t w = sc.env_w, h = sc.env_h;
  if (w == 0u || h == 0u || !sc.env_map) return ld3(sc.env);
  const float len_sq = d.x * d.x + d.y * d.y + d.z * d.z;
  if (len_sq <= 1.0e-12f) return ld3(sc.env);
  const float inv_len = 1.0f / __builtin_sqrtf(len_sq);
  const V3 dir = mk(d.x * inv_len, d.y * inv_len, d.z * inv_len);
  float u = 0.5f + ptm::atan2f_p(dir.z, dir.x) * 0.15915494309189533577f;
  u -= __builtin_floorf(u);
  const float v = ptm::acosf_p(sclamp(dir.y, -1.0f, 1.0f)) * 0.31830988618379067154f;
  const float xf = u * static_cast<float>(w);
  const float yf = sclamp(v, 0.0f, 1.0f) * static_cast<float>(h - 1u);
  const uint32_t x0 = static_cast<uint32_t>(__builtin_floorf(xf)) % w, x1 = (x0 + 1u) % w;
  const uint32_t y0 = static_cast<uint32_t>(__builtin_floorf(yf)), y1 = y0 + 1u < h - 1u ? y0 + 1u : h - 1u;
  const float tx = xf - __builtin_floorf(xf), ty = yf - __builtin_floorf(yf);
  const float* m = sc.env_map;
  const V3 c00 = ld3(m + 3u * (static_cast<size_t>(y0) * w + x0)), c10 = ld3(m + 3u * (static_cast<size_t>(y0) * w + x1));
  const V3 c01 = ld3(m + 3u * (static_cast<size_t>(y1) * w + x0)), c11 = ld3(m + 3u * (static_cast<size_t>(y1) * w + x1));
  const V3 top = mk(c00.x + (c10.x - c00.x) * tx, c00.y + (c10.y - c00.y) * tx, c00.z + (c10.z - c00.z) * tx);
  const V3 bot = mk(c01.x + (c11.x - c01.x) * tx, c01.y + (c11.y - c01.y) * tx, c01.z + (c11.z - c01.z) * tx);
  V3 color = mk(top.x + (bot.x - top.x) * ty, top.y + (bot.y - top.y) * ty, top.z + (bot.z - top.z) * ty);
  V3 scale = ld3(sc.env_scale);
  if (smax(scale.x, smax(scale.y, scale.z)) <= 1.0e-6f) scale = mk(1.0f, 1.0f, 1.0f);
  color.x *= scale.x;
  color.y *= scale.y;
  color.z *= scale.z;
  return color;
}

// spot_attenuation, ScalarCpuPathTracer.cpp:168-185
__device__ __forceinline__ float spot_attenuation(const PtLight& l, V3 from_light_dir) {
  if (l.spot_inner_cos <= -0.999f) return 1.0f;
  const float cone = dot(normalize(from_light_dir), normalize(ld3(l.direction)));
  if (cone <= l.spot_outer_cos) return 0.0f;
  if (cone >= l.spot_inner_cos || l.spot_inner_cos <= l.spot_outer_cos) return 1.0f;
  const float t = sclamp((cone - l.spot_outer_cos) / (l.spot_inner_cos - l.spot_outer_cos), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

__device__ __forceinline__ float mis_weight(float a, float b) {
  const float a2 = a * a, b2 = b * b;
  const float denom = a2 + b2;
  if (denom <= 0.0f) return 1.0f;
  return a2 / denom;
}

// light_pdf_for_direction, ScalarCpuPathTracer.cpp:1120-1138
__device__ __forceinline__ float light_pdf_for_direction(const DevScene& sc, V3 position, V3 direction) {
  if (sc.light_count == 0) return 0.0f;
  const float inv_n = 1.0f / static_cast<float>(sc.light_count);
  float total = 0.0f;
  for (uint32_t i = 0; i < sc.light_count; ++i) {
    const V3 to_light = sub(ld3(sc.lights[i].position), position);
    const float dist = __builtin_sqrtf(dot(to_light, to_light));
    if (dist < kEps) continue;
    const V3 dir = divs(to_light, dist);
    if (dot(dir, direction) > 0.9999f) total += inv_n / (dist * dist + kEps);
  }
  return total;
}

// sample_hemisphere, ScalarCpuPathTracer.cpp:978-988
__device__ __forceinline__ V3 sample_hemisphere(uint64_t& rng, V3 n) {
  const float u1 = next01(rng);
  const float u2 = next01(rng);
  const float r = __builtin_sqrtf(smax(0.0f, 1.0f - u1));
  const float phi = 2.0f * kPi * u2;
  float sp, cp;
  ptm::sincosf_p(phi, sp, cp);
  const V3 local = mk(r * cp, r * sp, __builtin_sqrtf(smax(0.0f, u1)));
  V3 tangent = cross((__builtin_fabsf(n.z) < 0.999f) ? mk(0.0f, 0.0f, 1.0f) : mk(0.0f, 1.0f, 0.0f), n);
  tangent = normalize(tangent);
  const V3 bitangent = cross(n, tangent);
  return normalize(add(add(muls(tangent, local.x), muls(bitangent, local.y)), muls(n, local.z)));
}

__device__ __forceinline__ V3 reflect_about(V3 d, V3 n) { return sub(d, muls(n, 2.0f * dot(n, d))); }

__device__ __forceinline__ float schlick_fresnel(float cosTheta, float ior) {
  const float safeIor = smax(1.01f, ior);
  float r0 = (1.0f - safeIor) / (1.0f + safeIor);
  r0 *= r0;
  const float m = 1.0f - smin(1.0f, smax(0.0f, cosTheta));
  return r0 + (1.0f - r0) * m * m * m * m * m;
}

__device__ __forceinline__ V3 refract_dir(V3 incoming, V3 normal, float eta, bool& tir) {
  const float cosi = smin(1.0f, smax(0.0f, dot(neg(incoming), normal)));
  const float k = 1.0f - eta * eta * (1.0f - cosi * cosi);
  if (k < 0.0f) {
    tir = true;
    return mk(0.0f, 0.0f, 0.0f);
  }
  tir = false;
  return normalize(add(muls(incoming, eta), muls(normal, eta * cosi - __builtin_sqrtf(k))));
}

__device__ __forceinline__ float pow5(float v) {
  const float v2 = v * v;
  return v2 * v2 * v;
}

// d log p(w) / d e for the direction the glossy branch returned: p is the density of
// sample_phong_lobe's output (the exponent-e lobe around refl plus the lobe samples that fell below
// the surface and were mirrored into w) mixed with the sheen branch's cosine hemisphere of weight
// ps: p(w) = (1 - ps) [q(w) + q(w')] + ps cos_n / pi, q = (e+1)/(2 pi) max(0, w.refl)^e, w' = w
// mirrored through the surface plane (the fallback for a mirrored sample still below the surface
// has probability zero). In the log domain, in double: e reaches 5e6.
__device__ double phong_score(double e, V3 refl, V3 n, V3 w, double ps) {
  const double wn = static_cast<double>(dot(w, n));
  const V3 wm = sub(w, muls(n, 2.0f * dot(w, n)));
  const double c[2] = {static_cast<double>(dot(w, refl)), static_cast<double>(dot(wm, refl))};
  double l[2], a[2];
  bool any = false;
  double L = -__builtin_huge_val();
  for (int i = 0; i < 2; ++i) {
    if (c[i] > 0.0) {
      l[i] = e * log(c[i]);
      a[i] = 1.0 / (e + 1.0) + log(c[i]);
      if (l[i] > L) L = l[i];
      any = true;
    } else {
      l[i] = -__builtin_huge_val();
      a[i] = 0.0;
    }
  }
  if (!any || ps >= 1.0) return 0.0;
  double num = 0.0, den = 0.0;
  for (int i = 0; i < 2; ++i)
    if (c[i] > 0.0) {
      const double wgt = exp(l[i] - L);
      num += wgt * a[i];
      den += wgt;
    }
  if (ps > 0.0 && wn > 0.0) {
    const double kPiD = 3.14159265358979323846;
    const double lx = log(ps * wn / kPiD) - log((1.0 - ps) * (e + 1.0) / (2.0 * kPiD)) - L;
    den += exp(fmin(lx, 700.0));
  }
  return num / den;
}

struct ShadeOut {
  bool cont, shadow;
  V3 no, nd;          // next extension ray
  V3 so, sd, contrib; // shadow ray and its radiance contribution if unoccluded
  float stmax;
  V3 galb, glight;    // adjoint: gradient pieces applied if the light is visible
  float grough;
  float2 guv;         // texture coordinate of the vertex (textured materials)
  uint32_t gmat, glight_index;
};

// d(clamped effect colour)/d(albedo): k inside the (0, 1.5) clamp, 0 on or outside it.
__device__ __forceinline__ float eff_grad(float pre_c, float k) { return (pre_c > 0.0f && pre_c < 1.5f) ? k : 0.0f; }

// One iteration of ScalarCpuPathTracer::trace (ScalarCpuPathTracer.cpp:1159-1287) for a path that
// arrived at a hit (prim >= 0) or a miss (prim < 0). kAdj replays the iteration for the adjoint.
// One iteration of trace's loop for the path in slot p whose extension ray (ro, rd) hit prim at t (-1:
// missed). T, L and rng are the path's throughput, radiance and RNG state, updated in place; the caller
// loads and stores them (the radiance is final when out.cont is false).
template <bool kAdj>
__device__ __forceinline__ void shade_path(const DevPaths& P, const DevScene& sc, const DevSettings& st, uint32_t depth,
                                           uint32_t p, V3 ro, V3 rd, float t, int32_t prim, V3& T, V3& L,
                                           uint64_t& rng, ShadeOut& out, uint32_t& shadow_tests, const DevAdjoint& A,
                                           float* glds, AdjItem& ai) {
  out.cont = false;
  out.shadow = false;
  uint32_t aflags = 0;
  V3 dl = mk(0.0f, 0.0f, 0.0f);
  if constexpr (kAdj) {
    if (!(finite(ai.ltot[0]) && finite(ai.ltot[1]) && finite(ai.ltot[2]))) return;  // dropped sample
    adj_flush(A, sc, glds, p, L, ai);
    aflags = ai.flags;
    dl = mk(ai.delta[0], ai.delta[1], ai.delta[2]);
  }
  if (prim == -1) {
    L = add(L, mulv(T, environment(sc, rd)));
    return;
  }
  const V3 pos = add(ro, muls(rd, t));
  V3 gN;
  uint32_t raw_mat;
  if (prim >= 0) {
    const float4 pn = sc.prims[prim];
    gN = mk(pn.x, pn.y, pn.z);
    raw_mat = static_cast<uint32_t>(__float_as_int(pn.w));
  } else {  // an SDF primitive (k_sdf_extend)
    const DevSdf& sp = sc.sdfs[-2 - prim];
    gN = sdf_hit_normal(sp, ro, rd, t);
    raw_mat = sp.material;
  }
  const bool entering = dot(rd, gN) < 0.0f;
  const V3 sN = entering ? gN : neg(gN);
  const uint32_t mat_index = raw_mat < sc.material_count - 1u ? raw_mat : sc.material_count - 1u;
  const PtMaterial& m = sc.materials[mat_index];
  float eff_k;
  V3 eff_pre;
  // the material's base colour, times its texture at the hit (no counterpart on the reference's
  // CPU path, which renders untextured)
  V3 base = ld3(m.albedo);
  float2 huv = make_float2(0.0f, 0.0f);
  if (m.base_color_texture_index < sc.texture_count && prim >= 0) {
    // the hit's barycentrics: the closest-hit test of the same ray against the same triangle record,
    // recomputed here (bit-identical) rather than carried through the traversal
    float ht = 0.0f, hu = 0.0f, hv = 0.0f;
    tri_test(sc.tris + 3u * sc.prim_tri[prim], ro, rd, kInf, ht, hu, hv);
    huv = uv_at(sc.prim_uv + 6u * static_cast<uint32_t>(prim), hu, hv);
    base = mulv(base, tex_tap(sc, m.base_color_texture_index, huv).color);
  }
  const V3 albedo = effect_albedo(m, base, pos, sN, rd, eff_k, eff_pre);
  const float roughness = clamp01(m.roughness);
  MatFlags f;
  f.emissive_model = m.material_model == 1u || m.emissive[0] > 0.0f || m.emissive[1] > 0.0f || m.emissive[2] > 0.0f;
  f.mirror = m.material_model == 2u || roughness <= 0.001f;
  f.metallic = m.material_model == 4u || m.metallic > 0.65f;
  f.transmissive = m.material_model == 5u || m.transmission > 0.05f;
  f.diffuse = roughness >= 0.999f && !f.mirror && !f.metallic && !f.transmissive;
  f.sheen = m.material_model == 6u;
  f.clearcoat = m.material_model == 7u || m.clearcoat > 0.05f;
  f.toon = m.material_model == 8u;
  const V3 emissive = ld3(m.emissive);

  float e_weight = -1.0f;
  if (!st.enable_nee || depth == 0u) {
    L = add(L, mulv(T, emissive));
    e_weight = 1.0f;
  } else if (st.enable_mis && f.emissive_model) {
    const float bsdf_pdf = smax(0.0f, dot(sN, neg(rd))) * kInvPi;
    const float l_pdf = light_pdf_for_direction(sc, ro, rd);
    const float w = mis_weight(bsdf_pdf, l_pdf);
    L = add(L, muls(mulv(T, emissive), w));
    e_weight = w;
  }
  if constexpr (kAdj) {
    if (e_weight >= 0.0f) {
      gadd(A, glds, kGradEmission, mat_index, 0, dl.x * T.x * e_weight);
      gadd(A, glds, kGradEmission, mat_index, 1, dl.y * T.y * e_weight);
      gadd(A, glds, kGradEmission, mat_index, 2, dl.z * T.z * e_weight);
    }
  }

  const bool last = depth + 1u >= st.max_depth;
  if (st.enable_nee && !last && sc.light_count > 0u) {
    // sample_direct_light, ScalarCpuPathTracer.cpp:1022-1118
    const uint32_t nl = sc.light_count;
    const uint64_t li = static_cast<uint64_t>(next01(rng) * static_cast<float>(nl));
    const uint32_t light_index = static_cast<uint32_t>(li < nl - 1u ? li : nl - 1u);
    const PtLight& light = sc.lights[light_index];
    V3 lp = ld3(light.position);
    if (light.radius > 1.0e-5f) {
      const float uz = 1.0f - 2.0f * next01(rng);
      const float upr = __builtin_sqrtf(smax(0.0f, 1.0f - uz * uz));
      const float phi = 6.28318530718f * next01(rng);
      float sp, cp;
      ptm::sincosf_p(phi, sp, cp);
      lp = add(lp, muls(mk(upr * cp, uz, upr * sp), light.radius));
    }
    const V3 to_light = sub(lp, pos);
    const float dist_sq = dot(to_light, to_light);
    const float dist = __builtin_sqrtf(dist_sq);
    if (!(dist < kEps)) {
      const V3 light_dir = divs(to_light, dist);
      const float cos_theta = dot(sN, light_dir);
      if (cos_theta > 0.0f) {
        const float spot = spot_attenuation(light, neg(light_dir));
        if (spot > 0.0f) {
          ++shadow_tests;
          const V3 view_dir = neg(rd);
          // the NEE albedo sees the light direction as the incoming one (ScalarCpuPathTracer.cpp:1089)
          float nk = eff_k;
          V3 npre = eff_pre;
          const V3 nalb = m.material_effect != 0u ? effect_albedo(m, base, pos, sN, neg(light_dir), nk, npre) : albedo;
          const float kscale = (light.intensity * spot) / (dist_sq + kEps);
          const V3 irradiance = muls(ld3(light.color), kscale);
          const float nlf = static_cast<float>(nl);
          V3 direct = mk(0.0f, 0.0f, 0.0f);
          V3 d_alb = mk(0.0f, 0.0f, 0.0f), d_col = mk(0.0f, 0.0f, 0.0f), d_rough = mk(0.0f, 0.0f, 0.0f);  // adjoint partials
          if (!f.mirror && !f.transmissive) {
            direct = add(direct, muls(muls(mulv(muls(nalb, kInvPi), irradiance), cos_theta), nlf));
            if constexpr (kAdj) {
              d_alb = muls(irradiance, kInvPi * cos_theta * nlf);
              d_col = muls(nalb, kInvPi * kscale * cos_theta * nlf);
            }
          }
          if (f.mirror || f.metallic || f.clearcoat || f.transmissive || roughness < 0.65f) {
            const V3 half_dir = normalize(add(light_dir, normalize(view_dir)));
            const float effR = smax(0.025f, roughness * (f.metallic ? 0.65f : 1.0f));
            const float a2 = effR * effR;
            const float specPower = sclamp(2.0f / smax(0.0005f, a2 * a2) - 2.0f, 4.0f, 96.0f);
            const float spec = ptm::powf_p(smax(0.0f, dot(sN, half_dir)), specPower);
            const float cosView = smax(0.0f, dot(sN, normalize(view_dir)));
            float f0 = (1.0f - m.ior) / (1.0f + m.ior);
            f0 *= f0;
            const float fresnel = f0 + (1.0f - f0) * pow5(1.0f - cosView);
            const float ss_raw = (1.0f - roughness) * 0.8f + m.metallic * 0.45f + m.clearcoat * 0.35f +
                                 (f.mirror ? 0.75f : 0.0f) + (f.transmissive ? fresnel : 0.0f);
            const float specStrength = clamp01(ss_raw);
            const V3 specTint = add(muls(mk(1.0f, 1.0f, 1.0f), f.metallic ? 0.15f : 0.88f),
                                    muls(nalb, f.metallic ? 0.85f : 0.12f));
            const float s = spec * cos_theta * nlf * smax(0.15f, specStrength);
            direct = add(direct, muls(mulv(specTint, irradiance), s));
            if constexpr (kAdj) {
              d_alb = add(d_alb, muls(irradiance, (f.metallic ? 0.85f : 0.12f) * s));
              d_col = add(d_col, muls(specTint, kscale * s));
              // d s / d roughness: through the Blinn-Phong exponent 2 / effR^4 - 2 (unclamped in
              // (4, 96)) and through specStrength (unclamped in (0, 1), above the 0.15 floor), times
              // the clamp01 derivative of the roughness parameter
              const double nh = static_cast<double>(smax(0.0f, dot(sN, half_dir)));
              const double raw = 2.0 / fmax(0.0005, static_cast<double>(a2) * static_cast<double>(a2)) - 2.0;
              const double kM = f.metallic ? 0.65 : 1.0;
              double dspec = 0.0;
              if (nh > 0.0 && raw > 4.0 && raw < 96.0 && static_cast<double>(roughness) * kM > 0.025)
                dspec = static_cast<double>(spec) * log(nh) * (-8.0 / pow(static_cast<double>(effR), 5.0)) * kM;
              double ds = dspec * static_cast<double>(smax(0.15f, specStrength));
              if (specStrength > 0.15f && ss_raw > 0.0f && ss_raw < 1.0f) ds += static_cast<double>(spec) * -0.8;
              ds *= static_cast<double>(cos_theta) * static_cast<double>(nlf) *
                    ((m.roughness > 0.0f && m.roughness < 1.0f) ? 1.0 : 0.0);
              d_rough = muls(mulv(specTint, irradiance), static_cast<float>(ds));
            }
          }
          V3 contrib;
          float w = 1.0f;
          if (st.enable_mis) {
            const V3 l_dir = normalize(sub(ld3(sc.lights[0].position), pos));
            const float ct = smax(0.0f, dot(sN, l_dir));
            const float bsdf_pdf = ct * kInvPi;
            const float l_pdf = 1.0f / nlf;
            w = mis_weight(l_pdf, bsdf_pdf);
            contrib = muls(mulv(T, direct), w);
          } else {
            contrib = mulv(T, direct);
          }
          V3 galb = mk(0.0f, 0.0f, 0.0f), glight = mk(0.0f, 0.0f, 0.0f);
          float grough = 0.0f;
          if constexpr (kAdj) {
            const V3 tw = muls(mulv(dl, T), w);
            galb = mk(tw.x * d_alb.x * eff_grad(npre.x, nk), tw.y * d_alb.y * eff_grad(npre.y, nk),
                      tw.z * d_alb.z * eff_grad(npre.z, nk));
            glight = mulv(tw, d_col);
            grough = tw.x * d_rough.x + tw.y * d_rough.y + tw.z * d_rough.z;
          }
          const float remaining = dist - 0.004f;
          if (remaining > 0.004f) {
            out.shadow = true;
            out.so = add(pos, muls(sN, 0.002f));
            out.sd = light_dir;
            out.stmax = remaining;
            out.contrib = contrib;
            out.galb = galb;
            out.glight = glight;
            out.grough = grough;
            out.guv = huv;
            out.gmat = mat_index;
            out.glight_index = light_index;
          } else {
            L = add(L, contrib);
            if constexpr (kAdj) {
              gadd_base(A, sc, glds, mat_index, huv, galb);
              gadd(A, glds, kGradLight, light_index, 0, glight.x);
              gadd(A, glds, kGradLight, light_index, 1, glight.y);
              gadd(A, glds, kGradLight, light_index, 2, glight.z);
              gadd(A, glds, kGradRough, mat_index, 0, grough);
            }
          }
        }
      }
    }
  }
  if (last) return;  // the loop ends after this iteration; nothing below reaches the radiance

  V3 outDir;
  bool refracted = false;
  float score = 0.0f;  // adjoint: d log(density of outDir) / d roughness (glossy branch only)
  if (f.mirror) {
    outDir = reflect_about(rd, sN);
  } else if (f.transmissive) {
    const float cosTheta = smax(0.0f, dot(sN, neg(rd)));
    const float fresnel = schlick_fresnel(cosTheta, m.ior);
    if (next01(rng) < smin(0.98f, fresnel + m.clearcoat * 0.015f)) {
      outDir = reflect_about(rd, sN);
    } else {
      bool tir = false;
      const float safeIor = smax(1.01f, m.ior);
      const float eta = entering ? (1.0f / safeIor) : safeIor;
      outDir = refract_dir(rd, sN, eta, tir);
      if (tir || dot(outDir, outDir) <= 1.0e-8f) outDir = reflect_about(rd, sN);
      else refracted = true;
    }
  } else {
    // sample_hemisphere (diffuse, toon; ScalarCpuPathTracer.cpp:978-988) or sample_phong_lobe (the
    // rest; :996-1020). Both draw u1 then u2, turn u2
    // into the angle 2 pi u2 and place (s cos, s sin, z) in a frame around an axis (the normal, or the
    // mirror direction) with the same operations, so the angle's sine and cosine and the frame are
    // evaluated once for lanes of either kind.
    const bool hemi = f.diffuse || f.toon;
    float effR = 0.0f, expt = 0.0f;
    V3 refl = sN;
    const float u1 = next01(rng);
    const float u2 = next01(rng);
    float radial, lz;
    V3 axis;
    if (hemi) {
      radial = __builtin_sqrtf(smax(0.0f, 1.0f - u1));
      lz = __builtin_sqrtf(smax(0.0f, u1));
      axis = sN;
    } else {
      effR = smax(0.025f, roughness * (f.metallic ? 0.75f : 1.0f) * (f.clearcoat ? 0.65f : 1.0f));
      const float a2 = effR * effR;
      expt = smax(0.0f, 2.0f / (a2 * a2) - 2.0f);
      refl = reflect_about(rd, sN);
      lz = ptm::powf_p(smax(0.0f, u1), 1.0f / (expt + 1.0f));  // cos theta of the lobe
      radial = __builtin_sqrtf(smax(0.0f, 1.0f - lz * lz));
      axis = refl;
    }
    const float phi = 2.0f * kPi * u2;
    float sp, cp;
    ptm::sincosf_p(phi, sp, cp);
    const V3 ref_t = (__builtin_fabsf(axis.z) < 0.999f) ? mk(0.0f, 0.0f, 1.0f) : mk(0.0f, 1.0f, 0.0f);
    const V3 tang = normalize(cross(ref_t, axis));
    const V3 btan = cross(axis, tang);
    outDir = normalize(add(add(muls(tang, radial * cp), muls(btan, radial * sp)), muls(axis, lz)));
    if (!hemi) {
      // the rest of sample_phong_lobe: a sample below the surface is mirrored, else resampled
      if (dot(outDir, sN) <= 0.0f) {
        outDir = sub(outDir, muls(sN, 2.0f * dot(outDir, sN)));
        if (dot(outDir, sN) <= 0.0f) outDir = sample_hemisphere(rng, sN);
      }
      if (f.sheen && next01(rng) < m.sheen * 0.35f) outDir = sample_hemisphere(rng, sN);
    }
    if constexpr (kAdj) {
      const double kk = (f.metallic ? 0.75 : 1.0) * (f.clearcoat ? 0.65 : 1.0);
      const double ps = f.sheen ? fmin(1.0, fmax(0.0, static_cast<double>(m.sheen * 0.35f))) : 0.0;
      if (!hemi && m.roughness > 0.0f && m.roughness < 1.0f && static_cast<double>(roughness) * kk > 0.025 &&
          expt > 0.0f)
        score = static_cast<float>(phong_score(static_cast<double>(expt), refl, sN, outDir, ps) *
                                   (-8.0 / pow(static_cast<double>(effR), 5.0)) * kk);
    }
  }
  if (!f.transmissive && dot(outDir, sN) <= 0.0f) return;
  V3 bw = albedo;
  if (f.metallic || f.mirror) {
    bw = muls(albedo, 0.65f + 0.35f * m.metallic);
  } else if (f.transmissive) {
    const float tint = clamp01(m.alpha) * m.transmission;
    bw = add(muls(mk(1.0f, 1.0f, 1.0f), 1.0f - tint), muls(albedo, tint));
    if (refracted) bw = muls(bw, smax(0.55f, m.transmission));
  }
  if (f.toon) bw = muls(bw, dot(outDir, sN) > 0.55f ? 1.0f : 0.45f);
  T = mulv(T, bw);
  if (vmax3(T) < 0.001f) return;
  const float maxT = vmax3(T);
  const float rr = smin(st.rr_max, smax(st.rr_min, maxT));
  if (depth >= st.rr_start && next01(rng) > rr) return;
  if constexpr (kAdj) {
    // d log bw / d albedo per channel, and the channel std::max picks for rr
    const int cstar = (T.x < smax(T.y, T.z)) ? ((T.y < T.z) ? 2 : 1) : 0;
    const bool trans_tint = f.transmissive && !(f.metallic || f.mirror);
    const float tint = trans_tint ? clamp01(m.alpha) * m.transmission : 0.0f;
    const float a3[3] = {albedo.x, albedo.y, albedo.z}, b3[3] = {bw.x, bw.y, bw.z};
    float dx[3];
    for (int c = 0; c < 3; ++c) {
      float v;
      if (trans_tint) {
        const float base = (1.0f - tint) + a3[c] * tint;
        v = base > 0.0f ? tint / base : 0.0f;
      } else {
        v = a3[c] > 0.0f ? 1.0f / a3[c] : 0.0f;
      }
      const float pre_c = c == 0 ? eff_pre.x : (c == 1 ? eff_pre.y : eff_pre.z);
      dx[c] = b3[c] > 0.0f ? v * eff_grad(pre_c, eff_k) : 0.0f;
      ai.pdx[c] = dx[c];
    }
    ai.mat = mat_index;
    ai.uv[0] = huv.x;
    ai.uv[1] = huv.y;
    ai.score = score;
    const bool active = depth < st.rr_start && st.rr_min < maxT && maxT < st.rr_max;
    aflags = (aflags & ~0xFu) | 1u | (active ? 2u : 0u) | (static_cast<uint32_t>(cstar) << 2);
    uint32_t nh = (aflags >> 4) & 7u;
    if (depth < st.rr_start && nh < static_cast<uint32_t>(kAdjHist)) {
      hist_mat(A, nh)[p] = mat_index;
      hist_uv(A, nh, 0)[p] = huv.x;
      hist_uv(A, nh, 1)[p] = huv.y;
      for (int c = 0; c < 3; ++c) hist_dx(A, nh, c)[p] = dx[c];
      ++nh;
      aflags = (aflags & ~0x70u) | (nh << 4);
    }
    ai.flags = aflags;
  }
  T = divs(T, rr);
  out.cont = true;
  out.no = add(pos, muls(outDir, 0.002f));
  out.nd = outDir;
}

template <bool kAdj>
__device__ __forceinline__ void shade_kernel(const DevPaths& P, const DevRays& in, const DevRays& out, const DevScene& sc,
                                             const DevSettings& st, uint32_t depth, const DevAdjoint& A, float* glds,
                                             uint32_t in_buf) {
  __shared__ uint32_t prefix[kSegments + 1];
  if constexpr (kAdj) {
    tex_table_init(glds);
    grad_lds_clear(A, glds);
  }
  load_prefix(P.queue + kQueueInCount, 1, prefix);
  const uint32_t n = prefix[kSegments];
  const uint32_t stride = gridDim.x * kBlock;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t shadow_tests = 0, shadow_rays = 0;
  for (uint32_t w0 = blockIdx.x * kBlock + (threadIdx.x & ~63u); w0 < n; w0 += stride) {
    const uint32_t i = w0 + lane;
    ShadeOut so;
    so.cont = false;
    so.shadow = false;
    uint32_t p = 0, sid = 0;
    V3 T = mk(0.0f, 0.0f, 0.0f), L = T;
    uint64_t rng = 0;
    AdjItem ai;
    if (i < n) {
      const uint32_t q = queue_pos(prefix, i, P.seg_cap);
      p = depth > 0u ? in.qpath[q] : q;  // the first bounce's queue is the identity
      sid = in.qscene[q];
      const V3 ro = mk(in.c[0][q], in.c[1][q], in.c[2][q]);
      const V3 rd = mk(in.c[3][q], in.c[4][q], in.c[5][q]);
      if (depth > 0u) {  // camera rays start with throughput 1 and radiance 0 (k_raygen stores neither)
        T = mk(in.tp[0][q], in.tp[1][q], in.tp[2][q]);
        L = mk(in.rad[0][q], in.rad[1][q], in.rad[2][q]);
      } else {
        T = mk(1.0f, 1.0f, 1.0f);
      }
      rng = in.rng[q];
      if constexpr (kAdj) ai = load_item(A.items[in_buf], A.stride, q);
      shade_path<kAdj>(P, scene_view(sc, sid), st, depth, p, ro, rd, P.hit_t[q], P.hit_prim[q], T, L, rng, so,
                       shadow_tests, A, glds, ai);
      if (!kAdj && !so.cont) {
        // The path ends here: its radiance goes to its slot, and so does the number of extension rays
        // it traced (depth + 1), which k_film takes back out of the ray counter for a dropped
        // (non-finite) sample (shade_pixel adds ray_count only for finite samples).
        P.rad[0][p] = L.x; P.rad[1][p] = L.y; P.rad[2][p] = L.z;
        P.ended_rays[p] = depth + 1u;
      }
    }
    const uint32_t seg = (w0 >> 6) & (kSegments - 1u);  // this chunk's output segment
    const uint32_t slot = segment_append(P.queue + kQueueOutCursor, seg, P.seg_cap, so.cont);
    if (so.cont) {
      out.c[0][slot] = so.no.x; out.c[1][slot] = so.no.y; out.c[2][slot] = so.no.z;
      out.c[3][slot] = so.nd.x; out.c[4][slot] = so.nd.y; out.c[5][slot] = so.nd.z;
      out.qpath[slot] = p;
      out.qscene[slot] = sid;
      out.tp[0][slot] = T.x; out.tp[1][slot] = T.y; out.tp[2][slot] = T.z;
      out.rad[0][slot] = L.x; out.rad[1][slot] = L.y; out.rad[2][slot] = L.z;
      out.rng[slot] = rng;
      if constexpr (kAdj) store_item(A.items[in_buf ^ 1u], A.stride, slot, ai);
    }
    const uint32_t sslot = segment_append(P.queue + kQueueShadowCursor, seg, P.seg_cap, so.shadow);
    if (so.shadow) {
      P.sray[0][sslot] = so.so.x; P.sray[1][sslot] = so.so.y; P.sray[2][sslot] = so.so.z;
      P.sray[3][sslot] = so.sd.x; P.sray[4][sslot] = so.sd.y; P.sray[5][sslot] = so.sd.z;
      P.stmax[sslot] = so.stmax;
      P.scontrib[0][sslot] = so.contrib.x; P.scontrib[1][sslot] = so.contrib.y; P.scontrib[2][sslot] = so.contrib.z;
      // where the contribution goes if the light is visible: an adjoint replay needs the radiance of
      // continuing paths only (the remainder of their pending bounce)
      P.spath[sslot] = so.cont ? (kRadQueued | slot) : (kAdj ? kRadNone : p);
      P.sscene[sslot] = sid;
      if constexpr (kAdj) {
        A.s_galb[0][sslot] = so.galb.x; A.s_galb[1][sslot] = so.galb.y; A.s_galb[2][sslot] = so.galb.z;
        A.s_glight[0][sslot] = so.glight.x; A.s_glight[1][sslot] = so.glight.y; A.s_glight[2][sslot] = so.glight.z;
        A.s_mat[sslot] = so.gmat;
        A.s_uv[0][sslot] = so.guv.x;
        A.s_uv[1][sslot] = so.guv.y;
        A.s_light[sslot] = so.glight_index;
        A.s_grough[sslot] = so.grough;
      }
      ++shadow_rays;
    }
    if constexpr (kAdj) tex_grad_flush(A, glds);  // the wave is converged here
  }
  wave_add(P.counters, kShadowTests, shadow_tests);
  wave_add(P.counters, kShadowRays, shadow_rays);
  if constexpr (kAdj) grad_lds_flush(A, glds);
}

// Waves per SIMD the forward shade is compiled for (its register budget).
#ifndef PT_SHADE_WAVES
#define PT_SHADE_WAVES 4
#endif
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PT_SHADE_WAVES, 8))) void k_shade(
    DevPaths P, DevRays in, DevRays out, DevScene sc, DevSettings st,
                                                 uint32_t depth) {
  shade_kernel<false>(P, in, out, sc, st, depth, DevAdjoint{}, nullptr, 0u);
}

// The adjoint shade holds about twice the forward's live values; letting it run PT_SHADE_ADJ_WAVES
// waves per SIMD (the compiler spills what does not fit) trades spills for latency hiding.
#ifndef PT_SHADE_ADJ_WAVES
#define PT_SHADE_ADJ_WAVES 2
#endif
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PT_SHADE_ADJ_WAVES, 8))) void k_shade_adj(
    DevPaths P, DevRays in, DevRays out, DevScene sc, DevSettings st, uint32_t depth, DevAdjoint A, uint32_t in_buf) {
  __shared__ float glds[kAdjLdsFloats + kTexTableFloats];
  shade_kernel<true>(P, in, out, sc, st, depth, A, glds, in_buf);
}

// NEE visibility over the bounce's shadow queue. Opaque scenes: an any-hit query with
// t_max = dist - 0.004. Scenes with transmissive materials: the reference's closest-hit loop
// that steps through up to 8 transmissive blockers (ScalarCpuPathTracer.cpp:1066-1087), each
// step a follow-up query on the same lane.
// Set in a shadow record's light index (adjoint passes) when the light turned out visible; the
// record's radiance and gradients are then applied by k_shadow_grad.
constexpr uint32_t kShadowVisible = 0x80000000u;

template <bool kAdj>
__device__ __forceinline__ void light_visible(const DevPaths& P, const DevScene& sc, const DevAdjoint& A, float* glds,
                                              const RadTarget& rq, uint32_t q) {
  const uint32_t target = P.spath[q];
  if (target != kRadNone) {
    float* const* rad = P.rad;
    uint32_t k = target;
    if (target & kRadQueued) {
      rad = rq.rad;
      k = target & ~kRadQueued;
    }
    rad[0][k] = rad[0][k] + P.scontrib[0][q];
    rad[1][k] = rad[1][k] + P.scontrib[1][q];
    rad[2][k] = rad[2][k] + P.scontrib[2][q];
  }
  if constexpr (kAdj) {
    const uint32_t mi = A.s_mat[q], li = A.s_light[q] & ~kShadowVisible;
    gadd_base(A, sc, glds, mi, make_float2(A.s_uv[0][q], A.s_uv[1][q]),
              mk(A.s_galb[0][q], A.s_galb[1][q], A.s_galb[2][q]));
    for (int c = 0; c < 3; ++c) gadd(A, glds, kGradLight, li, c, A.s_glight[c][q]);
    gadd(A, glds, kGradRough, mi, 0, A.s_grough[q]);
  }
}

// kAny: the scene has neither transmissive materials nor SDF primitives, so a shadow ray is one
// any-hit query; compiled apart from the skip loop so that it keeps extend-like register use.
// Top bit of a shadow record's scene index: the forward any-hit query found the light unoccluded.
constexpr uint32_t kShadowReached = 0x80000000u;

template <bool kAdj, bool kSdf, bool kAny>
__global__ __launch_bounds__(kBlock) void k_shadow(DevPaths P, DevScene sc, DevAdjoint A, RadTarget rq) {
  static_assert(!(kAny && kSdf), "SDF shadows need the closest-hit skip loop");
  __shared__ int32_t s_ref[kStackLds * kBlock];
  __shared__ float s_t[kStackLds * kBlock];
  __shared__ float4 s_hot[4 * (kHotNodes > 0 ? kHotNodes : 1)];
  load_hot(sc, s_hot);
  __shared__ uint32_t prefix[kSegments + 1];
  load_prefix(P.queue + kQueueShadowCursor, kSegLine, prefix);
  // forward: a visible light adds its contribution now; adjoint: the record is marked and
  // k_shadow_grad applies it (keeping this traversal kernel as lean as the forward one)
  auto visible = [&](uint32_t q) {
    if constexpr (kAdj) A.s_light[q] |= kShadowVisible;
    else light_visible<false>(P, sc, A, nullptr, rq, q);
  };
  const Stack stk = make_stack(P, sc, s_ref, s_t, s_hot);
  const uint32_t n = prefix[kSegments];
  TravStats st = {0, 0, 0, 0, 0};
  uint32_t q = 0;
  SdfStats sst = {0, 0, 0};
  if constexpr (kAny) {
    uint32_t sid = 0;
    auto begin = [&](uint32_t j, Trav& s) {
      q = queue_pos(prefix, j, P.seg_cap);
      sid = P.sscene[q];
      const V3 o = mk(P.sray[0][q], P.sray[1][q], P.sray[2][q]);
      const V3 d = mk(P.sray[3][q], P.sray[4][q], P.sray[5][q]);
      return trav_begin(sc.recs[sid], s, o, d, P.stmax[q], st);
    };
    // forward: only mark the record (a store, no read-modify-write of the path inside the
    // traversal loop); k_shadow_apply adds the contribution right after this kernel
    auto end = [&](uint32_t, Trav& s) {
      if (s.best_prim < 0) {
        if constexpr (kAdj) visible(q);
        else P.sscene[q] = sid | kShadowReached;
      }
      return false;
    };
    drive_queries<true>(sc, stk, P.queue, n, P.refill_min, st, begin, end);
  } else {
    V3 o, d;
    float remaining = 0.0f;
    uint32_t skips = 0, sid = 0;
    // Starts the next step of the loop; false when the loop is over (then `end` is called).
    auto begin = [&](uint32_t j, Trav& s) {
      q = queue_pos(prefix, j, P.seg_cap);
      o = mk(P.sray[0][q], P.sray[1][q], P.sray[2][q]);
      d = mk(P.sray[3][q], P.sray[4][q], P.sray[5][q]);
      remaining = P.stmax[q];
      skips = 0;
      sid = P.sscene[q];
      if (!(remaining > 0.004f)) {
        s.best_prim = -1;  // the loop body never runs: visible
        return false;
      }
      return trav_begin(sc.recs[sid], s, o, d, kInf, st);
    };
    auto end = [&](uint32_t, Trav& s) {
      while (true) {
        // intersect_scene's nearest surface: the triangle hit, then any nearer SDF primitive
        float bt = s.best_prim >= 0 ? s.best_t : kInf;
        uint32_t raw_mat = 0u;
        bool hit = s.best_prim >= 0;
        if (hit) raw_mat = static_cast<uint32_t>(__float_as_int(sc.prims[s.best_prim].w));
        if constexpr (kSdf) {
          const int32_t k = sdf_closest(sc, sc.recs[sid], o, d, bt, sst);
          if (k >= 0) {
            hit = true;
            raw_mat = sc.sdfs[k].material;
          }
        }
        if (!hit || bt >= remaining) break;  // nothing in the way
        const PtMaterial& m = material_at(sc, raw_mat);
        if (!(m.material_model == 5u || m.transmission > 0.05f)) return false;  // opaque blocker
        const float advance = bt + 0.004f;
        o = add(o, muls(d, advance));
        remaining -= advance;
        ++skips;
        if (!(skips < 8u && remaining > 0.004f)) break;  // loop exhausted: visible
        if (trav_begin(sc.recs[sid], s, o, d, kInf, st)) return true;
        // the follow-up query ended at once (root box missed): no triangle in the way
        s.best_prim = -1;
        if constexpr (!kSdf) break;
      }
      visible(q);
      return false;
    };
    drive_queries<false>(sc, stk, P.queue, n, P.refill_min, st, begin, end);
  }
  flush_stats(P.counters, st, true);
  if constexpr (kSdf) flush_sdf(P.counters, sst);
}

// Adjoint passes: the radiance and gradient pieces of every shadow record whose light was visible.
constexpr uint32_t kGradBlocks = 1024;
constexpr uint32_t kShadowApplyBlocks = 1024;  // grid-stride blocks of k_shadow_apply
static_assert(kGradBlocks <= kGradRows, "every block of an adjoint kernel owns a partial row");
// Adds the light contribution of every shadow record k_shadow's forward any-hit query marked. A path
// has at most one shadow record per bounce and this runs before the next shade, so every radiance
// sum sees its additions in the same order as when the traversal kernel added them itself.
__global__ __launch_bounds__(kBlock) void k_shadow_apply(DevPaths P, DevScene sc, RadTarget rq) {
  __shared__ uint32_t prefix[kSegments + 1];
  load_prefix(P.queue + kQueueShadowCursor, kSegLine, prefix);
  const uint32_t n = prefix[kSegments];
  for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
    const uint32_t q = queue_pos(prefix, j, P.seg_cap);
    if (P.sscene[q] & kShadowReached) light_visible<false>(P, sc, DevAdjoint{}, nullptr, rq, q);
  }
}

__global__ __launch_bounds__(kBlock) void k_shadow_grad(DevPaths P, DevScene sc, DevAdjoint A, RadTarget rq) {
  __shared__ float glds[kAdjLdsFloats + kTexTableFloats];
  __shared__ uint32_t prefix[kSegments + 1];
  tex_table_init(glds);
  grad_lds_clear(A, glds);
  load_prefix(P.queue + kQueueShadowCursor, kSegLine, prefix);
  const uint32_t n = prefix[kSegments];
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t w0 = blockIdx.x * kBlock + (threadIdx.x & ~63u); w0 < n; w0 += gridDim.x * kBlock) {
    const uint32_t i = w0 + lane;
    if (i < n) {
      const uint32_t q = queue_pos(prefix, i, P.seg_cap);
      if (A.s_light[q] & kShadowVisible) light_visible<true>(P, sc, A, glds, rq, q);
    }
    tex_grad_flush(A, glds);  // the wave is converged here
  }
  grad_lds_flush(A, glds);
}

// FilmBuffer::add_sample for every tile of a pass, after shade_pixel's non-finite drop
// (ScalarCpuPathTracer.Render.cpp:270-283, FilmBuffer.cpp:80-91).
// Film accumulation, one wave (one 64-thread block) per 64 pixels of a job, each lane one pixel: the
// pixel's samples are added in tile order (the reference's sample order), so the sums are those of
// adding one sample per render_tile call. A job that is one interleaved group keeps a pixel's
// samples in consecutive path slots; the wave stages them through LDS in chunks of kFilmChunk samples
// for half its pixels at a time, with coalesced loads (lanes 0-31 and 32-63 each read one pixel's
// 128-byte chunk per instruction), then those pixels' lanes sum their chunks in order. Half the
// pixels per stage keeps the LDS at 12.7 KiB per wave (twelve waves per CU) with whole-line loads.
constexpr uint32_t kFilmChunk = 32;
constexpr uint32_t kFilmHalf = 32;  // pixels staged at a time
__global__ __launch_bounds__(64) void k_film(DevPaths P, DevScene sc, DevFilm film, const DevTile* tiles,
                                            const FilmJob* jobs, uint32_t job_count, uint32_t job_waves) {
  __shared__ float stage[3][kFilmHalf * (kFilmChunk + 1)];
  const uint32_t w = blockIdx.x;
  if (w >= job_waves) return;
  uint32_t lo = 0, hi = job_count;  // last job with wave_begin <= w
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (jobs[mid].wave_begin <= w) lo = mid; else hi = mid;
  }
  const FilmJob J = jobs[lo];
  const uint32_t lane = threadIdx.x;
  const uint32_t jp = J.w * J.h;
  const uint32_t first = (w - J.wave_begin) * 64u;
  const uint32_t local = first + lane;
  const bool valid = local < jp;
  const uint32_t x = J.x0 + local % J.w, y = J.y0 + local / J.w;
  const SceneRec& R = sc.recs[J.scene];
  const uint32_t px = R.film_base + y * R.width + x;
  float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t cnt = 0, added = 0, dropped_rays = 0;
  bool touched = false;
  // Both tallies are updated on every sample (a select, not a branch between the two variables, which
  // the compiler would otherwise keep in a scratch array).
  auto add = [&](float r, float g, float b, uint32_t slot) {
    const bool ok = finite(r) && finite(g) && finite(b);  // shade_pixel drops it, rays included (Render.cpp:270-281)
    uint32_t drop = 0;
    if (!ok) drop = P.ended_rays[slot];
    dropped_rays += drop;
    added += ok ? 1u : 0u;
    if (ok) {
      if (!touched) {
        s = film.sum[px];
        cnt = film.count[px];
        touched = true;
      }
      s.x = s.x + r;
      s.y = s.y + g;
      s.z = s.z + b;
      ++cnt;
    }
  };
  if (J.group > 1u) {
    const uint32_t G = J.group, base = tiles[J.first_tile].path_begin;
    const uint32_t npx = jp - first < 64u ? jp - first : 64u;
    for (uint32_t kc = 0; kc < G; kc += kFilmChunk) {
      const uint32_t kn = G - kc < kFilmChunk ? G - kc : kFilmChunk;
      const uint32_t k = lane & (kFilmChunk - 1u);
      for (uint32_t h = 0; h < 64u; h += kFilmHalf) {  // pixels [h, h + kFilmHalf) of the wave
        for (uint32_t i = 0; i < kFilmHalf; i += 64u / kFilmChunk) {  // 64 / kFilmChunk pixels per load
          const uint32_t j = i + lane / kFilmChunk;
          if (h + j < npx && k < kn) {
            const uint32_t slot = base + (first + h + j) * G + kc + k;
            stage[0][j * (kFilmChunk + 1u) + k] = P.rad[0][slot];
            stage[1][j * (kFilmChunk + 1u) + k] = P.rad[1][slot];
            stage[2][j * (kFilmChunk + 1u) + k] = P.rad[2][slot];
          }
        }
        __syncthreads();
        if (valid && lane - h < kFilmHalf)
          for (uint32_t q = 0; q < kn; ++q) {
            const uint32_t o = (lane - h) * (kFilmChunk + 1u) + q;
            add(stage[0][o], stage[1][o], stage[2][o], base + local * G + kc + q);
          }
        __syncthreads();
      }
    }
  } else if (valid) {
    for (uint32_t k = 0; k < J.tile_count; ++k) {
      const DevTile& t = tiles[J.first_tile + k];
      if (x < t.x || y < t.y || x - t.x >= t.w || y - t.y >= t.h) continue;
      const uint32_t p = tile_slot(t, (y - t.y) * t.w + (x - t.x));
      add(P.rad[0][p], P.rad[1][p], P.rad[2][p], p);
    }
  }
  if (touched) {
    film.sum[px] = s;
    film.count[px] = cnt;
  }
  wave_add(P.counters, kSamples, added);
  // k_extend counted every extension ray; the u64 counter wraps back to the finite samples' rays
  wave_add(P.counters, kRays, 0ull - static_cast<uint64_t>(dropped_rays));
}

// Batched IRayAccelerator::intersect for the ray-query entry points (rays are 6 floats each). The
// output policy stores one finished query: t and primitive arrays, or whole RayQueryHit records.
struct HitArrays {
  float* t;
  int32_t* prim;
  __device__ __forceinline__ void put(uint32_t j, const float*, const Trav& s, const DevScene&) const {
    t[j] = s.best_prim >= 0 ? s.best_t : kInf;
    prim[j] = s.best_prim;
  }
};

// RayQueryHit as CpuBvhAccelerator::intersect fills it (CpuBvhAccelerator.cpp:446-452): position =
// add_vec3(origin, mul_vec3(direction, t)), the primitive's face normal, its instance's material
// index (inst_mat, unclamped) and primitive index; `out = {}` on a miss.
struct HitRecords {
  PtRayQueryHit* out;
  const uint32_t* inst_mat;  // per primitive
  __device__ __forceinline__ void put(uint32_t j, const float* r, const Trav& s, const DevScene& sc) const {
    PtRayQueryHit h = {};
    if (s.best_prim >= 0) {
      const float t = s.best_t;
      const float4 pn = sc.prims[s.best_prim];
      h.hit = 1u;
      h.t = t;
      h.position[0] = r[0] + r[3] * t;
      h.position[1] = r[1] + r[4] * t;
      h.position[2] = r[2] + r[5] * t;
      h.normal[0] = pn.x;
      h.normal[1] = pn.y;
      h.normal[2] = pn.z;
      h.material_index = inst_mat[s.best_prim];
      h.primitive_index = static_cast<uint32_t>(s.best_prim);
    }
    out[j] = h;
  }
};

template <bool kAny, typename Out>
__global__ __launch_bounds__(kBlock) void k_trace_rays(DevPaths P, DevScene sc, const float* rays, const float* tmax,
                                                      uint32_t n, Out out) {
  __shared__ int32_t s_ref[kStackLds * kBlock];
  __shared__ float s_t[kStackLds * kBlock];
  __shared__ float4 s_hot[4 * (kHotNodes > 0 ? kHotNodes : 1)];
  load_hot(sc, s_hot);
  __syncthreads();
  const Stack stk = make_stack(P, sc, s_ref, s_t, s_hot);
  TravStats st = {0, 0, 0, 0, 0};
  auto begin = [&](uint32_t j, Trav& s) {
    const float* r = rays + 6u * static_cast<size_t>(j);
    return trav_begin(sc.recs[0], s, mk(r[0], r[1], r[2]), mk(r[3], r[4], r[5]), tmax ? tmax[j] : kInf, st);
  };
  auto end = [&](uint32_t j, Trav& s) {
    out.put(j, rays + 6u * static_cast<size_t>(j), s, sc);
    return false;
  };
  drive_queries<kAny>(sc, stk, P.queue, n, P.refill_min, st, begin, end);
  flush_stats(P.counters, st, false);
}

}  // namespace

void launch_init_queue(hipStream_t stream, const DevPaths& paths, uint32_t live) {
  hipLaunchKernelGGL(k_init_queue, dim3(1), dim3(64), 0, stream, paths.queue, live, paths.seg_cap);
}

void launch_advance(hipStream_t stream, const DevPaths& paths) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, stream, paths.queue);
}

void launch_zero_counters(hipStream_t stream, unsigned long long* counters) {
  hipLaunchKernelGGL(k_zero_counters, dim3((kCounterWords + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, counters);
}

void launch_clear_film(hipStream_t stream, const DevFilm& film, uint32_t pixels) {
  if (pixels == 0) return;
  hipLaunchKernelGGL(k_clear_film, dim3((pixels + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, film, pixels);
}

void launch_refit_tris(hipStream_t stream, const RefitInst* insts, uint32_t inst_count, uint32_t tri_count,
                       const float* local_v, const uint32_t* local_i, const uint32_t* leaf_pos, float4* tris,
                       float4* prims, float* tri_box) {
  if (tri_count == 0 || inst_count == 0) return;
  hipLaunchKernelGGL(k_refit_tris, dim3((tri_count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, insts,
                     inst_count, tri_count, local_v, local_i, leaf_pos, tris, prims, tri_box);
}

void launch_refit_nodes(hipStream_t stream, float4* nodes, const uint32_t* order, uint32_t begin, uint32_t count,
                        const float* tri_box, uint32_t* nonfinite) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_refit_nodes, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, nodes, order,
                     begin, count, tri_box, nonfinite);
}

void launch_refit_root(hipStream_t stream, const float4* nodes, int32_t root_ref, const float* tri_box,
                       SceneRec* rec) {
  hipLaunchKernelGGL(k_refit_root, dim3(1), dim3(64), 0, stream, nodes, root_ref, tri_box, rec);
}

void launch_hot_refresh(hipStream_t stream, float4* nodes, const uint32_t* src, uint32_t hot_base, uint32_t count) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_hot_refresh, dim3((3u * count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, nodes, src,
                     hot_base, count);
}

void launch_resolve(hipStream_t stream, const DevFilm& film, uint32_t pixels, const DevResolve* rs, void* out) {
  if (pixels == 0) return;
  const dim3 grid((pixels + kBlock - 1) / kBlock);
  if (rs)
    hipLaunchKernelGGL(k_resolve<true>, grid, dim3(kBlock), 0, stream, film, pixels, *rs, nullptr,
                       static_cast<uint32_t*>(out));
  else
    hipLaunchKernelGGL(k_resolve<false>, grid, dim3(kBlock), 0, stream, film, pixels, DevResolve{},
                       static_cast<float*>(out), nullptr);
}

void launch_raygen(hipStream_t stream, const DevPaths& paths, const DevScene& scene, const DevSettings& st,
                   const DevTile* tiles, uint32_t tile_count, uint32_t path_count) {
  if (path_count == 0) return;
  hipLaunchKernelGGL(k_raygen<false>, dim3((path_count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, paths,
                     paths.rays[0], scene, st, tiles, tile_count, path_count, DevAdjoint{});
}

void launch_raygen_adjoint(hipStream_t stream, const DevPaths& paths, const DevAdjoint& adj, const DevScene& scene,
                           const DevSettings& st, const DevTile* tiles, uint32_t tile_count, uint32_t path_count) {
  if (path_count == 0) return;
  hipLaunchKernelGGL(k_raygen<true>, dim3((path_count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, paths,
                     paths.rays[0], scene, st, tiles, tile_count, path_count, adj);
}

void launch_grad_clear(hipStream_t stream, const DevAdjoint& adj) {
  if (!adj.use_lds) return;
  const uint32_t count = adj.partial_rows * (7u * adj.nmat + 3u * adj.nlight);
  hipLaunchKernelGGL(k_grad_clear, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, adj, count);
}

void launch_grad_reduce(hipStream_t stream, const DevAdjoint& adj) {
  if (!adj.use_lds) return;
  const uint32_t n = 7u * adj.nmat + 3u * adj.nlight;
  hipLaunchKernelGGL(k_grad_reduce, dim3((n + kBlock - 1) / kBlock, kReduceGroups), dim3(kBlock), 0, stream, adj);
}

void launch_extend(hipStream_t stream, const DevPaths& paths, const DevScene& scene, int in_buf, uint32_t grid) {
  hipLaunchKernelGGL(k_extend, dim3(grid), dim3(kBlock), 0, stream, paths, paths.rays[in_buf], scene);
  if (scene.has_sdf)
    hipLaunchKernelGGL(k_sdf_extend, dim3(grid), dim3(kBlock), 0, stream, paths, paths.rays[in_buf], scene);
}

void launch_shade(hipStream_t stream, const DevPaths& paths, const DevScene& scene, const DevSettings& st, int in_buf,
                  uint32_t depth, uint32_t grid) {
  hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kBlock), 0, stream, paths, paths.rays[in_buf],
                     paths.rays[in_buf ^ 1], scene, st, depth);
}

void launch_shade_adjoint(hipStream_t stream, const DevPaths& paths, const DevAdjoint& adj, const DevScene& scene,
                          const DevSettings& st, int in_buf, uint32_t depth, uint32_t grid) {
  grid = grid < kGradRows ? grid : kGradRows;  // one partial row per block
  hipLaunchKernelGGL(k_shade_adj, dim3(grid), dim3(kBlock), 0, stream, paths, paths.rays[in_buf],
                     paths.rays[in_buf ^ 1], scene, st, depth, adj, static_cast<uint32_t>(in_buf));
}

void launch_shadow(hipStream_t stream, const DevPaths& paths, const DevScene& scene, int in_buf, uint32_t grid) {
  const DevRays& o = paths.rays[in_buf ^ 1];
  const RadTarget rq{{o.rad[0], o.rad[1], o.rad[2]}};
  if (scene.has_sdf)
    hipLaunchKernelGGL((k_shadow<false, true, false>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, DevAdjoint{},
                       rq);
  else if (scene.has_transmissive)
    hipLaunchKernelGGL((k_shadow<false, false, false>), dim3(grid), dim3(kBlock), 0, stream, paths, scene,
                       DevAdjoint{}, rq);
  else {
    hipLaunchKernelGGL((k_shadow<false, false, true>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, DevAdjoint{},
                       rq);
    hipLaunchKernelGGL(k_shadow_apply, dim3(grid < kShadowApplyBlocks ? grid : kShadowApplyBlocks), dim3(kBlock), 0,
                       stream, paths, scene, rq);
  }
}

void launch_shadow_adjoint(hipStream_t stream, const DevPaths& paths, const DevAdjoint& adj, const DevScene& scene,
                           int in_buf, uint32_t grid) {
  const DevRays& o = paths.rays[in_buf ^ 1];
  const RadTarget rq{{o.rad[0], o.rad[1], o.rad[2]}};
  if (scene.has_sdf)
    hipLaunchKernelGGL((k_shadow<true, true, false>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, adj, rq);
  else if (scene.has_transmissive)
    hipLaunchKernelGGL((k_shadow<true, false, false>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, adj, rq);
  else
    hipLaunchKernelGGL((k_shadow<true, false, true>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, adj, rq);
  hipLaunchKernelGGL(k_shadow_grad, dim3(grid < kGradBlocks ? grid : kGradBlocks), dim3(kBlock), 0, stream, paths,
                     scene, adj, rq);
}

void launch_film(hipStream_t stream, const DevPaths& paths, const DevScene& scene, const DevFilm& film,
                 const DevTile* tiles, const FilmJob* jobs, uint32_t job_count, uint32_t job_waves) {
  if (job_waves == 0) return;
  hipLaunchKernelGGL(k_film, dim3(job_waves), dim3(64), 0, stream, paths, scene, film, tiles, jobs, job_count,
                     job_waves);
}

void launch_trace_rays(hipStream_t stream, const DevPaths& paths, const DevScene& scene, int any_hit,
                       const float* rays, const float* tmax, uint32_t count, float* out_t, int32_t* out_prim,
                       uint32_t grid) {
  const HitArrays out{out_t, out_prim};
  if (any_hit)
    hipLaunchKernelGGL((k_trace_rays<true, HitArrays>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, rays, tmax,
                       count, out);
  else
    hipLaunchKernelGGL((k_trace_rays<false, HitArrays>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, rays, tmax,
                       count, out);
}

void launch_intersect(hipStream_t stream, const DevPaths& paths, const DevScene& scene, int any_hit, const float* rays,
                      const float* tmax, uint32_t count, PtRayQueryHit* hits, const uint32_t* inst_mat, uint32_t grid) {
  const HitRecords out{hits, inst_mat};
  if (any_hit)
    hipLaunchKernelGGL((k_trace_rays<true, HitRecords>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, rays, tmax,
                       count, out);
  else
    hipLaunchKernelGGL((k_trace_rays<false, HitRecords>), dim3(grid), dim3(kBlock), 0, stream, paths, scene, rays,
                       tmax, count, out);
}

static uint32_t occupancy_grid(int device, const void* fn) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, kBlock, 0) != hipSuccess || blocks <= 0) blocks = 4;
  return static_cast<uint32_t>(cus) * static_cast<uint32_t>(blocks);
}

// ------------------------------------------------------------------ allocation-free ray queries
// pt_hip_trace_rays_noalloc keeps all of its traversal scratch in one static device array (each GPU
// holds its own copy from module load until the process ends): the counter stripes, the queue block
// (chunk cursor and waves-done count), the error flag, and the stack spill area of a grid of at most
// kNoAllocBlocks workgroups with kNoAllocSpill spilled entries per lane (one int32 reference array, then
// one float t_near array).
constexpr size_t kNoAllocQueueOff = sizeof(unsigned long long) * kCounterWords;          // 8 KiB of counters first
constexpr size_t kNoAllocErrorOff = kNoAllocQueueOff + ((sizeof(uint32_t) * kQueueWords + 255u) & ~size_t(255u));
constexpr size_t kNoAllocSpillOff = 32768;
constexpr size_t kNoAllocLanes = static_cast<size_t>(kNoAllocBlocks) * kBlock;
constexpr size_t kNoAllocSpillEntries = static_cast<size_t>(kNoAllocSpill) * kNoAllocLanes;
constexpr size_t kNoAllocBytes = kNoAllocSpillOff + 8u * kNoAllocSpillEntries;
static_assert(kNoAllocErrorOff + 256u <= kNoAllocSpillOff, "scratch header overlaps the spill area");
#if PT_STACK_LDS == 8  // the default ring
static_assert(kNoAllocBytes == 32768u + 26u * 1024u * 1024u, "the documented scratch size is 26 MiB + 32 KiB");
#endif

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// Clears what the queries read before writing: counters, queue block, error flag.
__global__ __launch_bounds__(kBlock) void k_noalloc_clear(unsigned long long* counters, uint32_t* queue,
                                                          uint32_t* error) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < static_cast<uint32_t>(kCounterWords)) counters[i] = 0u;
  if (i < static_cast<uint32_t>(kQueueWords)) queue[i] = 0u;
  if (i == 0u) *error = 0u;
}

constexpr int kMaxNoAllocDevices = 64;
std::mutex g_noalloc_mutex;
unsigned char* g_noalloc_base[kMaxNoAllocDevices] = {};
hipError_t g_noalloc_status[kMaxNoAllocDevices] = {};
bool g_noalloc_tried[kMaxNoAllocDevices] = {};

}  // namespace

hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxNoAllocDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_tried[device]) {
    // `scratch` names a host-side stand-in; the device copy's address comes from the runtime, per GPU
    void* p = nullptr;
    g_noalloc_status[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_noalloc_base[device] = g_noalloc_status[device] == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
    g_noalloc_tried[device] = true;
  }
  return g_noalloc_status[device];
}

NoAllocScratch noalloc_scratch(int device) {
  NoAllocScratch s{};
  if (device < 0 || device >= kMaxNoAllocDevices) return s;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  unsigned char* b = g_noalloc_base[device];
  if (!b) return s;
  s.counters = reinterpret_cast<unsigned long long*>(b);
  s.queue = reinterpret_cast<uint32_t*>(b + kNoAllocQueueOff);
  s.error = reinterpret_cast<uint32_t*>(b + kNoAllocErrorOff);
  s.stack_ref = reinterpret_cast<int32_t*>(b + kNoAllocSpillOff);
  s.stack_t = reinterpret_cast<float*>(b + kNoAllocSpillOff + 4u * kNoAllocSpillEntries);
  return s;
}

void launch_trace_rays_noalloc(hipStream_t stream, const NoAllocScratch& sc, uint32_t refill_min, const DevScene& scene,
                               int any_hit, const float* rays, const float* tmax, uint32_t count, float* out_t,
                               int32_t* out_prim, uint32_t grid) {
  grid = grid < kNoAllocBlocks ? grid : kNoAllocBlocks;
  DevPaths P{};
  P.queue = sc.queue;
  P.counters = sc.counters;
  P.error = sc.error;
  P.stack_ref = sc.stack_ref;
  P.stack_t = sc.stack_t;
  P.stack_overflow_depth = kNoAllocSpill;  // entry i of lane g at i * stack_threads + g < kNoAllocSpillEntries
  P.stack_threads = grid * kBlock;
  P.refill_min = refill_min;
  constexpr uint32_t clear_words = kQueueWords > kCounterWords ? kQueueWords : kCounterWords;
  hipLaunchKernelGGL(k_noalloc_clear, dim3((clear_words + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, sc.counters,
                     sc.queue, sc.error);
  if (count) launch_trace_rays(stream, P, scene, any_hit, rays, tmax, count, out_t, out_prim, grid);
}

Grids resident_grids(int device) {
  Grids g;
  g.extend = occupancy_grid(device, reinterpret_cast<const void*>(&k_extend));
  g.shadow = occupancy_grid(device, reinterpret_cast<const void*>(&k_shadow<false, false, true>));
  g.shadow_adj = occupancy_grid(device, reinterpret_cast<const void*>(&k_shadow<true, false, true>));
  const void* queries[4] = {reinterpret_cast<const void*>(&k_trace_rays<false, HitArrays>),
                            reinterpret_cast<const void*>(&k_trace_rays<true, HitArrays>),
                            reinterpret_cast<const void*>(&k_trace_rays<false, HitRecords>),
                            reinterpret_cast<const void*>(&k_trace_rays<true, HitRecords>)};
  g.query = ~0u;
  for (const void* fn : queries) {
    const uint32_t q = occupancy_grid(device, fn);
    g.query = q < g.query ? q : g.query;
  }
  g.shade = occupancy_grid(device, reinterpret_cast<const void*>(&k_shade));
  g.shade_adj = occupancy_grid(device, reinterpret_cast<const void*>(&k_shade_adj));
  return g;
}

}  // namespace ptd
