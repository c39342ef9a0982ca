/*
 * ORACLE — test infrastructure only (see pto.h). CPU restatement of the
 * reference's scalar path-tracing hot path. Each function cites the reference
 * file:line it follows. Compile with -ffp-contract=off: every float expression
 * is evaluated in the operand order the reference's C++ evaluates it.
 *
 * PTO_REAL    float (default) or double (the fp64 oracle)
 * PTO_LIBM    1: glibc sinf/cosf/powf like the reference; 0: pto_math.h
 * PTO_AVX2    1: leaves are intersected 8 triangles at a time with AVX2 (fp32 only; the CPU
 *             baseline's AVX2 mode, as the reference's leaf batch intersector when built with
 *             PT_ENABLE_AVX2, BvhTriangleIntersector.cpp:156-171). Same results as the scalar loop.
 *
 * How each part is pinned (the reference cannot be built here and ships no vectors for this path):
 *   splitmix64 / Rng / next01 ........ published SplitMix64 known answers (tests/test_oracle_kat.py)
 *   sin cos pow atan2 acos ............ double-precision values, and the glibc build (test_oracle_kat.py,
 *                                        test_oracle_integrator.py: portable vs glibc on every config)
 *   Moller-Trumbore, traversal ........ hand cases, brute force over all triangles (test_bvh_host.py)
 *   midpoint / SAH builders ........... invariants, costs (test_bvh_host.py)
 *   camera_rays, RNG key .............. pinhole geometry, key layout (test_oracle_integrator.py)
 *   trace at depth 1 / 2, NEE, spot,
 *   MIS with lights[0], light pdf,
 *   roulette, samplers, effects, env .. closed forms, sampler moments, a second transcription of
 *                                        the effects (test_oracle_integrator.py)
 *   trace at depth 8, every branch .... a second, float64 transcription of camera_rays, trace,
 *                                        sample_direct_light and the shadow skip loop: every sample
 *                                        within 1e-4 (test_oracle_trace_transcription.py)
 *   dropped samples and ray counts .... the reference's counting rule (test_oracle_counters.py)
 *   film resolve ...................... closed forms and the libm build (test_oracle_resolve.py)
 *   adjoint ........................... finite differences (test_oracle_adjoint.py, test_oracle_textures.py)
 */
#include "pto.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "pto_math.h"

#ifndef PTO_REAL_IS_DOUBLE
#define PTO_REAL_IS_DOUBLE 0
#endif
#ifndef PTO_LIBM
#define PTO_LIBM 0
#endif
#ifndef PTO_AVX2
#define PTO_AVX2 0
#endif
#if PTO_AVX2
#if PTO_REAL_IS_DOUBLE
#error "PTO_AVX2 is an fp32 build"
#endif
#include <immintrin.h>
#endif

#if PTO_REAL_IS_DOUBLE
typedef double real;
#define RC(x) (x)
#define R_SQRT(x) sqrt(x)
#define R_SIN(x) sin(x)
#define R_COS(x) cos(x)
#define R_POW(x, y) pow((x), (y))
#define R_FABS(x) fabs(x)
#define R_MAXV DBL_MAX
#else
typedef float real;
#define RC(x) (x##f)
#define R_SQRT(x) sqrtf(x)
#if PTO_LIBM
#define R_SIN(x) sinf(x)
#define R_COS(x) cosf(x)
#define R_POW(x, y) powf((x), (y))
#else
#define R_SIN(x) pm_sinf(x)
#define R_COS(x) pm_cosf(x)
#define R_POW(x, y) pm_powf((x), (y))
#endif
#define R_FABS(x) fabsf(x)
#define R_MAXV FLT_MAX
#endif

/* std::min / std::max semantics: min(a,b) = (b < a) ? b : a; max(a,b) = (a < b) ? b : a. */
static inline real smin(real a, real b) { return (b < a) ? b : a; }
static inline real smax(real a, real b) { return (a < b) ? b : a; }

static const real kEps = RC(1e-4);                      /* ScalarCpuPathTracer.cpp:16 */
static const real kPi = RC(3.14159265358979323846);    /* :14 */
static const real kInvPi = RC(0.31830988618);          /* :15 */

typedef struct { real x, y, z; } v3;

static inline v3 mk(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 muls(v3 a, real s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 mulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 divs(v3 a, real s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline real dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real length_sq(v3 a) { return dot(a, a); }
static inline real length(v3 a) { return R_SQRT(length_sq(a)); }
/* normalize(): ScalarCpuPathTracer.cpp:104-110 (divide by length, l <= eps -> +Y). */
static inline v3 normalize(v3 a) {
  const real l = length(a);
  if (l <= kEps) return mk(RC(0.0), RC(1.0), RC(0.0));
  return divs(a, l);
}
/* normalize_vec3(): CpuBvhAccelerator.cpp:76-82 (multiply by 1/sqrt). */
static inline v3 normalize_accel(v3 a) {
  const real len_sq = dot(a, a);
  if (len_sq <= kEps * kEps) return mk(RC(0.0), RC(1.0), RC(0.0));
  return muls(a, RC(1.0) / R_SQRT(len_sq));
}
static inline real vmax3(v3 a) { return smax(a.x, smax(a.y, a.z)); }
static inline int is_finite_r(real v) { return isfinite(v); }
/* clamp01: ScalarCpuPathTracer.cpp:161-166 */
static inline real clamp01(real v) {
  if (!is_finite_r(v)) return RC(0.0);
  return smin(RC(1.0), smax(RC(0.0), v));
}
/* std::clamp(v, lo, hi) */
static inline real sclamp(real v, real lo, real hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

static v3 v3f(const float* p) { return mk((real)p[0], (real)p[1], (real)p[2]); }

/* ------------------------------------------------------------------ math */
uint32_t pto_real_bytes(void) { return (uint32_t)sizeof(real); }
uint32_t pto_uses_libm(void) { return PTO_LIBM || PTO_REAL_IS_DOUBLE; }
float pto_sinf(float x) { return (float)R_SIN((real)x); }
float pto_cosf(float x) { return (float)R_COS((real)x); }
float pto_powf(float x, float y) { return (float)R_POW((real)x, (real)y); }

float pto_atan2f(float y, float x) {
#if PTO_REAL_IS_DOUBLE
  return (float)atan2((double)y, (double)x);
#elif PTO_LIBM
  return atan2f(y, x);
#else
  return pm_atan2f(y, x);
#endif
}
float pto_acosf(float x) {
#if PTO_REAL_IS_DOUBLE
  return (float)acos((double)x);
#elif PTO_LIBM
  return acosf(x);
#else
  return pm_acosf(x);
#endif
}

uint32_t pto_sizeof(int which) {
  switch (which) {
    case 0: return sizeof(PtoMaterial);
    case 1: return sizeof(PtoLight);
    case 2: return sizeof(PtoInstance);
    case 3: return sizeof(PtoCamera);
    case 4: return sizeof(PtoScene);
    case 5: return sizeof(PtoSettings);
    case 6: return sizeof(PtoCounters);
    case 7: return sizeof(PtoTextureDesc);
    default: return 0;
  }
}

/* ------------------------------------------------------------------ RNG */
/* splitmix64: ScalarCpuPathTracer.cpp:36-41 */
uint64_t pto_splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}
#define sm pto_splitmix64

/* Rng::Rng(SampleKey): ScalarCpuPathTracer.cpp:383-388 */
uint64_t pto_rng_seed(uint64_t seed, uint64_t frame, uint64_t sample, uint64_t path_id,
                      uint64_t dimension, uint64_t depth, uint64_t pixel) {
  const uint64_t mix = sm(seed ^ sm(frame) ^ sm(sample) ^ sm(path_id) ^ sm(dimension) ^ sm(depth));
  return sm(mix + sm(pixel));
}

/* Rng::next01: ScalarCpuPathTracer.cpp:390-393 */
static inline real next01(uint64_t* state) {
  *state = sm(*state);
  return (real)((*state >> 8) & 0x00ffffffu) * (RC(1.0) / RC(16777215.0));
}

void pto_rng_draws(uint64_t state, uint32_t n, float* out, uint64_t* out_state) {
  for (uint32_t i = 0; i < n; ++i) out[i] = (float)next01(&state);
  if (out_state) *out_state = state;
}

/* ------------------------------------------------------------------ scene flattening + BVH */
typedef struct {
  real mn[3], mx[3];
} aabb_t;

typedef struct {
  aabb_t box;      /* padded by kEps, as CpuBvhAccelerator's make_triangle_aabb */
  int32_t left, right, first, count;
} node_t;

typedef struct {
  uint32_t material_index, primitive_index;
  uint32_t vi[3]; /* vertex indices into PtoScene.vertices; ~0u for instance-local geometry */
  v3 v0, e1, e2, normal;
} prim_t;

struct pto_accel {
  prim_t* prims;
  uint32_t prim_count;
  node_t* nodes;
  uint32_t node_count;
  uint32_t* order; /* leaf order -> primitive index */
  uint32_t max_depth;
  /* textures (pto_accel_set_textures): texture coordinates of each primitive's three vertices
   * (NULL: none), texels (rgb) and texture records */
  float* prim_uv;
  float* texels;
  uint64_t texel_count;
  float* soa;         /* PTO_AVX2: per leaf batch of 8, v0 xyz, e1 xyz, e2 xyz as 9 x 8 floats */
  uint32_t* soa_base; /* PTO_AVX2: per node, its first batch in soa (leaves only) */
  PtoTextureDesc* textures;
  uint32_t texture_count;
  PtoSdf* sdfs; /* pto_accel_set_sdfs: the SDF primitives intersect_scene tests after the triangles */
  uint32_t sdf_count;
};

/* rotate_quat / transform_instance_vertex: CpuBvhAccelerator.cpp:48-74 */
static v3 rotate_quat(v3 value, const float* q) {
  const real qx = (real)q[0], qy = (real)q[1], qz = (real)q[2], qw0 = (real)q[3];
  const real len_sq = qx * qx + qy * qy + qz * qz + qw0 * qw0;
  if (len_sq <= kEps * kEps) return value;
  const real inv_len = RC(1.0) / R_SQRT(len_sq);
  const v3 qv = mk(qx * inv_len, qy * inv_len, qz * inv_len);
  const real qw = qw0 * inv_len;
  const v3 t = muls(cross(qv, value), RC(2.0));
  return add(add(value, muls(t, qw)), cross(qv, t));
}
static v3 transform_instance_vertex(v3 local, const PtoInstance* inst) {
  const v3 scaled = mk(local.x * (real)inst->scale[0], local.y * (real)inst->scale[1], local.z * (real)inst->scale[2]);
  return add(rotate_quat(scaled, inst->rotation), v3f(inst->translation));
}

/* make_triangle_aabb: CpuBvhAccelerator.cpp:84-95. std::min({a,b,c}) picks the first smallest. */
static inline real min3(real a, real b, real c) { real m = a; if (b < m) m = b; if (c < m) m = c; return m; }
static inline real max3(real a, real b, real c) { real m = a; if (m < b) m = b; if (m < c) m = c; return m; }

/* Flatten instances into primitives: CpuBvhAccelerator::build, CpuBvhAccelerator.cpp:161-241. */
static int flatten(const PtoScene* sc, prim_t** out_prims, uint32_t* out_n, v3** raw /* 3 per prim, unpadded vertices */) {
  uint64_t cap = 0;
  for (uint32_t i = 0; i < sc->instance_count; ++i) cap += sc->instances[i].triangle_count;
  prim_t* prims = (prim_t*)malloc(sizeof(prim_t) * (cap ? cap : 1));
  v3* verts = (v3*)malloc(sizeof(v3) * 3 * (cap ? cap : 1));
  if (!prims || !verts) { free(prims); free(verts); return -1; }
  uint32_t n = 0;
  for (uint32_t ii = 0; ii < sc->instance_count; ++ii) {
    const PtoInstance* inst = &sc->instances[ii];
    const int use_local = (inst->flags & 1u) != 0u && inst->local_vertex_count > 0u && inst->local_index_count > 0u;
    for (uint32_t tri = 0; tri < inst->triangle_count; ++tri) {
      v3 a = mk(0, 0, 0), b = a, c = a;
      int loaded = 0;
      uint32_t vi[3] = {~0u, ~0u, ~0u};
      if (use_local) {
        const uint32_t local_base = inst->local_first_index + tri * 3u;
        const uint32_t local_end = inst->local_first_vertex + inst->local_vertex_count;
        if ((uint64_t)local_base + 2u < sc->local_index_count) {
          const uint32_t li0 = sc->local_indices[local_base + 0u];
          const uint32_t li1 = sc->local_indices[local_base + 1u];
          const uint32_t li2 = sc->local_indices[local_base + 2u];
          const uint32_t f = inst->local_first_vertex;
          if (f + li0 < local_end && f + li1 < local_end && f + li2 < local_end &&
              f + li0 < sc->local_vertex_count && f + li1 < sc->local_vertex_count && f + li2 < sc->local_vertex_count) {
            a = transform_instance_vertex(v3f(sc->local_vertices + 3u * (f + li0)), inst);
            b = transform_instance_vertex(v3f(sc->local_vertices + 3u * (f + li1)), inst);
            c = transform_instance_vertex(v3f(sc->local_vertices + 3u * (f + li2)), inst);
            loaded = 1;
          }
        }
      }
      if (!loaded) {
        const uint64_t base_tri = ((uint64_t)inst->first_triangle + tri) * 3u;
        if (base_tri + 2u >= sc->index_count) continue;
        const uint32_t i0 = sc->indices[base_tri], i1 = sc->indices[base_tri + 1], i2 = sc->indices[base_tri + 2];
        if (i0 >= sc->vertex_count || i1 >= sc->vertex_count || i2 >= sc->vertex_count) continue;
        a = v3f(sc->vertices + 3u * i0);
        b = v3f(sc->vertices + 3u * i1);
        c = v3f(sc->vertices + 3u * i2);
        vi[0] = i0;
        vi[1] = i1;
        vi[2] = i2;
      }
      prim_t* p = &prims[n];
      p->material_index = inst->material_index;
      p->primitive_index = n;
      p->vi[0] = vi[0];
      p->vi[1] = vi[1];
      p->vi[2] = vi[2];
      p->v0 = a;
      p->e1 = sub(b, a);
      p->e2 = sub(c, a);
      p->normal = normalize_accel(cross(p->e1, p->e2));
      verts[3 * n + 0] = a;
      verts[3 * n + 1] = b;
      verts[3 * n + 2] = c;
      ++n;
    }
  }
  *out_prims = prims;
  *out_n = n;
  *raw = verts;
  return 0;
}

static aabb_t padded_box(const v3* t) {
  aabb_t b;
  b.mn[0] = min3(t[0].x, t[1].x, t[2].x) - kEps;
  b.mn[1] = min3(t[0].y, t[1].y, t[2].y) - kEps;
  b.mn[2] = min3(t[0].z, t[1].z, t[2].z) - kEps;
  b.mx[0] = max3(t[0].x, t[1].x, t[2].x) + kEps;
  b.mx[1] = max3(t[0].y, t[1].y, t[2].y) + kEps;
  b.mx[2] = max3(t[0].z, t[1].z, t[2].z) + kEps;
  return b;
}

/* ---- midpoint builder: ParallelBvhBuilder.cpp:14-163 (deterministic: stable_partition) */
typedef struct {
  const aabb_t* boxes;
  node_t* nodes;
  uint32_t* idx_base;
  uint32_t* tmp;
  int32_t alloc;
  uint32_t leaf;
  uint32_t max_depth;
} mid_ctx;

static void mid_build(mid_ctx* c, int32_t ni, uint32_t* idx, uint32_t count, uint32_t depth) {
  if (depth > c->max_depth) c->max_depth = depth;
  aabb_t b;
  for (int k = 0; k < 3; ++k) { b.mn[k] = R_MAXV; b.mx[k] = -R_MAXV; }
  for (uint32_t i = 0; i < count; ++i) {
    const aabb_t* a = &c->boxes[idx[i]];
    for (int k = 0; k < 3; ++k) { b.mn[k] = smin(b.mn[k], a->mn[k]); b.mx[k] = smax(b.mx[k], a->mx[k]); }
  }
  c->nodes[ni].box = b;
  if (count <= c->leaf) {
    c->nodes[ni].left = -1;
    c->nodes[ni].right = -1;
    c->nodes[ni].first = (int32_t)(idx - c->idx_base);
    c->nodes[ni].count = (int32_t)count;
    return;
  }
  aabb_t cb;
  for (int k = 0; k < 3; ++k) { cb.mn[k] = R_MAXV; cb.mx[k] = -R_MAXV; }
  for (uint32_t i = 0; i < count; ++i) {
    const aabb_t* a = &c->boxes[idx[i]];
    for (int k = 0; k < 3; ++k) {
      const real ck = RC(0.5) * (a->mn[k] + a->mx[k]);
      cb.mn[k] = smin(cb.mn[k], ck);
      cb.mx[k] = smax(cb.mx[k], ck);
    }
  }
  int axis = 0;
  real best = cb.mx[0] - cb.mn[0];
  for (int k = 1; k < 3; ++k) {
    const real e = cb.mx[k] - cb.mn[k];
    if (e > best) { best = e; axis = k; }
  }
  const real split = RC(0.5) * (cb.mn[axis] + cb.mx[axis]);
  uint32_t nl = 0, nr = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const aabb_t* a = &c->boxes[idx[i]];
    const real ck = RC(0.5) * (a->mn[axis] + a->mx[axis]);
    if (ck < split) idx[nl++] = idx[i]; else c->tmp[nr++] = idx[i];
  }
  memcpy(idx + nl, c->tmp, sizeof(uint32_t) * nr);
  const uint32_t left_count = nl, right_count = count - nl;
  const uint32_t safe_left = (left_count == 0 || right_count == 0) ? count / 2 : left_count;
  const int32_t li = c->alloc++;
  const int32_t ri = c->alloc++;
  c->nodes[ni].left = li;
  c->nodes[ni].right = ri;
  c->nodes[ni].first = -1;
  c->nodes[ni].count = 0;
  mid_build(c, li, idx, safe_left, depth + 1);
  mid_build(c, ri, idx + safe_left, count - safe_left, depth + 1);
}

/* ---- bucketed SAH builder: D3D12GpuPathTracer.SceneBvh.cpp:61-65 (bvh_sa), :73-238 (bvh_build).
 * Split decisions use unpadded triangle bounds and centroids (D3D12GpuPathTracer.Scene.cpp:1083-1094).
 * Node bounds are stored padded by kEps so traversal follows CpuBvhAccelerator's contract
 * (union of padded triangle boxes == padded union, since rounding is monotonic). */
typedef struct { uint32_t orig; real c[3], mn[3], mx[3]; } ref_t;

typedef struct {
  ref_t* refs;
  ref_t* tmp;
  node_t* nodes;
  uint32_t node_count;
  uint32_t* order;
  uint32_t order_n;
  uint32_t leaf, buckets;
  uint32_t max_depth;
} sah_ctx;

static real bvh_sa(const real* mn, const real* mx) {
  real dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
  if (dx < 0) dx = 0;
  if (dy < 0) dy = 0;
  if (dz < 0) dz = 0;
  return RC(2.0) * (dx * dy + dy * dz + dz * dx);
}

static uint32_t sah_build(sah_ctx* c, uint32_t start, uint32_t count, uint32_t depth) {
  if (depth > c->max_depth) c->max_depth = depth;
  const uint32_t ni = c->node_count++;
  real bmn[3] = {RC(1e30), RC(1e30), RC(1e30)}, bmx[3] = {RC(-1e30), RC(-1e30), RC(-1e30)};
  for (uint32_t i = start; i < start + count; ++i) {
    for (int k = 0; k < 3; ++k) {
      bmn[k] = smin(bmn[k], c->refs[i].mn[k]);
      bmx[k] = smax(bmx[k], c->refs[i].mx[k]);
    }
  }
  node_t* node = &c->nodes[ni];
  for (int k = 0; k < 3; ++k) { node->box.mn[k] = bmn[k] - kEps; node->box.mx[k] = bmx[k] + kEps; }
  if (count <= c->leaf) {
    node->left = -1;
    node->right = -1;
    node->first = (int32_t)c->order_n;
    node->count = (int32_t)count;
    for (uint32_t i = start; i < start + count; ++i) c->order[c->order_n++] = c->refs[i].orig;
    return ni;
  }
  const int B = (int)(c->buckets < 2u ? 2u : (c->buckets > 16u ? 16u : c->buckets));
  const real node_sa = smax(RC(1e-9), bvh_sa(bmn, bmx));
  int best_axis = 0;
  real best_split = RC(0.0);
  real best_cost = RC(1e30);
  int found = 0;
  for (int axis = 0; axis < 3; ++axis) {
    real cmin = RC(1e30), cmax = RC(-1e30);
    for (uint32_t i = start; i < start + count; ++i) {
      cmin = smin(cmin, c->refs[i].c[axis]);
      cmax = smax(cmax, c->refs[i].c[axis]);
    }
    const real extent = cmax - cmin;
    if (extent < RC(1e-6)) continue;
    real bmn_b[16][3], bmx_b[16][3];
    uint32_t bn[16];
    for (int b = 0; b < B; ++b) {
      bn[b] = 0;
      for (int k = 0; k < 3; ++k) { bmn_b[b][k] = RC(1e30); bmx_b[b][k] = RC(-1e30); }
    }
    for (uint32_t i = start; i < start + count; ++i) {
      /* the reference's unchecked int conversion is undefined for a NaN or infinite value (a
       * non-finite vertex): those primitives take bucket 0, or the last for +inf, as the product does */
      real x = (c->refs[i].c[axis] - cmin) / extent * (real)B;
      if (!(x >= RC(0.0))) x = RC(0.0);
      int b = x < (real)B ? (int)x : B - 1;
      if (B - 1 < b) b = B - 1;
      bn[b]++;
      for (int k = 0; k < 3; ++k) {
        bmn_b[b][k] = smin(bmn_b[b][k], c->refs[i].mn[k]);
        bmx_b[b][k] = smax(bmx_b[b][k], c->refs[i].mx[k]);
      }
    }
    for (int s = 1; s < B; ++s) {
      real lmn[3] = {RC(1e30), RC(1e30), RC(1e30)}, lmx[3] = {RC(-1e30), RC(-1e30), RC(-1e30)};
      real rmn[3] = {RC(1e30), RC(1e30), RC(1e30)}, rmx[3] = {RC(-1e30), RC(-1e30), RC(-1e30)};
      uint32_t lc = 0, rc = 0;
      for (int b = 0; b < s; ++b) {
        if (!bn[b]) continue;
        lc += bn[b];
        for (int k = 0; k < 3; ++k) { lmn[k] = smin(lmn[k], bmn_b[b][k]); lmx[k] = smax(lmx[k], bmx_b[b][k]); }
      }
      for (int b = s; b < B; ++b) {
        if (!bn[b]) continue;
        rc += bn[b];
        for (int k = 0; k < 3; ++k) { rmn[k] = smin(rmn[k], bmn_b[b][k]); rmx[k] = smax(rmx[k], bmx_b[b][k]); }
      }
      if (!lc || !rc) continue;
      const real cost = (bvh_sa(lmn, lmx) * (real)lc + bvh_sa(rmn, rmx) * (real)rc) / node_sa;
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = axis;
        best_split = cmin + (real)s / (real)B * extent;
        found = 1;
      }
    }
  }
  uint32_t mid;
  if (found) {
    uint32_t nl = 0, nr = 0;
    for (uint32_t i = start; i < start + count; ++i) {
      if (c->refs[i].c[best_axis] < best_split) c->refs[start + nl++] = c->refs[i];
      else c->tmp[nr++] = c->refs[i];
    }
    memcpy(c->refs + start + nl, c->tmp, sizeof(ref_t) * nr);
    mid = start + nl;
    if (mid <= start || mid >= start + count) mid = start + count / 2u;
  } else {
    mid = start + count / 2u;
  }
  const uint32_t l = sah_build(c, start, mid - start, depth + 1);
  const uint32_t r = sah_build(c, mid, start + count - mid, depth + 1);
  c->nodes[ni].left = (int32_t)l;
  c->nodes[ni].right = (int32_t)r;
  c->nodes[ni].first = -1;
  c->nodes[ni].count = 0;
  return ni;
}

#if PTO_AVX2
/* structure-of-arrays copies of every leaf's triangles, 8 per batch, unused lanes zero */
static void soa_fill(pto_accel* a) {
  free(a->soa);
  free(a->soa_base);
  uint64_t batches = 0;
  a->soa_base = (uint32_t*)calloc(a->node_count ? a->node_count : 1u, sizeof(uint32_t));
  for (uint32_t i = 0; i < a->node_count; ++i)
    if (a->nodes[i].count > 0) {
      a->soa_base[i] = (uint32_t)batches;
      batches += ((uint32_t)a->nodes[i].count + 7u) / 8u;
    }
  a->soa = (float*)aligned_alloc(32, sizeof(float) * 72u * (size_t)(batches ? batches : 1u));
  memset(a->soa, 0, sizeof(float) * 72u * (size_t)(batches ? batches : 1u));
  for (uint32_t i = 0; i < a->node_count; ++i) {
    const node_t* nd = &a->nodes[i];
    for (int32_t k = 0; k < nd->count; ++k) {
      const prim_t* p = &a->prims[a->order[nd->first + k]];
      float* b = a->soa + 72u * ((size_t)a->soa_base[i] + (uint32_t)k / 8u);
      const uint32_t lane = (uint32_t)k & 7u;
      const float f[9] = {p->v0.x, p->v0.y, p->v0.z, p->e1.x, p->e1.y, p->e1.z, p->e2.x, p->e2.y, p->e2.z};
      for (int c = 0; c < 9; ++c) b[8 * c + lane] = f[c];
    }
  }
}
#endif

pto_accel* pto_accel_build(const PtoScene* scene, int builder, uint32_t leaf_size, uint32_t bucket_count) {
  pto_accel* a = (pto_accel*)calloc(1, sizeof(pto_accel));
  if (!a) return NULL;
  v3* raw = NULL;
  if (flatten(scene, &a->prims, &a->prim_count, &raw) != 0) { free(a); return NULL; }
  const uint32_t n = a->prim_count;
  if (n == 0) { free(raw); return a; }
  const uint32_t max_nodes = 2u * n - 1u;
  a->nodes = (node_t*)calloc(max_nodes, sizeof(node_t));
  a->order = (uint32_t*)malloc(sizeof(uint32_t) * n);
  if (builder == PTO_BUILDER_MIDPOINT) {
    aabb_t* boxes = (aabb_t*)malloc(sizeof(aabb_t) * n);
    uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) { boxes[i] = padded_box(raw + 3 * i); a->order[i] = i; }
    mid_ctx c = {boxes, a->nodes, a->order, tmp, 1, leaf_size ? leaf_size : 4u, 0};
    mid_build(&c, 0, a->order, n, 0);
    a->node_count = (uint32_t)c.alloc;
    a->max_depth = c.max_depth;
    free(boxes);
    free(tmp);
  } else {
    ref_t* refs = (ref_t*)malloc(sizeof(ref_t) * n);
    ref_t* tmp = (ref_t*)malloc(sizeof(ref_t) * n);
    for (uint32_t i = 0; i < n; ++i) {
      const v3* t = raw + 3 * i;
      refs[i].orig = i;
      refs[i].mn[0] = min3(t[0].x, t[1].x, t[2].x);
      refs[i].mn[1] = min3(t[0].y, t[1].y, t[2].y);
      refs[i].mn[2] = min3(t[0].z, t[1].z, t[2].z);
      refs[i].mx[0] = max3(t[0].x, t[1].x, t[2].x);
      refs[i].mx[1] = max3(t[0].y, t[1].y, t[2].y);
      refs[i].mx[2] = max3(t[0].z, t[1].z, t[2].z);
      for (int k = 0; k < 3; ++k) refs[i].c[k] = (refs[i].mn[k] + refs[i].mx[k]) * RC(0.5);
    }
    sah_ctx c = {refs, tmp, a->nodes, 0, a->order, 0, leaf_size ? leaf_size : 4u, bucket_count ? bucket_count : 8u, 0};
    sah_build(&c, 0, n, 0);
    a->node_count = c.node_count;
    a->max_depth = c.max_depth;
    free(refs);
    free(tmp);
  }
  free(raw);
#if PTO_AVX2
  soa_fill(a);
#endif
  return a;
}

void pto_accel_free(pto_accel* a) {
  if (!a) return;
  free(a->soa);
  free(a->soa_base);
  free(a->prims);
  free(a->nodes);
  free(a->order);
  free(a->prim_uv);
  free(a->texels);
  free(a->textures);
  free(a->sdfs);
  free(a);
}

int pto_accel_set_textures(pto_accel* a, const float* texcoords, uint32_t texcoord_count, const float* texels,
                           uint64_t texel_count, const PtoTextureDesc* textures, uint32_t texture_count) {
  if (!a) return -1;
  free(a->prim_uv);
  free(a->texels);
  free(a->textures);
  a->prim_uv = NULL;
  a->texels = NULL;
  a->textures = NULL;
  a->texture_count = 0;
  a->texel_count = 0;
  for (uint32_t i = 0; i < texture_count; ++i)
    if (textures[i].width == 0 || textures[i].height == 0 ||
        (uint64_t)textures[i].offset + (uint64_t)textures[i].width * textures[i].height > texel_count)
      return -1;
  if (!texcoords || !texcoord_count) return 0;
  a->prim_uv = (float*)calloc(6u * (size_t)(a->prim_count ? a->prim_count : 1u), sizeof(float));
  a->texels = (float*)malloc(sizeof(float) * 3u * (size_t)(texel_count ? texel_count : 1u));
  a->textures = (PtoTextureDesc*)malloc(sizeof(PtoTextureDesc) * (texture_count ? texture_count : 1u));
  if (!a->prim_uv || !a->texels || !a->textures) return -1;
  for (uint32_t p = 0; p < a->prim_count; ++p)
    for (int k = 0; k < 3; ++k) {
      const uint32_t v = a->prims[p].vi[k];
      if (v < texcoord_count) {
        a->prim_uv[6u * p + 2u * k] = texcoords[2u * v];
        a->prim_uv[6u * p + 2u * k + 1u] = texcoords[2u * v + 1u];
      }
    }
  if (texel_count) memcpy(a->texels, texels, sizeof(float) * 3u * (size_t)texel_count);
  if (texture_count) memcpy(a->textures, textures, sizeof(PtoTextureDesc) * texture_count);
  a->texture_count = texture_count;
  a->texel_count = texel_count;
  return 0;
}

/* ------------------------------------------------------------------ dynamic-instance refit
 * CpuBvhAccelerator::apply_instance_transform_update (CpuBvhAccelerator.cpp:310-358): the primitives of
 * the updated instances are recomputed from their local geometry through the new transforms
 * (refit_instance_primitives, :527-584), then every node box is refit bottom-up as the union of its
 * children's, leaves from their padded primitive boxes (refit_node, :586-603). The topology stays.
 * Here the scene passed in already holds the updated instance records; re-flattening it recomputes
 * exactly the updated instances' primitives (the others come out as they were). */
static aabb_t union_box(aabb_t l, aabb_t r) {
  aabb_t o;
  for (int k = 0; k < 3; ++k) {
    o.mn[k] = smin(l.mn[k], r.mn[k]);
    o.mx[k] = smax(l.mx[k], r.mx[k]);
  }
  return o;
}

static aabb_t refit_node(pto_accel* a, const aabb_t* boxes, int32_t ni) {
  node_t* nd = &a->nodes[ni];
  if (nd->count > 0) {
    aabb_t b = boxes[a->order[nd->first]];
    for (int32_t i = 1; i < nd->count; ++i) b = union_box(b, boxes[a->order[nd->first + i]]);
    nd->box = b;
    return b;
  }
  const aabb_t l = refit_node(a, boxes, nd->left);
  const aabb_t r = refit_node(a, boxes, nd->right);
  nd->box = union_box(l, r);
  return nd->box;
}

int pto_accel_refit(pto_accel* a, const PtoScene* scene) {
  if (!a) return -1;
  prim_t* prims = NULL;
  uint32_t n = 0;
  v3* raw = NULL;
  if (flatten(scene, &prims, &n, &raw) != 0) return -1;
  if (n != a->prim_count) {
    free(prims);
    free(raw);
    return -2;
  }
  free(a->prims);
  a->prims = prims;
  if (a->node_count) {
    aabb_t* boxes = (aabb_t*)malloc(sizeof(aabb_t) * n);
    for (uint32_t i = 0; i < n; ++i) boxes[i] = padded_box(raw + 3 * i);
    refit_node(a, boxes, 0);
    free(boxes);
  }
  free(raw);
#if PTO_AVX2
  soa_fill(a);
#endif
  return 0;
}

uint32_t pto_accel_node_count(const pto_accel* a) { return a ? a->node_count : 0; }
uint32_t pto_accel_prim_count(const pto_accel* a) { return a ? a->prim_count : 0; }
uint32_t pto_accel_max_depth(const pto_accel* a) { return a ? a->max_depth : 0; }

int pto_accel_export(const pto_accel* a, void* bounds, int32_t* meta, uint32_t* prim_order,
                     void* prim_geo, uint32_t* prim_material) {
  if (!a) return -1;
  real* bd = (real*)bounds;
  real* pg = (real*)prim_geo;
  for (uint32_t i = 0; i < a->node_count; ++i) {
    const node_t* nd = &a->nodes[i];
    if (bd) for (int k = 0; k < 3; ++k) { bd[6 * i + k] = nd->box.mn[k]; bd[6 * i + 3 + k] = nd->box.mx[k]; }
    if (meta) { meta[4 * i] = nd->left; meta[4 * i + 1] = nd->right; meta[4 * i + 2] = nd->first; meta[4 * i + 3] = nd->count; }
  }
  if (prim_order) memcpy(prim_order, a->order, sizeof(uint32_t) * a->prim_count);
  for (uint32_t i = 0; i < a->prim_count; ++i) {
    const prim_t* p = &a->prims[i];
    if (pg) {
      real* d = pg + 12 * i;
      d[0] = p->v0.x; d[1] = p->v0.y; d[2] = p->v0.z;
      d[3] = p->e1.x; d[4] = p->e1.y; d[5] = p->e1.z;
      d[6] = p->e2.x; d[7] = p->e2.y; d[8] = p->e2.z;
      d[9] = p->normal.x; d[10] = p->normal.y; d[11] = p->normal.z;
    }
    if (prim_material) prim_material[i] = p->material_index;
  }
  return 0;
}

/* ---- traversal: CpuBvhAccelerator::intersect, CpuBvhAccelerator.cpp:360-458 */
typedef struct { v3 o; real inv[3]; int par[3]; } rayq_t;

/* make_ray_aabb_query: CpuBvhAccelerator.cpp:113-123 */
static rayq_t make_query(v3 o, v3 d) {
  rayq_t q;
  q.o = o;
  const real dd[3] = {d.x, d.y, d.z};
  for (int k = 0; k < 3; ++k) {
    q.par[k] = R_FABS(dd[k]) <= kEps;
    q.inv[k] = q.par[k] ? RC(0.0) : RC(1.0) / dd[k];
  }
  return q;
}

/* intersect_aabb: CpuBvhAccelerator.cpp:125-154 */
static int slab(const aabb_t* b, const rayq_t* q, real max_t, real* t_near) {
  real tmin = kEps, tmax = max_t;
  const real o[3] = {q->o.x, q->o.y, q->o.z};
  for (int k = 0; k < 3; ++k) {
    if (q->par[k]) {
      if (o[k] < b->mn[k] || o[k] > b->mx[k]) return 0;
      continue;
    }
    real t0 = (b->mn[k] - o[k]) * q->inv[k];
    real t1 = (b->mx[k] - o[k]) * q->inv[k];
    if (t0 > t1) { const real s = t0; t0 = t1; t1 = s; }
    tmin = smax(tmin, t0);
    tmax = smin(tmax, t1);
    if (tmin > tmax) return 0;
  }
  *t_near = tmin;
  return tmax > kEps;
}

typedef struct { int hit; real t; uint32_t prim; real u, v; } qhit_t;

/* Triangle tests of this thread rejected by |det| < 1e-4 (BvhTriangleIntersector.cpp:48), for the
 * benchmark's report of how often the scale-dependent cut-off fires (pto_det_rejections). */
static _Thread_local uint64_t tls_det_rejects;

/* Leaf batch: IntersectBvhTriangleBatch / intersect_scalar / commit_ordered_lane_hit,
 * BvhTriangleIntersector.cpp:25-74, :186-209. Returns accepted count; updates *best. */
static uint32_t leaf_batch(const pto_accel* a, const uint32_t* prims, uint32_t cnt, v3 o, v3 d,
                           real* best_t, int* lane_out, real* u_out, real* v_out) {
  uint32_t accepted = 0;
  *lane_out = -1;
  if (cnt == 0 || !(*best_t > kEps)) return 0;
  real out_t = *best_t;
  for (uint32_t lane = 0; lane < cnt; ++lane) {
    const prim_t* p = &a->prims[prims[lane]];
    const real hx = d.y * p->e2.z - d.z * p->e2.y;
    const real hy = d.z * p->e2.x - d.x * p->e2.z;
    const real hz = d.x * p->e2.y - d.y * p->e2.x;
    const real det = p->e1.x * hx + p->e1.y * hy + p->e1.z * hz;
    if (R_FABS(det) < kEps) {
      ++tls_det_rejects;
      continue;
    }
    const real inv_det = RC(1.0) / det;
    const real sx = o.x - p->v0.x, sy = o.y - p->v0.y, sz = o.z - p->v0.z;
    const real u = (sx * hx + sy * hy + sz * hz) * inv_det;
    if (u < RC(0.0) || u > RC(1.0)) continue;
    const real qx = sy * p->e1.z - sz * p->e1.y;
    const real qy = sz * p->e1.x - sx * p->e1.z;
    const real qz = sx * p->e1.y - sy * p->e1.x;
    const real v = (d.x * qx + d.y * qy + d.z * qz) * inv_det;
    if (v < RC(0.0) || u + v > RC(1.0)) continue;
    const real t = (p->e2.x * qx + p->e2.y * qy + p->e2.z * qz) * inv_det;
    if (t <= kEps || t >= out_t) continue;
    ++accepted;
    *lane_out = (int)lane;
    *u_out = u;
    *v_out = v;
    out_t = t;
  }
  *best_t = out_t;
  return accepted;
}

#if PTO_AVX2
/* leaf_batch for 8 triangles at once: the same Moller-Trumbore expressions lane-wise (mul / add /
 * sub only, no fused ops), the same acceptance tests, then the lanes committed in order so the
 * closest t, its lane (first lane on ties), the accepted count and u, v equal the scalar loop's. */
static uint32_t leaf_batch_avx2(const float* b, uint32_t cnt, v3 o, v3 d, real* best_t, int* lane_out, real* u_out,
                                real* v_out) {
  *lane_out = -1;
  if (cnt == 0 || !(*best_t > kEps)) return 0;
  const __m256 v0x = _mm256_load_ps(b), v0y = _mm256_load_ps(b + 8), v0z = _mm256_load_ps(b + 16);
  const __m256 e1x = _mm256_load_ps(b + 24), e1y = _mm256_load_ps(b + 32), e1z = _mm256_load_ps(b + 40);
  const __m256 e2x = _mm256_load_ps(b + 48), e2y = _mm256_load_ps(b + 56), e2z = _mm256_load_ps(b + 64);
  const __m256 dx = _mm256_set1_ps(d.x), dy = _mm256_set1_ps(d.y), dz = _mm256_set1_ps(d.z);
  const __m256 hx = _mm256_sub_ps(_mm256_mul_ps(dy, e2z), _mm256_mul_ps(dz, e2y));
  const __m256 hy = _mm256_sub_ps(_mm256_mul_ps(dz, e2x), _mm256_mul_ps(dx, e2z));
  const __m256 hz = _mm256_sub_ps(_mm256_mul_ps(dx, e2y), _mm256_mul_ps(dy, e2x));
  const __m256 det =
      _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(e1x, hx), _mm256_mul_ps(e1y, hy)), _mm256_mul_ps(e1z, hz));
  const __m256 inv = _mm256_div_ps(_mm256_set1_ps(1.0f), det);
  const __m256 sx = _mm256_sub_ps(_mm256_set1_ps(o.x), v0x);
  const __m256 sy = _mm256_sub_ps(_mm256_set1_ps(o.y), v0y);
  const __m256 sz = _mm256_sub_ps(_mm256_set1_ps(o.z), v0z);
  const __m256 u = _mm256_mul_ps(
      _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(sx, hx), _mm256_mul_ps(sy, hy)), _mm256_mul_ps(sz, hz)), inv);
  const __m256 qx = _mm256_sub_ps(_mm256_mul_ps(sy, e1z), _mm256_mul_ps(sz, e1y));
  const __m256 qy = _mm256_sub_ps(_mm256_mul_ps(sz, e1x), _mm256_mul_ps(sx, e1z));
  const __m256 qz = _mm256_sub_ps(_mm256_mul_ps(sx, e1y), _mm256_mul_ps(sy, e1x));
  const __m256 v = _mm256_mul_ps(
      _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(dx, qx), _mm256_mul_ps(dy, qy)), _mm256_mul_ps(dz, qz)), inv);
  const __m256 t = _mm256_mul_ps(
      _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(e2x, qx), _mm256_mul_ps(e2y, qy)), _mm256_mul_ps(e2z, qz)), inv);
  const __m256 zero = _mm256_setzero_ps(), one = _mm256_set1_ps(1.0f), eps = _mm256_set1_ps(kEps);
  const __m256 adet = _mm256_andnot_ps(_mm256_set1_ps(-0.0f), det);
  /* the scalar loop's rejections, negated (unordered compares keep NaN lanes as the scalar loop does) */
  __m256 ok = _mm256_cmp_ps(adet, eps, _CMP_NLT_UQ);
  ok = _mm256_and_ps(ok, _mm256_cmp_ps(u, zero, _CMP_NLT_UQ));
  ok = _mm256_and_ps(ok, _mm256_cmp_ps(u, one, _CMP_NGT_UQ));
  ok = _mm256_and_ps(ok, _mm256_cmp_ps(v, zero, _CMP_NLT_UQ));
  ok = _mm256_and_ps(ok, _mm256_cmp_ps(_mm256_add_ps(u, v), one, _CMP_NGT_UQ));
  ok = _mm256_and_ps(ok, _mm256_cmp_ps(t, eps, _CMP_NLE_UQ));
  uint32_t mask = (uint32_t)_mm256_movemask_ps(ok) & ((1u << cnt) - 1u);
  if (!mask) return 0;
  float tl[8], ul[8], vl[8];
  _mm256_storeu_ps(tl, t);
  _mm256_storeu_ps(ul, u);
  _mm256_storeu_ps(vl, v);
  uint32_t accepted = 0;
  real out_t = *best_t;
  for (uint32_t lane = 0; lane < cnt; ++lane) {
    if (!(mask >> lane & 1u) || tl[lane] >= out_t) continue;
    ++accepted;
    out_t = tl[lane];
    *lane_out = (int)lane;
    *u_out = ul[lane];
    *v_out = vl[lane];
  }
  *best_t = out_t;
  return accepted;
}
#endif

/* Closest hit (any_hit == 0) or any hit with t < max_t (any_hit == 1). The any-hit form
 * is the GPU shadow query's contract: it stops at the first accepted triangle. */
static qhit_t traverse(const pto_accel* a, v3 o, v3 d, real max_t, int any_hit, PtoCounters* c) {
  qhit_t h = {0, RC(0.0), 0, RC(0.0), RC(0.0)};
  if (a->node_count == 0 || a->prim_count == 0) return h;
  real best_t = max_t;
  const rayq_t q = make_query(o, d);
  int32_t stack_small[128];
  int32_t* stack = stack_small;
  uint32_t cap = 128, sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int32_t ni = stack[--sp];
    const node_t* node = &a->nodes[ni];
    c->bvh_node_visits++;
    real tn;
    if (!slab(&node->box, &q, best_t, &tn)) continue;
    if (node->count > 0) {
      c->bvh_leaf_visits++;
      for (int32_t i = 0; i < node->count;) {
        uint32_t lanes[8];
        uint32_t cnt = 0;
        while (i < node->count && cnt < 8u) lanes[cnt++] = a->order[node->first + i++];
        c->triangle_tests += cnt;
        int lane;
        real bu = RC(0.0), bv = RC(0.0);
#if PTO_AVX2
        const uint32_t batch = (uint32_t)(i - (int32_t)cnt) / 8u;
        const uint32_t acc = leaf_batch_avx2(a->soa + 72u * ((size_t)a->soa_base[ni] + batch), cnt, o, d, &best_t,
                                             &lane, &bu, &bv);
#else
        const uint32_t acc = leaf_batch(a, lanes, cnt, o, d, &best_t, &lane, &bu, &bv);
#endif
        if (lane < 0) continue;
        c->triangle_hits += acc;
        h.hit = 1;
        h.t = best_t;
        h.prim = lanes[lane];
        h.u = bu;
        h.v = bv;
        if (any_hit) goto done;
      }
      continue;
    }
    real lt = 0, rt = 0;
    const int hl = node->left >= 0 && slab(&a->nodes[node->left].box, &q, best_t, &lt);
    const int hr = node->right >= 0 && slab(&a->nodes[node->right].box, &q, best_t, &rt);
    if (sp + 2 > cap) {
      const uint32_t ncap = cap * 2;
      int32_t* ns = (int32_t*)malloc(sizeof(int32_t) * ncap);
      memcpy(ns, stack, sizeof(int32_t) * sp);
      if (stack != stack_small) free(stack);
      stack = ns;
      cap = ncap;
    }
    if (hl && hr) {
      if (lt <= rt) { stack[sp++] = node->right; stack[sp++] = node->left; }
      else { stack[sp++] = node->left; stack[sp++] = node->right; }
    } else if (hl) {
      stack[sp++] = node->left;
    } else if (hr) {
      stack[sp++] = node->right;
    }
  }
done:
  if (stack != stack_small) free(stack);
  return h;
}

int pto_trace_rays(const pto_accel* a, const float* rays, uint32_t n, int any_hit, const float* tmax,
                   float* out_t, int32_t* out_prim, PtoCounters* counters) {
  PtoCounters c;
  memset(&c, 0, sizeof(c));
  for (uint32_t i = 0; i < n; ++i) {
    const v3 o = v3f(rays + 6 * i), d = v3f(rays + 6 * i + 3);
    const real mt = tmax ? (real)tmax[i] : (real)INFINITY;
    const qhit_t h = traverse(a, o, d, mt, any_hit, &c);
    out_t[i] = h.hit ? (float)h.t : (float)INFINITY;
    out_prim[i] = h.hit ? (int32_t)a->prims[h.prim].primitive_index : -1;
  }
  if (counters) *counters = c;
  return 0;
}

/* pto_trace_rays (closest or any hit) that also reports how many of the triangle tests the |det| <
 * 1e-4 cut-off rejected: out[0] = triangle tests, out[1] = rejections. Single-threaded. */
int pto_det_rejections(const pto_accel* a, const float* rays, uint32_t n, int any_hit, uint64_t out[2]) {
  PtoCounters c;
  memset(&c, 0, sizeof(c));
  tls_det_rejects = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const v3 o = v3f(rays + 6 * i), d = v3f(rays + 6 * i + 3);
    (void)traverse(a, o, d, (real)INFINITY, any_hit, &c);
  }
  out[0] = c.triangle_tests;
  out[1] = tls_det_rejects;
  return 0;
}

/* Brute force: every primitive, strict < keeps the first of equal t (ScalarCpuPathTracer.Intersect.cpp:52-89). */
int pto_brute_force_rays(const pto_accel* a, const float* rays, uint32_t n, float* out_t, int32_t* out_prim) {
  for (uint32_t i = 0; i < n; ++i) {
    const v3 o = v3f(rays + 6 * i), d = v3f(rays + 6 * i + 3);
    real best = (real)INFINITY;
    int32_t bp = -1;
    for (uint32_t p = 0; p < a->prim_count; ++p) {
      real bt = best, bu, bv;
      int lane;
      leaf_batch(a, &p, 1, o, d, &bt, &lane, &bu, &bv);
      if (lane == 0) { best = bt; bp = (int32_t)p; }
    }
    out_t[i] = bp >= 0 ? (float)best : (float)INFINITY;
    out_prim[i] = bp;
  }
  return 0;
}

/* ------------------------------------------------------------------ integrator */
/* A bilinear tap of a texture: its four texels (indices into the texel array), weights and the
 * filtered colour. tex = ~0u: no texture. */
typedef struct {
  uint32_t tex;
  uint32_t idx[4];
  real w[4];
  v3 color;
} tap_t;

typedef struct {
  int hit;
  real t;
  v3 position, normal;
  uint32_t material_index;
  uint32_t prim;   /* flattened primitive */
  real bu, bv;     /* barycentrics of the hit */
  v3 base;         /* base colour: material albedo times its texture (set by the shading code) */
  tap_t tap;
} hit_t;

typedef struct {
  const PtoScene* sc;
  const pto_accel* a;
  const PtoSettings* s;
  uint32_t max_depth;
  int shadow_mode;
} ctx_t;

/* ---- SDF primitives: sdf_distance / sdf_normal / intersect_sphere / intersect_box and its aliases
 * (ScalarCpuPathTracer.cpp:809-962), rotate_x/y/z and rotate_euler(_inv) (:113-137), divide_by_scale
 * (:154-159). */
#define MIN_MARCH_STEP RC(1.0e-3) /* ScalarCpuPathTracer.cpp:17-19 */
#define MAX_MARCH_STEPS 192u
#define MAX_MARCH_DISTANCE RC(10000.0)

static v3 rot_x(v3 v, real a) { const real s = R_SIN(a), c = R_COS(a); return mk(v.x, c * v.y - s * v.z, s * v.y + c * v.z); }
static v3 rot_y(v3 v, real a) { const real s = R_SIN(a), c = R_COS(a); return mk(c * v.x + s * v.z, v.y, -s * v.x + c * v.z); }
static v3 rot_z(v3 v, real a) { const real s = R_SIN(a), c = R_COS(a); return mk(c * v.x - s * v.y, s * v.x + c * v.y, v.z); }
static v3 euler_inv(v3 v, const PtoSdf* p) {
  return rot_z(rot_y(rot_x(v, -(real)p->rotation[0]), -(real)p->rotation[1]), -(real)p->rotation[2]);
}
static v3 euler_fwd(v3 v, const PtoSdf* p) {
  return rot_x(rot_y(rot_z(v, (real)p->rotation[2]), (real)p->rotation[1]), (real)p->rotation[0]);
}
static v3 div_scale(v3 v, const float* s) {
  return mk(v.x / (s[0] != 0.0f ? (real)s[0] : RC(1.0)), v.y / (s[1] != 0.0f ? (real)s[1] : RC(1.0)),
            v.z / (s[2] != 0.0f ? (real)s[2] : RC(1.0)));
}

static real sdf_distance(const PtoSdf* p, v3 w) {
  const v3 l = div_scale(euler_inv(sub(w, v3f(p->position)), p), p->scale);
  const real sx = (real)p->scale[0], sy = (real)p->scale[1], sz = (real)p->scale[2];
  switch (p->shape) {
    case 1u: {
      const v3 q = mk(R_FABS(l.x) - sx, R_FABS(l.y) - sy, R_FABS(l.z) - sz);
      const v3 mq = mk(smax(RC(0.0), q.x), smax(RC(0.0), q.y), smax(RC(0.0), q.z));
      return length(mq) + smin(smax(q.x, smax(q.y, q.z)), RC(0.0));
    }
    case 2u: {
      const real r = (real)p->param_a;
      const v3 q = mk(R_FABS(l.x) - sx + r, R_FABS(l.y) - sy + r, R_FABS(l.z) - sz + r);
      const v3 mq = mk(smax(RC(0.0), q.x), smax(RC(0.0), q.y), smax(RC(0.0), q.z));
      return length(mq) + smin(smax(q.x, smax(q.y, q.z)), RC(0.0)) - (real)p->param_a;
    }
    case 3u:
      return l.y + (real)p->param_b;
    case 4u: {
      const real a = smax(RC(0.05), (real)p->param_a);
      const real b = smax(RC(0.02), (real)p->radius);
      const real qx = length(mk(l.x, RC(0.0), l.z)) - a;
      return length(mk(qx, l.y, RC(0.0))) - b;
    }
    case 5u: {
      const v3 a = mk(RC(0.0), -(real)p->param_a, RC(0.0));
      const v3 b = mk(RC(0.0), (real)p->param_a, RC(0.0));
      const v3 ab = sub(b, a);
      const real t = dot(sub(l, a), ab) / smax(kEps, dot(ab, ab));
      const real c = smax(RC(0.0), smin(RC(1.0), t));
      return length(sub(l, add(a, muls(ab, c)))) - (real)p->radius;
    }
    default:
      return length(l) - (real)p->radius;
  }
}

static v3 sdf_normal(const PtoSdf* p, v3 w) {
  const real e = RC(0.0005);
  const real dx = sdf_distance(p, add(w, mk(e, RC(0.0), RC(0.0)))) - sdf_distance(p, sub(w, mk(e, RC(0.0), RC(0.0))));
  const real dy = sdf_distance(p, add(w, mk(RC(0.0), e, RC(0.0)))) - sdf_distance(p, sub(w, mk(RC(0.0), e, RC(0.0))));
  const real dz = sdf_distance(p, add(w, mk(RC(0.0), RC(0.0), e))) - sdf_distance(p, sub(w, mk(RC(0.0), RC(0.0), e)));
  return normalize(muls(mk(dx, dy, dz), RC(0.5)));
}

static int sdf_test(const PtoSdf* p, v3 o, v3 d, real* t_out, v3* n_out, PtoCounters* c) {
  ++c->sdf_tests;
  if (p->shape == 0u) {
    const real radius = smax(RC(0.01), (real)p->radius);
    const v3 lo = div_scale(euler_inv(sub(o, v3f(p->position)), p), p->scale);
    const v3 ld = div_scale(euler_inv(d, p), p->scale);
    const real a = dot(ld, ld);
    ++c->sdf_steps;
    if (a <= kEps * kEps) { ++c->sdf_misses; return 0; }
    const real hb = dot(lo, ld);
    const real cc = dot(lo, lo) - radius * radius;
    const real disc = hb * hb - a * cc;
    if (disc < RC(0.0)) { ++c->sdf_misses; return 0; }
    const real root = R_SQRT(disc);
    real cand = (-hb - root) / a;
    if (cand <= kEps) cand = (-hb + root) / a;
    if (cand <= kEps || !isfinite(cand)) { ++c->sdf_misses; return 0; }
    *t_out = cand;
    const v3 lhit = add(lo, muls(ld, cand));
    *n_out = normalize(euler_fwd(div_scale(normalize(lhit), p->scale), p));
    ++c->sdf_hits;
    return 1;
  }
  real t = RC(0.0);
  for (uint32_t i = 0; i < MAX_MARCH_STEPS; ++i) {
    ++c->sdf_steps;
    const v3 pt = add(o, muls(d, t));
    const real dd = sdf_distance(p, pt);
    if (dd <= kEps) {
      *n_out = sdf_normal(p, pt);
      *t_out = t;
      ++c->sdf_hits;
      return 1;
    }
    t += smax(MIN_MARCH_STEP, dd);
    if (t > MAX_MARCH_DISTANCE) { ++c->sdf_misses; return 0; }
  }
  ++c->sdf_misses;
  return 0;
}

/* intersect_scene: ScalarCpuPathTracer.Intersect.cpp:28-132 (the accelerator, then every SDF primitive;
 * an SDF hit replaces the surface when it is nearer). SDF hits carry prim = ~0u. */
static hit_t intersect_scene(const ctx_t* x, v3 o, v3 d, PtoCounters* c) {
  hit_t h;
  memset(&h, 0, sizeof(h));
  real best = (real)INFINITY;
  const qhit_t q = traverse(x->a, o, d, (real)INFINITY, 0, c);
  if (q.hit) {
    const prim_t* p = &x->a->prims[q.prim];
    h.hit = 1;
    h.t = q.t;
    h.position = add(o, muls(d, q.t));
    h.normal = p->normal;
    h.material_index = p->material_index;
    h.prim = q.prim;
    h.bu = q.u;
    h.bv = q.v;
    best = q.t;
  }
  for (uint32_t k = 0; k < x->a->sdf_count; ++k) {
    const PtoSdf* p = &x->a->sdfs[k];
    real t = RC(0.0);
    v3 n = mk(RC(0.0), RC(0.0), RC(0.0));
    if (!sdf_test(p, o, d, &t, &n, c) || t >= best) continue;
    best = t;
    h.hit = 1;
    h.t = t;
    h.position = add(o, muls(d, t));
    h.normal = normalize(n);
    h.material_index = p->material_index;
    h.prim = ~0u;
    h.bu = h.bv = RC(0.0);
  }
  return h;
}

/* intersect_scene for a batch of rays (o xyz, d xyz): t (inf on a miss), the reported normal, the
 * material index and whether the surface is an SDF primitive. */
int pto_intersect_scene_rays(const PtoScene* sc, const pto_accel* a, const float* rays, uint32_t n, float* out_t,
                             float* out_normal, uint32_t* out_material, uint8_t* out_sdf, PtoCounters* c) {
  if (!sc || !a || (n && !rays)) return -1;
  PtoSettings st;
  memset(&st, 0, sizeof(st));
  ctx_t x = {sc, a, &st, 1u, PTO_SHADOW_CLOSEST};
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = rays + 6u * i;
    const hit_t h = intersect_scene(&x, v3f(r), v3f(r + 3), c);
    out_t[i] = h.hit ? (float)h.t : INFINITY;
    out_normal[3 * i] = (float)h.normal.x;
    out_normal[3 * i + 1] = (float)h.normal.y;
    out_normal[3 * i + 2] = (float)h.normal.z;
    out_material[i] = h.material_index;
    out_sdf[i] = (uint8_t)(h.hit && h.prim == ~0u);
  }
  return 0;
}

int pto_accel_set_sdfs(pto_accel* a, const PtoSdf* sdfs, uint32_t count) {
  if (!a || (count && !sdfs)) return -1;
  free(a->sdfs);
  a->sdfs = NULL;
  a->sdf_count = 0;
  if (!count) return 0;
  a->sdfs = (PtoSdf*)malloc(sizeof(PtoSdf) * count);
  if (!a->sdfs) return -1;
  for (uint32_t i = 0; i < count; ++i)
    if (sdfs[i].shape <= 5u) a->sdfs[a->sdf_count++] = sdfs[i]; /* Unknown shapes are never tested */
  return 0;
}

static const PtoMaterial* material_at(const PtoScene* sc, uint32_t index) {
  static const PtoMaterial kDefault = {{1.0f, 1.0f, 1.0f}, {0.0f, 0.0f, 0.0f}, 1.0f, 0.0f, 1.5f, 0.0f, 0.0f, 0.0f,
                                       0.0f, 1.0f, 0u, 0u, 0u, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u};
  if (sc->material_count == 0) return &kDefault; /* load_scene_snapshot adds one default material (:469-471) */
  const uint32_t last = sc->material_count - 1u;
  return &sc->materials[index < last ? index : last];
}

#if PTO_REAL_IS_DOUBLE
#define R_FLOOR(x) floor(x)
#define R_TRUNC(x) trunc(x)
#define R_ATAN2(y, x) atan2((y), (x))
#define R_ACOS(x) acos(x)
#else
#define R_FLOOR(x) floorf(x)
#define R_TRUNC(x) truncf(x)
#if PTO_LIBM
#define R_ATAN2(y, x) atan2f((y), (x))
#define R_ACOS(x) acosf(x)
#else
#define R_ATAN2(y, x) pm_atan2f((y), (x))
#define R_ACOS(x) pm_acosf(x)
#endif
#endif

/* ---- textures (no counterpart in the reference's CPU path, whose Hit carries no texture
 * coordinates, PathTracer.h:1283-1289). uv interpolation at the hit barycentrics, then a bilinear
 * tap with repeat addressing and texel centres at half-integers. The operation order is fixed so
 * that fp32 builds reproduce the device bit for bit. */
static tap_t texture_tap(const pto_accel* a, uint32_t ti, uint32_t prim, real bu, real bv) {
  tap_t r;
  const float* uvp = a->prim_uv + 6u * prim;
  const real w = RC(1.0) - bu - bv;
  const real uu = (real)uvp[0] * w + (real)uvp[2] * bu + (real)uvp[4] * bv;
  const real vv = (real)uvp[1] * w + (real)uvp[3] * bu + (real)uvp[5] * bv;
  const PtoTextureDesc* T = &a->textures[ti];
  real uf = uu - R_FLOOR(uu), vf = vv - R_FLOOR(vv);
  if (!(uf >= RC(0.0) && uf < RC(1.0))) uf = RC(0.0);
  if (!(vf >= RC(0.0) && vf < RC(1.0))) vf = RC(0.0);
  const real x = uf * (real)T->width - RC(0.5), y = vf * (real)T->height - RC(0.5);
  const real x0 = R_FLOOR(x), y0 = R_FLOOR(y);
  const real fx = x - x0, fy = y - y0;
  int ix0 = (int)x0, iy0 = (int)y0;
  if (ix0 < 0) ix0 += (int)T->width;
  if (iy0 < 0) iy0 += (int)T->height;
  const int ix1 = ix0 + 1 == (int)T->width ? 0 : ix0 + 1;
  const int iy1 = iy0 + 1 == (int)T->height ? 0 : iy0 + 1;
  r.tex = ti;
  r.idx[0] = T->offset + (uint32_t)iy0 * T->width + (uint32_t)ix0;
  r.idx[1] = T->offset + (uint32_t)iy0 * T->width + (uint32_t)ix1;
  r.idx[2] = T->offset + (uint32_t)iy1 * T->width + (uint32_t)ix0;
  r.idx[3] = T->offset + (uint32_t)iy1 * T->width + (uint32_t)ix1;
  r.w[0] = (RC(1.0) - fx) * (RC(1.0) - fy);
  r.w[1] = fx * (RC(1.0) - fy);
  r.w[2] = (RC(1.0) - fx) * fy;
  r.w[3] = fx * fy;
  v3 col = mk(RC(0.0), RC(0.0), RC(0.0));
  for (int k = 0; k < 4; ++k) col = add(col, muls(v3f(a->texels + 3u * r.idx[k]), r.w[k]));
  r.color = col;
  return r;
}

int pto_texture_sample(const pto_accel* a, uint32_t tex, uint32_t prim, float bu, float bv, double* out_rgb,
                       uint32_t* out_idx, double* out_w) {
  if (!a || !a->prim_uv || tex >= a->texture_count || prim >= a->prim_count) return -1;
  const tap_t t = texture_tap(a, tex, prim, (real)bu, (real)bv);
  out_rgb[0] = (double)t.color.x;
  out_rgb[1] = (double)t.color.y;
  out_rgb[2] = (double)t.color.z;
  for (int k = 0; k < 4; ++k) {
    out_idx[k] = t.idx[k];
    out_w[k] = (double)t.w[k];
  }
  return 0;
}

/* The material's base colour at a hit: albedo, times the texture when the material names one. */
static void hit_base_color(const pto_accel* a, const PtoMaterial* m, hit_t* h) {
  h->base = v3f(m->albedo);
  h->tap.tex = ~0u;
  if (a->prim_uv && m->base_color_texture_index < a->texture_count && h->prim != ~0u) {
    h->tap = texture_tap(a, m->base_color_texture_index, h->prim, h->bu, h->bv);
    h->base = mulv(h->base, h->tap.color);
  }
}

/* std::fmod(x, 2) for the integer-valued sums the effects pass (exact). */
static inline real fmod2(real x) { return x - RC(2.0) * R_TRUNC(x * RC(0.5)); }

/* hash01 / smoothstep: ScalarCpuPathTracer.cpp:187-196 */
static inline real hash01(real x, real y, real z, real seed) {
  const real v = R_SIN(x * RC(12.9898) + y * RC(78.233) + z * RC(37.719) + seed * RC(19.19)) * RC(43758.5453);
  return v - R_FLOOR(v);
}
static inline real smoothstep_r(real e0, real e1, real x) {
  const real t = clamp01((x - e0) / smax(RC(1.0e-6), e1 - e0));
  return t * t * (RC(3.0) - RC(2.0) * t);
}

/* procedural_wood_albedo: ScalarCpuPathTracer.cpp:198-262. Linear in `base`: *k receives the factor. */
static v3 wood_albedo(v3 base, v3 p, real roughness, real clearcoat, uint32_t style, real* k) {
  const real rough = clamp01(roughness), coat = clamp01(clearcoat);
  const real polish = clamp01(coat * (RC(1.25) - rough * RC(0.75)));
  const real parquet = style == 17u ? RC(1.0) : RC(0.0);
  const real walnut = style == 18u ? RC(1.0) : RC(0.0);
  const real sandal = style == 19u ? RC(1.0) : RC(0.0);
  const real plank_width = RC(0.40) + RC(0.16) * parquet + RC(0.08) * polish - RC(0.18) * walnut;
  const real plank_length = RC(1.15) + RC(0.85) * parquet + RC(0.55) * polish - RC(0.55) * walnut;
  const real px = (p.x + sandal * p.y * RC(0.35)) / smax(RC(0.12), plank_width);
  const real pz = (p.z + walnut * p.y * RC(0.70) + sandal * p.x * RC(0.22)) / smax(RC(0.30), plank_length);
  const real cell_x = R_FLOOR(px), cell_z = R_FLOOR(pz);
  const real local_x = px - cell_x, local_z = pz - cell_z;
  const real seed = hash01(cell_x, cell_z, RC(0.0), RC(16.0));
  const real macro_noise = hash01(R_FLOOR(p.x * RC(3.7)), R_FLOOR(p.y * RC(3.7)), R_FLOOR(p.z * RC(3.7)),
                                  RC(31.0) + (real)style);
  const real micro_noise = hash01(R_FLOOR(p.x * RC(34.0)), R_FLOOR(p.y * RC(34.0)), R_FLOOR(p.z * RC(34.0)),
                                  RC(47.0) + (real)style);
  const int rotate = parquet > RC(0.5) && fmod2(cell_x + cell_z) >= RC(1.0);
  const real along = rotate ? local_z : local_x, across = rotate ? local_x : local_z;
  const real grain_wobble = (macro_noise - RC(0.5)) * (RC(0.06) + rough * RC(0.12) + walnut * RC(0.08) + sandal * RC(0.04));
  const real warped_along = along + grain_wobble + R_SIN(across * RC(6.0) + seed * RC(5.0)) * RC(0.018);
  const real grain_freq = RC(12.0) + rough * RC(8.0) + walnut * RC(22.0) + sandal * RC(36.0);
  const real wave = R_SIN((warped_along * grain_freq + seed * RC(6.2831853)) +
                          R_SIN(across * (RC(6.0) + rough * RC(8.0)) + seed * RC(4.0)) *
                              (RC(0.25) + rough * RC(0.45) + walnut * RC(0.45)));
  const real fine = R_SIN(warped_along * (RC(55.0) + rough * RC(55.0) + sandal * RC(110.0)) +
                          across * (RC(9.0) + walnut * RC(25.0)) + seed * RC(12.0));
  const real fine_weight = RC(0.10) + rough * RC(0.18) + walnut * RC(0.18) + sandal * RC(0.22);
  const real ring = smoothstep_r(RC(-0.35) + polish * RC(0.18), RC(0.82), wave) * (RC(1.0) - fine_weight) +
                    (RC(0.5) + RC(0.5) * fine) * fine_weight;
  const real seam_x = smin(local_x, RC(1.0) - local_x), seam_z = smin(local_z, RC(1.0) - local_z);
  const real seam = (RC(1.0) - sandal) *
                    (RC(1.0) - smoothstep_r(RC(0.0), RC(0.028) + polish * RC(0.018), smin(seam_x, seam_z)));
  const real kd = RC(0.46) - rough * RC(0.10) - walnut * RC(0.18);
  const real ka = RC(1.14) + coat * RC(0.22) + sandal * RC(0.18);
  const v3 dark = muls(base, kd);
  const v3 amber = add(muls(base, ka), mk(RC(0.08) + sandal * RC(0.18), RC(0.055) + sandal * RC(0.06), RC(0.025)));
  v3 wood = add(muls(dark, RC(1.0) - ring), muls(amber, ring));
  const real board_tint = RC(0.92) + (RC(0.08) + walnut * RC(0.12) + sandal * RC(0.06)) * seed +
                          (macro_noise - RC(0.5)) * (RC(0.08) + rough * RC(0.06));
  const real pores = smoothstep_r(RC(0.74) - walnut * RC(0.08), RC(1.0), micro_noise) *
                     (RC(0.025) + rough * RC(0.055) + walnut * RC(0.055) + sandal * RC(0.018));
  const real seam_f = RC(1.0) - seam * (RC(0.10) + rough * RC(0.35) + parquet * RC(0.18)) * (RC(1.0) - coat * RC(0.45));
  wood = muls(wood, board_tint);
  wood = muls(wood, RC(1.0) - pores);
  wood = muls(wood, seam_f);
  const real wood_mix = RC(0.30) + rough * RC(0.34) - polish * RC(0.14) + walnut * RC(0.26) + sandal * RC(0.18);
  *k = (RC(1.0) - wood_mix) + (kd * (RC(1.0) - ring) + ka * ring) * board_tint * (RC(1.0) - pores) * seam_f * wood_mix;
  return add(muls(base, RC(1.0) - wood_mix), muls(wood, wood_mix));
}

/* apply_material_effect: ScalarCpuPathTracer.cpp:264-355. Every effect is affine in the albedo
 * (colour = k * albedo + b before the final [0, 1.5] clamp): *k receives k and pre[] the colour
 * before the clamp, which the adjoint needs. */
static v3 effect_albedo(const PtoMaterial* m, v3 base, v3 pos, v3 n, v3 incoming, real* k_out, v3* pre_out) {
  v3 c = base;
  real k = RC(1.0);
  const uint32_t e = m->material_effect;
  if (e != 0u) {
    const real h = hash01(pos.x, pos.y, pos.z, (real)e);
    const real ndi = dot(n, neg(incoming));
    switch (e) {
      case 1u: {
        const real r0 = smax(RC(0.0), RC(1.0) - R_FABS(ndi));
        const real rim = r0 * r0;
        k = RC(0.65) + RC(0.25) * h;
        c = add(muls(c, k), muls(mk(RC(0.25), RC(0.22), RC(0.28)), rim * (RC(0.4) + (real)m->sheen)));
        break;
      }
      case 2u: {
        const real stripes = RC(0.5) + RC(0.5) * R_SIN(pos.x * RC(7.0) + pos.z * RC(5.0) + h * RC(6.0));
        k = RC(0.45) + RC(0.55) * stripes;
        c = muls(c, k);
        break;
      }
      case 3u:
        k = h > RC(0.72) ? RC(0.18) : RC(1.0);
        c = muls(c, k);
        break;
      case 4u: {
        const real vein = RC(0.5) + RC(0.5) * R_SIN((pos.x + pos.y * RC(0.4) + pos.z * RC(0.7)) * RC(9.0) + h * RC(3.0));
        k = RC(0.55) + RC(0.45) * vein;
        c = add(muls(c, k), muls(mk(RC(0.18), RC(0.20), RC(0.23)), RC(1.0) - vein));
        break;
      }
      case 5u:
        k = RC(0.55) + RC(0.35) * h;
        c = add(muls(c, k), muls(mk(RC(0.45), RC(0.16), RC(0.04)), RC(0.25) + RC(0.25) * h));
        break;
      case 6u:
        k = RC(0.65) + RC(0.35) * h;
        c = muls(c, k);
        break;
      case 7u:
        k = RC(0.78);
        c = add(muls(c, k), muls(mk(RC(1.0), RC(0.62), RC(0.42)), RC(0.16)));
        break;
      case 8u:
        k = RC(0.55);
        c = add(muls(c, k), muls(mk(RC(0.35) + RC(0.45) * h, RC(0.25) + RC(0.35) * (RC(1.0) - h), RC(0.85)), RC(0.45)));
        break;
      case 9u: {
        const real check = fmod2(R_FLOOR(pos.x * RC(4.0)) + R_FLOOR(pos.z * RC(4.0)));
        k = check < RC(0.5) ? RC(1.0) : RC(0.28);
        c = muls(c, k);
        break;
      }
      case 10u:
        k = RC(0.45);
        c = add(muls(c, k), muls(mk(RC(1.0), RC(0.35) + RC(0.3) * h, RC(0.08)), RC(0.55)));
        break;
      case 11u: {
        const real streak = RC(0.5) + RC(0.5) * R_SIN(pos.y * RC(18.0) + h * RC(5.0));
        k = RC(0.55) + RC(0.45) * streak;
        c = muls(c, k);
        break;
      }
      case 12u: {
        const real v = smax(RC(0.0), ndi);
        const real v2 = v * v;
        c = add(c, muls(mk(RC(0.55), RC(0.65), RC(0.95)), v2 * v2 * v2));
        break;
      }
      case 13u:
        k = RC(0.65) + RC(0.35) * R_FABS(R_SIN(pos.x * RC(10.0)) * R_COS(pos.z * RC(10.0)));
        c = muls(c, k);
        break;
      case 14u: {
        const real rim = R_POW(smax(RC(0.0), RC(1.0) - R_FABS(ndi)), RC(0.8));
        k = RC(0.0);
        c = muls(mk(RC(0.15), RC(0.75), RC(1.0)), RC(0.2) + RC(0.8) * rim);
        break;
      }
      case 15u: {
        const real bands = RC(0.5) + RC(0.5) * R_SIN(pos.y * RC(11.0) + pos.x * RC(3.0) + h * RC(7.0));
        const real rim = R_POW(smax(RC(0.0), RC(1.0) - R_FABS(ndi)), RC(1.5));
        k = RC(0.25) + RC(0.45) * bands;
        c = add(muls(c, k), muls(mk(RC(0.48), RC(0.56), RC(0.68)), RC(0.18) + RC(0.32) * rim));
        break;
      }
      case 16u:
      case 17u:
      case 18u:
      case 19u:
        c = wood_albedo(c, pos, (real)m->roughness, (real)m->clearcoat, e, &k);
        break;
      default:
        break;
    }
  }
  if (k_out) *k_out = k;
  if (pre_out) *pre_out = c;
  return mk(smin(RC(1.5), smax(RC(0.0), c.x)), smin(RC(1.5), smax(RC(0.0), c.y)), smin(RC(1.5), smax(RC(0.0), c.z)));
}

static int is_emissive(const PtoMaterial* m) { return m->emissive[0] > 0.0f || m->emissive[1] > 0.0f || m->emissive[2] > 0.0f; }

/* spot_attenuation: ScalarCpuPathTracer.cpp:168-185 */
static real spot_attenuation(const PtoLight* l, v3 from_light_dir) {
  const real inner = (real)l->spot_inner_cos, outer = (real)l->spot_outer_cos;
  if (inner <= RC(-0.999)) return RC(1.0);
  const real cone = dot(normalize(from_light_dir), normalize(v3f(l->direction)));
  if (cone <= outer) return RC(0.0);
  if (cone >= inner || inner <= outer) return RC(1.0);
  const real t = sclamp((cone - outer) / (inner - outer), RC(0.0), RC(1.0));
  return t * t * (RC(3.0) - RC(2.0) * t);
}

static inline real pow5(real v) { const real v2 = v * v; return v2 * v2 * v; }

/* radiance/valid: NeeResult (PathTracer.h:181-184). The adjoint also reads the partial
 * derivatives of `radiance` by the albedo channel (d_albedo) and by the light colour (d_color)
 * and the index of the sampled light. */
typedef struct {
  v3 radiance;
  int valid;
  v3 d_albedo, d_color, d_rough;  /* d_rough: d radiance / d(material roughness parameter) */
  uint32_t light;
  real eff_k;  /* d(effect colour)/d(albedo) before the clamp, and that pre-clamp colour */
  v3 eff_pre;
} nee_t;

static int shadow_visible(const ctx_t* x, v3 o, v3 d, real dist, PtoCounters* c) {
  if (x->shadow_mode == PTO_SHADOW_ANYHIT) {
    /* Opaque-blocker any-hit query with t_max = dist - 0.004; only valid when the scene has
     * no transmissive material (the caller checks). */
    const real remaining = dist - RC(0.004);
    if (!(remaining > RC(0.004))) return 1;
    const qhit_t q = traverse(x->a, o, d, remaining, 1, c);
    return !q.hit;
  }
  /* ScalarCpuPathTracer.cpp:1066-1087 */
  v3 vo = o;
  real remaining = dist - RC(0.004);
  for (uint32_t skips = 0; skips < 8u && remaining > RC(0.004); ++skips) {
    const hit_t sh = intersect_scene(x, vo, d, c);
    if (!sh.hit || sh.t >= remaining) break;
    const PtoMaterial* m = material_at(x->sc, sh.material_index);
    const int transmissive = m->material_model == 5u || m->transmission > 0.05f;
    if (!transmissive) return 0;
    const real advance = sh.t + RC(0.004);
    vo = add(vo, muls(d, advance));
    remaining -= advance;
  }
  return 1;
}

/* sample_direct_light: ScalarCpuPathTracer.cpp:1022-1118 (hit normal is the shading normal). */
static nee_t sample_direct_light(const ctx_t* x, const hit_t* hit, v3 view_dir, uint64_t* rng, PtoCounters* c) {
  nee_t r;
  memset(&r, 0, sizeof(r));
  const PtoScene* sc = x->sc;
  if (sc->light_count == 0) return r;
  const uint32_t nl = sc->light_count;
  const uint64_t li = (uint64_t)(next01(rng) * (real)nl);
  r.light = (uint32_t)(li < nl - 1u ? li : nl - 1u);
  const PtoLight* light = &sc->lights[r.light];
  v3 lp = v3f(light->position);
  if (light->radius > 1.0e-5f) {
    const real uz = RC(1.0) - RC(2.0) * next01(rng);
    const real upr = R_SQRT(smax(RC(0.0), RC(1.0) - uz * uz));
    const real phi = RC(6.28318530718) * next01(rng);
    lp = add(lp, muls(mk(upr * R_COS(phi), uz, upr * R_SIN(phi)), (real)light->radius));
  }
  const v3 to_light = sub(lp, hit->position);
  const real dist_sq = length_sq(to_light);
  const real dist = R_SQRT(dist_sq);
  if (dist < kEps) return r;
  const v3 light_dir = divs(to_light, dist);
  const real cos_theta = dot(hit->normal, light_dir);
  if (cos_theta <= RC(0.0)) return r;
  const real spot = spot_attenuation(light, neg(light_dir));
  if (spot <= RC(0.0)) return r;
  const PtoMaterial* m = material_at(sc, hit->material_index);
  const real roughness = clamp01((real)m->roughness);
  const int isMirror = m->material_model == 2u || roughness <= RC(0.001);
  const int isMetallic = m->material_model == 4u || m->metallic > 0.65f;
  const int isTransmissive = m->material_model == 5u || m->transmission > 0.05f;
  const int isClearcoat = m->material_model == 7u || m->clearcoat > 0.05f;
  c->shadow_tests++;
  const v3 so = add(hit->position, muls(hit->normal, RC(0.002)));
  if (!shadow_visible(x, so, light_dir, dist, c)) return r;
  const v3 albedo = effect_albedo(m, hit->base, hit->position, hit->normal, neg(light_dir), &r.eff_k, &r.eff_pre);
  const real kscale = ((real)light->intensity * spot) / (dist_sq + kEps);
  const v3 irradiance = muls(v3f(light->color), kscale);
  v3 direct = mk(0, 0, 0);
  const real nlf = (real)nl;
  r.d_albedo = mk(0, 0, 0);
  r.d_color = mk(0, 0, 0);
  if (!isMirror && !isTransmissive) {
    direct = add(direct, muls(muls(mulv(muls(albedo, kInvPi), irradiance), cos_theta), nlf));
    r.d_albedo = muls(irradiance, kInvPi * cos_theta * nlf);
    r.d_color = muls(albedo, kInvPi * kscale * cos_theta * nlf);
  }
  if (isMirror || isMetallic || isClearcoat || isTransmissive || roughness < RC(0.65)) {
    const v3 half_dir = normalize(add(light_dir, normalize(view_dir)));
    const real effR = smax(RC(0.025), roughness * (isMetallic ? RC(0.65) : RC(1.0)));
    const real a2 = effR * effR;
    const real specPower = sclamp(RC(2.0) / smax(RC(0.0005), a2 * a2) - RC(2.0), RC(4.0), RC(96.0));
    const real spec = R_POW(smax(RC(0.0), dot(hit->normal, half_dir)), specPower);
    const real cosView = smax(RC(0.0), dot(hit->normal, normalize(view_dir)));
    const real ior = (real)m->ior;
    real f0 = (RC(1.0) - ior) / (RC(1.0) + ior);
    f0 *= f0;
    const real fresnel = f0 + (RC(1.0) - f0) * pow5(RC(1.0) - cosView);
    const real ss_raw = (RC(1.0) - roughness) * RC(0.8) + (real)m->metallic * RC(0.45) +
                        (real)m->clearcoat * RC(0.35) + (isMirror ? RC(0.75) : RC(0.0)) +
                        (isTransmissive ? fresnel : RC(0.0));
    const real specStrength = clamp01(ss_raw);
    const v3 white = mk(RC(1.0), RC(1.0), RC(1.0));
    const v3 specTint = add(muls(white, isMetallic ? RC(0.15) : RC(0.88)), muls(albedo, isMetallic ? RC(0.85) : RC(0.12)));
    const real s = spec * cos_theta * nlf * smax(RC(0.15), specStrength);
    direct = add(direct, muls(mulv(specTint, irradiance), s));
    r.d_albedo = add(r.d_albedo, muls(irradiance, (isMetallic ? RC(0.85) : RC(0.12)) * s));
    r.d_color = add(r.d_color, muls(specTint, kscale * s));
    /* d s / d roughness: through the Blinn-Phong exponent (2 / effR^4 - 2, unclamped in (4, 96))
     * and through specStrength (unclamped in (0, 1) and above the 0.15 floor); times the clamp01
     * derivative of the roughness parameter. */
    const double nh = (double)smax(RC(0.0), dot(hit->normal, half_dir));
    const double raw = 2.0 / fmax(0.0005, (double)a2 * (double)a2) - 2.0;
    const double kM = isMetallic ? 0.65 : 1.0;
    double dspec = 0.0;
    if (nh > 0.0 && raw > 4.0 && raw < 96.0 && (double)roughness * kM > 0.025)
      dspec = (double)spec * log(nh) * (-8.0 / pow((double)effR, 5.0)) * kM;
    double ds = dspec * (double)smax(RC(0.15), specStrength);
    if (specStrength > RC(0.15) && ss_raw > RC(0.0) && ss_raw < RC(1.0)) ds += (double)spec * -0.8;
    ds *= (double)cos_theta * (double)nlf * ((m->roughness > 0.0f && m->roughness < 1.0f) ? 1.0 : 0.0);
    r.d_rough = muls(mulv(specTint, irradiance), (real)ds);
  }
  r.radiance = direct;
  r.valid = 1;
  return r;
}

/* light_pdf_for_direction: ScalarCpuPathTracer.cpp:1120-1138 */
static real light_pdf_for_direction(const PtoScene* sc, v3 position, v3 direction) {
  if (sc->light_count == 0) return RC(0.0);
  const real inv_n = RC(1.0) / (real)sc->light_count;
  real total = RC(0.0);
  for (uint32_t i = 0; i < sc->light_count; ++i) {
    const v3 to_light = sub(v3f(sc->lights[i].position), position);
    const real dist = length(to_light);
    if (dist < kEps) continue;
    const v3 dir = divs(to_light, dist);
    if (dot(dir, direction) > RC(0.9999)) total += inv_n / (dist * dist + kEps);
  }
  return total;
}

/* MisWeight: ScalarCpuPathTracer.cpp:1140-1146 */
static real mis_weight(real a, real b) {
  const real a2 = a * a, b2 = b * b;
  const real denom = a2 + b2;
  if (denom <= RC(0.0)) return RC(1.0);
  return a2 / denom;
}

/* sample_hemisphere: ScalarCpuPathTracer.cpp:978-988 */
static v3 sample_hemisphere(uint64_t* rng, v3 n) {
  const real u1 = next01(rng);
  const real u2 = next01(rng);
  const real r = R_SQRT(smax(RC(0.0), RC(1.0) - u1));
  const real phi = RC(2.0) * kPi * u2;
  const v3 local = mk(r * R_COS(phi), r * R_SIN(phi), R_SQRT(smax(RC(0.0), u1)));
  v3 tangent = cross((R_FABS(n.z) < RC(0.999)) ? mk(0, 0, 1) : mk(0, 1, 0), n);
  tangent = normalize(tangent);
  const v3 bitangent = cross(n, tangent);
  return normalize(add(add(muls(tangent, local.x), muls(bitangent, local.y)), muls(n, local.z)));
}

/* sample_phong_lobe: ScalarCpuPathTracer.cpp:996-1020 */
static v3 sample_phong_lobe(uint64_t* rng, v3 refl, real exponent, v3 n) {
  const real u1 = next01(rng);
  const real u2 = next01(rng);
  const real cosT = R_POW(smax(RC(0.0), u1), RC(1.0) / (exponent + RC(1.0)));
  const real sinT = R_SQRT(smax(RC(0.0), RC(1.0) - cosT * cosT));
  const real phi = RC(2.0) * kPi * u2;
  const v3 ref_t = (R_FABS(refl.z) < RC(0.999)) ? mk(0, 0, 1) : mk(0, 1, 0);
  const v3 tang = normalize(cross(ref_t, refl));
  const v3 btan = cross(refl, tang);
  v3 out = normalize(add(add(muls(tang, sinT * R_COS(phi)), muls(btan, sinT * R_SIN(phi))), muls(refl, cosT)));
  if (dot(out, n) <= RC(0.0)) {
    out = sub(out, muls(n, RC(2.0) * dot(out, n)));
    if (dot(out, n) <= RC(0.0)) return sample_hemisphere(rng, n);
  }
  return out;
}

/* schlick_fresnel / refract_dir: ScalarCpuPathTracer.cpp:357-377 */
static real schlick_fresnel(real cosTheta, real ior) {
  const real safeIor = smax(RC(1.01), ior);
  real r0 = (RC(1.0) - safeIor) / (RC(1.0) + safeIor);
  r0 *= r0;
  const real m = RC(1.0) - smin(RC(1.0), smax(RC(0.0), cosTheta));
  return r0 + (RC(1.0) - r0) * m * m * m * m * m;
}
static v3 refract_dir(v3 incoming, v3 normal, real eta, int* tir) {
  const real cosi = smin(RC(1.0), smax(RC(0.0), dot(neg(incoming), normal)));
  const real k = RC(1.0) - eta * eta * (RC(1.0) - cosi * cosi);
  if (k < RC(0.0)) { *tir = 1; return mk(0, 0, 0); }
  *tir = 0;
  return normalize(add(muls(incoming, eta), muls(normal, eta * cosi - R_SQRT(k))));
}

static v3 reflect_about(v3 d, v3 n) { return sub(d, muls(n, RC(2.0) * dot(n, d))); }

/* SampleSceneEnvironment, PathTracer.cpp:213-272: constant colour, or a bilinear lookup in an
 * equirectangular RGB map. */
static v3 environment(const PtoScene* sc, v3 d) {
  const uint32_t w = sc->environment_map_width, h = sc->environment_map_height;
  if (w == 0u || h == 0u || !sc->environment_map) return v3f(sc->environment_color);
  const real len_sq = d.x * d.x + d.y * d.y + d.z * d.z;
  if (len_sq <= RC(1.0e-12)) return v3f(sc->environment_color);
  const real inv_len = RC(1.0) / R_SQRT(len_sq);
  const v3 dir = mk(d.x * inv_len, d.y * inv_len, d.z * inv_len);
  const real kInvTwoPi = RC(0.15915494309189533577), kInvPiE = RC(0.31830988618379067154);
  real u = RC(0.5) + R_ATAN2(dir.z, dir.x) * kInvTwoPi;
  u -= R_FLOOR(u);
  const real v = R_ACOS(sclamp(dir.y, RC(-1.0), RC(1.0))) * kInvPiE;
  const real xf = u * (real)w;
  const real yf = sclamp(v, RC(0.0), RC(1.0)) * (real)(h - 1u);
  const uint32_t x0 = (uint32_t)R_FLOOR(xf) % w, x1 = (x0 + 1u) % w;
  const uint32_t y0 = (uint32_t)R_FLOOR(yf), y1 = y0 + 1u < h - 1u ? y0 + 1u : h - 1u;
  const real tx = xf - R_FLOOR(xf), ty = yf - R_FLOOR(yf);
  const float* m = sc->environment_map;
  const v3 c00 = v3f(m + 3u * ((size_t)y0 * w + x0)), c10 = v3f(m + 3u * ((size_t)y0 * w + x1));
  const v3 c01 = v3f(m + 3u * ((size_t)y1 * w + x0)), c11 = v3f(m + 3u * ((size_t)y1 * w + x1));
  const v3 top = mk(c00.x + (c10.x - c00.x) * tx, c00.y + (c10.y - c00.y) * tx, c00.z + (c10.z - c00.z) * tx);
  const v3 bot = mk(c01.x + (c11.x - c01.x) * tx, c01.y + (c11.y - c01.y) * tx, c01.z + (c11.z - c01.z) * tx);
  v3 color = mk(top.x + (bot.x - top.x) * ty, top.y + (bot.y - top.y) * ty, top.z + (bot.z - top.z) * ty);
  v3 scale = v3f(sc->environment_map_scale);
  if (smax(scale.x, smax(scale.y, scale.z)) <= RC(1.0e-6)) scale = mk(RC(1.0), RC(1.0), RC(1.0));
  color.x *= scale.x;
  color.y *= scale.y;
  color.z *= scale.z;
  return color;
}

/* ---- path-replay adjoint (no reference counterpart: the reference has no differentiation).
 *
 * The replay re-traces the path with the same random numbers and, knowing the path's total
 * radiance L from the forward pass, derives the gradient of f = sum_c delta_c L_c with respect to
 * material albedos, material emission and light colours (Vicini, Speierer & Jakob 2021, "Path
 * replay backpropagation"):
 *   - emission term T*e*w             -> d/de      = T*w
 *   - NEE term T*D*w (if visible)     -> d/dalbedo = T*w*dD/dalbedo, d/dcolor = T*w*dD/dcolor
 *   - bounce T' = T*bw/rr             -> d/dalbedo via d log bw * R, R = L - (radiance so far)
 *   - the reference divides by rr at every depth (ScalarCpuPathTracer.cpp:1274-1280); below the
 *     roulette start depth that division is deterministic and differentiated (rr = max channel of
 *     T*bw when unclamped, which depends on every earlier albedo); from the start depth on the
 *     roulette is unbiased (survival probability rr, weight 1/rr) and its expectation does not
 *     depend on rr, so no derivative flows through it there.
 * With x_j = log bw_j and rho_k = log rr_k the exact derivative of the realisation is
 *   df/dx_{j,c} = delta_c R_{j,c} + sum_{k >= j active, c*_k = c} Psi_k,
 *   Psi_k = -S_k + S_{next active after k},  S_k = sum_c delta_c R_{k,c},
 * accumulated incrementally below with a history of the vertices before the roulette depth. */
#define PTO_ADJ_HIST 8
typedef struct {
  v3 L_total;
  double delta[3];
  double *g_albedo, *g_emission, *g_light, *g_rough;
  double* g_tex;  /* 3 per texel (NULL: not requested) */
  const PtoScene* sc;
  int pending;
  uint32_t p_mat;
  tap_t p_tap;  /* texture tap of the pending bounce's hit */
  double p_dx[3];
  double p_score;  /* d log(density of the sampled direction) / d roughness of the pending bounce */
  int p_active, p_cstar;
  int nhist;
  uint32_t h_mat[PTO_ADJ_HIST];
  tap_t h_tap[PTO_ADJ_HIST];
  double h_dx[PTO_ADJ_HIST][3];
  int prev_active, prev_cstar, prev_nhist;
} adj_t;

static inline real v3c(v3 v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

/* Adds G = df/d(base colour channel c) of a vertex on material mat: base = albedo * texture tap
 * (the tap's colour is 1 when untextured), so df/dalbedo_c = G * tap_c and
 * df/dtexel_{j,c} = G * albedo_c * w_j for the tap's four texels. */
static void adj_base(adj_t* a, uint32_t mat, const tap_t* tap, int c, double G) {
  if (tap->tex == ~0u) {
    a->g_albedo[3u * mat + c] += G;
    return;
  }
  a->g_albedo[3u * mat + c] += G * (double)v3c(tap->color, c);
  if (!a->g_tex) return;
  const double alb = (double)material_at(a->sc, mat)->albedo[c];
  for (int j = 0; j < 4; ++j) a->g_tex[3u * tap->idx[j] + c] += G * alb * (double)tap->w[j];
}

static void adj_flush_pending(adj_t* a, v3 radiance) {
  if (!a || !a->pending) return;
  const double R[3] = {(double)(a->L_total.x - radiance.x), (double)(a->L_total.y - radiance.y),
                       (double)(a->L_total.z - radiance.z)};
  double S = 0.0;
  for (int c = 0; c < 3; ++c) {
    S += a->delta[c] * R[c];
    adj_base(a, a->p_mat, &a->p_tap, c, a->delta[c] * R[c] * a->p_dx[c]);
  }
  a->g_rough[a->p_mat] += S * a->p_score;
  if (a->p_active) {
    for (int h = 0; h < a->nhist; ++h)
      adj_base(a, a->h_mat[h], &a->h_tap[h], a->p_cstar, -S * a->h_dx[h][a->p_cstar]);
    if (a->prev_active)
      for (int h = 0; h < a->prev_nhist; ++h)
        adj_base(a, a->h_mat[h], &a->h_tap[h], a->prev_cstar, S * a->h_dx[h][a->prev_cstar]);
    a->prev_active = 1;
    a->prev_cstar = a->p_cstar;
    a->prev_nhist = a->nhist;
  }
  a->pending = 0;
}

/* d(clamped effect colour)/d(albedo): k inside the (0, 1.5) clamp, 0 on or outside it. */
static inline double eff_grad(real pre_c, real k) {
  return (pre_c > RC(0.0) && pre_c < RC(1.5)) ? (double)k : 0.0;
}

/* d log p(w) / d e for the direction the glossy branch returned: p is the density of
 * sample_phong_lobe's output (the exponent-e lobe around refl, plus the lobe samples that fell
 * below the surface and were mirrored into w) mixed with the sheen branch's cosine hemisphere of
 * weight ps: p(w) = (1 - ps) [q(w) + q(w')] + ps cos_n / pi, q = (e+1)/(2 pi) max(0, w.refl)^e,
 * w' = w mirrored through the surface plane. The hemisphere fallback for a mirrored sample still
 * below the surface has probability zero. Evaluated in the log domain (e reaches 5e6). */
static double phong_score(double e, v3 refl, v3 n, v3 w, double ps) {
  const double wn = (double)dot(w, n);
  const v3 wm = sub(w, muls(n, RC(2.0) * dot(w, n)));
  const double c[2] = {(double)dot(w, refl), (double)dot(wm, refl)};
  double l[2], a[2];
  int any = 0;
  double L = -INFINITY;
  for (int i = 0; i < 2; ++i) {
    if (c[i] > 0.0) {
      l[i] = e * log(c[i]);
      a[i] = 1.0 / (e + 1.0) + log(c[i]);
      if (l[i] > L) L = l[i];
      any = 1;
    } else {
      l[i] = -INFINITY;
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
    const double lx = log(ps * wn / 3.14159265358979323846) - log((1.0 - ps) * (e + 1.0) / (2.0 * 3.14159265358979323846)) - L;
    den += exp(fmin(lx, 700.0));
  }
  return num / den;
}

/* The glossy branch's direction sampling alone (sample_phong_lobe, then the sheen branch's
 * hemisphere with probability ps), `count` times from one RNG stream, with each direction's
 * phong_score. For pinning the score: its mean over the sampler's own directions must vanish and
 * its covariance with a smooth function must equal that function's mean's derivative in e. */
int pto_phong_samples(float exponent, const float refl_in[3], const float n_in[3], float ps, uint64_t state,
                      uint32_t count, float* out_dirs, double* out_score) {
  const v3 refl = v3f(refl_in), n = v3f(n_in);
  for (uint32_t i = 0; i < count; ++i) {
    v3 w = sample_phong_lobe(&state, refl, (real)exponent, n);
    if (ps > 0.0f && next01(&state) < (real)ps) w = sample_hemisphere(&state, n);
    out_dirs[3 * i] = (float)w.x;
    out_dirs[3 * i + 1] = (float)w.y;
    out_dirs[3 * i + 2] = (float)w.z;
    out_score[i] = phong_score((double)exponent, refl, n, w, (double)ps);
  }
  return 0;
}

/* trace: ScalarCpuPathTracer.cpp:1148-1289; adj != NULL replays the path for the adjoint. */
static v3 trace(const ctx_t* x, v3 ro, v3 rd, uint64_t* rng, uint64_t* rays, PtoCounters* c, adj_t* adj) {
  const PtoScene* sc = x->sc;
  const PtoSettings* st = x->s;
  v3 radiance = mk(0, 0, 0);
  v3 throughput = mk(RC(1.0), RC(1.0), RC(1.0));
  const real rr_min = (real)st->rr_min_survival, rr_max = (real)st->rr_max_survival;
  for (uint32_t depth = 0; depth < x->max_depth; ++depth) {
    ++*rays;
    hit_t hit = intersect_scene(x, ro, rd, c);
    adj_flush_pending(adj, radiance);
    if (!hit.hit) {
      radiance = add(radiance, mulv(throughput, environment(sc, rd)));
      break;
    }
    const v3 gN = hit.normal;
    const int entering = dot(rd, gN) < RC(0.0);
    const v3 sN = entering ? gN : neg(gN);
    const PtoMaterial* m = material_at(sc, hit.material_index);
    real eff_k;
    v3 eff_pre;
    hit_base_color(x->a, m, &hit);
    const v3 albedo = effect_albedo(m, hit.base, hit.position, sN, rd, &eff_k, &eff_pre);
    const real roughness = clamp01((real)m->roughness);
    const int isEmissiveModel = m->material_model == 1u || is_emissive(m);
    const int isMirror = m->material_model == 2u || roughness <= RC(0.001);
    const int isMetallic = m->material_model == 4u || m->metallic > 0.65f;
    const int isTransmissive = m->material_model == 5u || m->transmission > 0.05f;
    const int isDiffuse = roughness >= RC(0.999) && !isMirror && !isMetallic && !isTransmissive;
    const int isSheen = m->material_model == 6u;
    const int isClearcoat = m->material_model == 7u || m->clearcoat > 0.05f;
    const int isToon = m->material_model == 8u;
    const v3 emissive = v3f(m->emissive);

    const uint32_t mat_index = sc->material_count ? (hit.material_index < sc->material_count - 1u
                                                         ? hit.material_index : sc->material_count - 1u) : 0u;
    real e_weight = RC(-1.0);
    if (!st->enable_nee || depth == 0) {
      radiance = add(radiance, mulv(throughput, emissive));
      e_weight = RC(1.0);
    } else if (st->enable_mis && isEmissiveModel) {
      const real bsdf_pdf = smax(RC(0.0), dot(sN, neg(rd))) * kInvPi;
      const real l_pdf = light_pdf_for_direction(sc, ro, rd);
      const real w = mis_weight(bsdf_pdf, l_pdf);
      radiance = add(radiance, muls(mulv(throughput, emissive), w));
      e_weight = w;
    }
    if (adj && e_weight >= RC(0.0) && sc->material_count) {
      const real t3[3] = {throughput.x, throughput.y, throughput.z};
      for (int ch = 0; ch < 3; ++ch) adj->g_emission[3u * mat_index + ch] += adj->delta[ch] * (double)t3[ch] * (double)e_weight;
    }

    if (st->enable_nee && depth < x->max_depth - 1u) {
      hit_t lh = hit;
      lh.normal = sN;
      const nee_t nee = sample_direct_light(x, &lh, neg(rd), rng, c);
      if (nee.valid) {
        real w = RC(1.0);
        if (st->enable_mis) {
          const v3 to_approx = sc->light_count == 0 ? mk(0, 0, 0) : sub(v3f(sc->lights[0].position), hit.position);
          const v3 l_dir = normalize(to_approx);
          const real cos_theta = smax(RC(0.0), dot(sN, l_dir));
          const real bsdf_pdf = cos_theta * kInvPi;
          const real l_pdf = sc->light_count == 0 ? RC(1.0) : (RC(1.0) / (real)sc->light_count);
          w = mis_weight(l_pdf, bsdf_pdf);
          radiance = add(radiance, muls(mulv(throughput, nee.radiance), w));
        } else {
          radiance = add(radiance, mulv(throughput, nee.radiance));
        }
        if (adj) {
          const real t3[3] = {throughput.x, throughput.y, throughput.z};
          const real da[3] = {nee.d_albedo.x, nee.d_albedo.y, nee.d_albedo.z};
          const real dc[3] = {nee.d_color.x, nee.d_color.y, nee.d_color.z};
          for (int ch = 0; ch < 3; ++ch) {
            const double tw = adj->delta[ch] * (double)t3[ch] * (double)w;
            const real pre[3] = {nee.eff_pre.x, nee.eff_pre.y, nee.eff_pre.z};
            adj_base(adj, mat_index, &hit.tap, ch, tw * (double)da[ch] * eff_grad(pre[ch], nee.eff_k));
            adj->g_light[3u * nee.light + ch] += tw * (double)dc[ch];
          }
          const real dr3[3] = {nee.d_rough.x, nee.d_rough.y, nee.d_rough.z};
          for (int ch = 0; ch < 3; ++ch)
            adj->g_rough[mat_index] += adj->delta[ch] * (double)t3[ch] * (double)w * (double)dr3[ch];
        }
      }
    }

    v3 outDir;
    int refracted = 0;
    double score = 0.0;
    if (isMirror) {
      outDir = reflect_about(rd, sN);
    } else if (isTransmissive) {
      const real cosTheta = smax(RC(0.0), dot(sN, neg(rd)));
      const real fresnel = schlick_fresnel(cosTheta, (real)m->ior);
      if (next01(rng) < smin(RC(0.98), fresnel + (real)m->clearcoat * RC(0.015))) {
        outDir = reflect_about(rd, sN);
      } else {
        int tir = 0;
        const real safeIor = smax(RC(1.01), (real)m->ior);
        const real eta = entering ? (RC(1.0) / safeIor) : safeIor;
        outDir = refract_dir(rd, sN, eta, &tir);
        if (tir || length_sq(outDir) <= RC(1.0e-8)) outDir = reflect_about(rd, sN);
        else refracted = 1;
      }
    } else if (isDiffuse || isToon) {
      outDir = sample_hemisphere(rng, sN);
    } else {
      const real effR = smax(RC(0.025), roughness * (isMetallic ? RC(0.75) : RC(1.0)) * (isClearcoat ? RC(0.65) : RC(1.0)));
      const real a2 = effR * effR;
      const real expt = smax(RC(0.0), RC(2.0) / (a2 * a2) - RC(2.0));
      const v3 refl = reflect_about(rd, sN);
      outDir = sample_phong_lobe(rng, refl, expt, sN);
      if (isSheen && next01(rng) < (real)m->sheen * RC(0.35)) outDir = sample_hemisphere(rng, sN);
      if (adj) {
        const double kk = (isMetallic ? 0.75 : 1.0) * (isClearcoat ? 0.65 : 1.0);
        const double ps = isSheen ? fmin(1.0, fmax(0.0, (double)((real)m->sheen * RC(0.35)))) : 0.0;
        const double drp = (m->roughness > 0.0f && m->roughness < 1.0f) ? 1.0 : 0.0;
        if (drp != 0.0 && (double)roughness * kk > 0.025 && expt > RC(0.0))
          score = phong_score((double)expt, refl, sN, outDir, ps) * (-8.0 / pow((double)effR, 5.0)) * kk;
      }
    }

    if (!isTransmissive && dot(outDir, sN) <= RC(0.0)) break;
    v3 bw = albedo;
    if (isMetallic || isMirror) {
      bw = muls(albedo, RC(0.65) + RC(0.35) * (real)m->metallic);
    } else if (isTransmissive) {
      const real tint = clamp01((real)m->alpha) * (real)m->transmission;
      const v3 white = mk(RC(1.0), RC(1.0), RC(1.0));
      bw = add(muls(white, RC(1.0) - tint), muls(albedo, tint));
      if (refracted) bw = muls(bw, smax(RC(0.55), (real)m->transmission));
    }
    if (isToon) bw = muls(bw, dot(outDir, sN) > RC(0.55) ? RC(1.0) : RC(0.45));
    throughput = mulv(throughput, bw);
    if (vmax3(throughput) < RC(0.001)) break;
    const real maxT = vmax3(throughput);
    /* the channel std::max(x, std::max(y, z)) returns: x unless x < max(y, z); then z if y < z, else y */
    const int cstar = (throughput.x < smax(throughput.y, throughput.z)) ? ((throughput.y < throughput.z) ? 2 : 1) : 0;
    const real rr = smin(rr_max, smax(rr_min, maxT));
    if (depth >= st->rr_start_depth && next01(rng) > rr) break;
    throughput = divs(throughput, rr);
    if (adj && sc->material_count) {
      const real a3[3] = {albedo.x, albedo.y, albedo.z};
      const real b3[3] = {bw.x, bw.y, bw.z};
      real tint = RC(0.0);
      if (isTransmissive && !(isMetallic || isMirror)) tint = clamp01((real)m->alpha) * (real)m->transmission;
      adj->pending = 1;
      adj->p_mat = mat_index;
      adj->p_tap = hit.tap;
      adj->p_score = score;
      const real pre3[3] = {eff_pre.x, eff_pre.y, eff_pre.z};
      for (int ch = 0; ch < 3; ++ch) {
        double dx = 0.0;
        if (isTransmissive && !(isMetallic || isMirror)) {
          const real base = (RC(1.0) - tint) + a3[ch] * tint; /* bw before the refraction / toon factors */
          dx = base > RC(0.0) ? (double)tint / (double)base : 0.0;
        } else {
          dx = a3[ch] > RC(0.0) ? 1.0 / (double)a3[ch] : 0.0;
        }
        adj->p_dx[ch] = b3[ch] > RC(0.0) ? dx * eff_grad(pre3[ch], eff_k) : 0.0;
      }
      adj->p_active = (depth < st->rr_start_depth && rr_min < maxT && maxT < rr_max) ? 1 : 0;
      adj->p_cstar = cstar;
      if (depth < st->rr_start_depth && adj->nhist < PTO_ADJ_HIST) {
        adj->h_mat[adj->nhist] = mat_index;
        adj->h_tap[adj->nhist] = hit.tap;
        for (int ch = 0; ch < 3; ++ch) adj->h_dx[adj->nhist][ch] = adj->p_dx[ch];
        adj->nhist++;
      }
    }
    ro = add(hit.position, muls(outDir, RC(0.002)));
    rd = outDir;
  }
  return radiance;
}

typedef struct {
  v3 cam_pos, fwd, right, up;
  real aspect, tan_half_fov, aperture, focus;
  uint32_t iris_blades;
  real iris_roundness, iris_rotation, anamorphic;
} camera_t;

/* compute_camera_basis: ScalarCpuPathTracer.cpp:139-152; per-frame constants of camera_rays
 * (ScalarCpuPathTracer.Render.cpp:74-83). tanf runs on the host in both oracle and product. */
static camera_t make_camera(const PtoScene* sc, const PtoSettings* s) {
  camera_t cam;
  const v3 pos = v3f(sc->camera.position), tgt = v3f(sc->camera.target), upv = v3f(sc->camera.up);
  cam.cam_pos = pos;
  cam.fwd = normalize(sub(tgt, pos));
  cam.right = normalize(cross(cam.fwd, upv));
  if (length_sq(cam.right) <= kEps * kEps) cam.right = mk(1, 0, 0);
  cam.up = normalize(cross(cam.right, cam.fwd));
  if (length_sq(cam.up) <= kEps * kEps) cam.up = mk(0, 1, 0);
  const real hf = (real)(s->height > 0 ? s->height : 0);
  cam.aspect = (real)s->width / smax(RC(1.0), hf);
  const real radians = (real)sc->camera.fov_deg * (kPi / RC(180.0));
#if PTO_REAL_IS_DOUBLE
  cam.tan_half_fov = tan(RC(0.5) * radians);
#else
  cam.tan_half_fov = tanf(RC(0.5) * radians);
#endif
  cam.aperture = sc->camera.aperture_radius > 0.0f ? (real)sc->camera.aperture_radius : (real)s->camera_aperture_radius;
  cam.focus = sc->camera.focus_distance > 0.0f ? (real)sc->camera.focus_distance : (real)s->camera_focus_distance;
  cam.iris_blades = sc->camera.iris_blade_count < 64u ? sc->camera.iris_blade_count : 64u;
  cam.iris_roundness = clamp01((real)sc->camera.iris_roundness);
  cam.iris_rotation = (real)sc->camera.iris_rotation_degrees * (kPi / RC(180.0));
  cam.anamorphic = isfinite(sc->camera.anamorphic_squeeze) ? smax(RC(0.01), (real)sc->camera.anamorphic_squeeze) : RC(1.0);
  return cam;
}

/* The pinhole direction of camera_rays (ScalarCpuPathTracer.Render.cpp:74-81) for the pixel jitter (u, v). */
static v3 camera_dir(const camera_t* cam, const PtoSettings* s, uint32_t x, uint32_t y, real u, real v) {
  const real fx = ((real)x + u) / (real)(s->width > 1u ? s->width : 1u);
  const real fy = ((real)y + v) / (real)(s->height > 1u ? s->height : 1u);
  const real nx = (RC(2.0) * fx - RC(1.0)) * cam->aspect * cam->tan_half_fov;
  const real ny = (RC(1.0) - RC(2.0) * fy) * cam->tan_half_fov;
  return normalize(add(add(cam->fwd, muls(cam->right, nx)), muls(cam->up, ny)));
}

/* camera_rays: ScalarCpuPathTracer.Render.cpp:55-119. jitter (may be NULL) receives the two pixel draws. */
static void camera_ray(const camera_t* cam, const PtoSettings* s, uint32_t x, uint32_t y, uint32_t sample,
                       uint64_t frame, uint32_t path_id, uint64_t* sample_seed, v3* ro, v3* rd, real* jitter) {
  uint64_t rng = pto_rng_seed(s->seed, frame, sample, path_id, 0, 0, (uint64_t)y * s->width + x);
  const real u = next01(&rng);
  const real v = next01(&rng);
  if (jitter) { jitter[0] = u; jitter[1] = v; }
  *sample_seed ^= (uint64_t)(next01(&rng) * (real)0xffffu);
  const v3 dir = camera_dir(cam, s, x, y, u, v);
  if (cam->aperture > RC(0.0) && cam->focus > kEps) {
    const real lens_radius_sample = R_SQRT(next01(&rng));
    const real lens_phi = RC(2.0) * kPi * next01(&rng);
    real boundary = RC(1.0);
    if (cam->iris_blades >= 3u && cam->iris_roundness < RC(0.999)) {
      const real sector = RC(2.0) * kPi / (real)cam->iris_blades;
      const real local_phi = lens_phi - cam->iris_rotation;
      const real wrapped = local_phi - sector * R_FLOOR(local_phi / sector);
      const real centered = wrapped > sector * RC(0.5) ? wrapped - sector : wrapped;
      const real polygon = R_COS(sector * RC(0.5)) / smax(RC(0.1), R_COS(centered));
      boundary = polygon * (RC(1.0) - cam->iris_roundness) + cam->iris_roundness;
    }
    const real lens_r = cam->aperture * lens_radius_sample * boundary;
    const v3 offset = add(muls(cam->right, lens_r * R_COS(lens_phi) * cam->anamorphic), muls(cam->up, lens_r * R_SIN(lens_phi)));
    const v3 focus_point = add(cam->cam_pos, muls(dir, cam->focus));
    const v3 origin = add(cam->cam_pos, offset);
    *ro = origin;
    *rd = normalize(sub(focus_point, origin));
    return;
  }
  *ro = cam->cam_pos;
  *rd = dir;
}

/* shade_pixel: ScalarCpuPathTracer.Render.cpp:228-284 */
static int shade_pixel(const ctx_t* x, const camera_t* cam, uint32_t pixel, uint32_t sample, uint64_t frame,
                       v3* out, uint64_t* rays, PtoCounters* c, adj_t* adj) {
  const PtoSettings* s = x->s;
  const uint32_t px = pixel % s->width, py = pixel / s->width;
  uint64_t seed = 0;
  const uint64_t sample_seed = ((uint64_t)py << 32) | px;
  const uint64_t path_id = sample_seed + (uint64_t)sample + frame;
  v3 ro, rd;
  camera_ray(cam, s, px, py, sample, frame, (uint32_t)path_id, &seed, &ro, &rd, NULL);
  uint64_t rng = pto_rng_seed(s->seed + sample_seed + seed, frame, sample, path_id, 0, 0, sample_seed);
  const v3 v = trace(x, ro, rd, &rng, rays, c, adj);
  *out = v;
  return is_finite_r(v.x) && is_finite_r(v.y) && is_finite_r(v.z);
}

/* Probes of single oracle functions, for tests that check them against closed forms. */

/* The primary ray shade_pixel makes for one sample of each listed pixel (camera_rays with the path id
 * and RNG key of Render.cpp:246-251): origin and direction (6 per pixel) and the two pixel jitter draws
 * (2 per pixel, may be NULL). */
int pto_camera_rays(const PtoScene* scene, const PtoSettings* s, const uint32_t* pixels, uint32_t n,
                    uint32_t sample, uint64_t frame, float* out_rays, float* out_jitter) {
  if (!scene || !s || s->width == 0 || s->height == 0) return -1;
  const camera_t cam = make_camera(scene, s);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t px = pixels[i] % s->width, py = pixels[i] / s->width;
    const uint64_t sample_seed = ((uint64_t)py << 32) | px;
    const uint64_t path_id = sample_seed + (uint64_t)sample + frame;
    uint64_t seed = 0;
    v3 ro, rd;
    real jit[2];
    camera_ray(&cam, s, px, py, sample, frame, (uint32_t)path_id, &seed, &ro, &rd, jit);
    float* r = out_rays + 6u * i;
    r[0] = (float)ro.x; r[1] = (float)ro.y; r[2] = (float)ro.z;
    r[3] = (float)rd.x; r[4] = (float)rd.y; r[5] = (float)rd.z;
    if (out_jitter) { out_jitter[2u * i] = (float)jit[0]; out_jitter[2u * i + 1u] = (float)jit[1]; }
  }
  return 0;
}

/* The pinhole direction of camera_rays through pixel (x, y) at jitter (u, v). */
int pto_camera_dir(const PtoScene* scene, const PtoSettings* s, uint32_t x, uint32_t y, float u, float v,
                   float out_dir[3]) {
  if (!scene || !s || s->width == 0 || s->height == 0) return -1;
  const camera_t cam = make_camera(scene, s);
  const v3 d = camera_dir(&cam, s, x, y, (real)u, (real)v);
  out_dir[0] = (float)d.x; out_dir[1] = (float)d.y; out_dir[2] = (float)d.z;
  return 0;
}

/* sample_hemisphere directions about n from the RNG state `state` (each draw advances it). */
int pto_hemisphere_samples(const float n_in[3], uint64_t state, uint32_t count, float* out_dirs) {
  const v3 n = v3f(n_in);
  for (uint32_t i = 0; i < count; ++i) {
    const v3 w = sample_hemisphere(&state, n);
    out_dirs[3 * i] = (float)w.x; out_dirs[3 * i + 1] = (float)w.y; out_dirs[3 * i + 2] = (float)w.z;
  }
  return 0;
}

/* light_pdf_for_direction (ScalarCpuPathTracer.cpp:1120-1138) at a position for a direction. */
double pto_light_pdf(const PtoScene* scene, const float pos[3], const float dir[3]) {
  return (double)light_pdf_for_direction(scene, v3f(pos), v3f(dir));
}

static int check_supported(const PtoScene* sc) {
  if (sc->environment_map && sc->environment_map_width && sc->environment_map_height == 0u) return -3;
  return 0;
}

static int any_transmissive(const PtoScene* sc) {
  for (uint32_t i = 0; i < sc->material_count; ++i)
    if (sc->materials[i].material_model == 5u || sc->materials[i].transmission > 0.05f) return 1;
  return 0;
}

static void add_counters(PtoCounters* d, const PtoCounters* s) {
  d->sdf_tests += s->sdf_tests; d->sdf_steps += s->sdf_steps; d->sdf_hits += s->sdf_hits;
  d->sdf_misses += s->sdf_misses;
  d->samples += s->samples; d->rays += s->rays; d->triangle_tests += s->triangle_tests;
  d->triangle_hits += s->triangle_hits; d->bvh_node_visits += s->bvh_node_visits;
  d->bvh_leaf_visits += s->bvh_leaf_visits; d->shadow_tests += s->shadow_tests;
}

int pto_render_pixels(const PtoScene* scene, const pto_accel* a, const PtoSettings* s,
                      const uint32_t* pixels, uint32_t n, uint32_t sample_index, uint64_t frame_index,
                      int threads, int shadow_mode, void* out_rgb, uint8_t* out_valid, PtoCounters* counters) {
  if (!scene || !a || !s || s->width == 0 || s->height == 0) return -1;
  const int sup = check_supported(scene);
  if (sup) return sup;
  if (shadow_mode == PTO_SHADOW_ANYHIT && (any_transmissive(scene) || a->sdf_count)) shadow_mode = PTO_SHADOW_CLOSEST;
  ctx_t x = {scene, a, s, s->max_depth > 1u ? s->max_depth : 1u, shadow_mode};
  const camera_t cam = make_camera(scene, s);
  real* rgb = (real*)out_rgb;
  PtoCounters total;
  memset(&total, 0, sizeof(total));
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel
  {
    PtoCounters lc;
    memset(&lc, 0, sizeof(lc));
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      v3 v;
      uint64_t rays = 0;
      const uint64_t total_px = (uint64_t)s->width * s->height;
      if ((uint64_t)pixels[i] >= total_px) { out_valid[i] = 0; continue; }
      const int ok = shade_pixel(&x, &cam, pixels[i], sample_index, frame_index, &v, &rays, &lc, NULL);
      rgb[3 * i] = v.x; rgb[3 * i + 1] = v.y; rgb[3 * i + 2] = v.z;
      out_valid[i] = (uint8_t)ok;
      /* shade_pixel adds ray_count only for a finite sample (Render.cpp:270-281) */
      if (ok) {
        lc.rays += rays;
        lc.samples++;
      }
    }
#pragma omp critical
    add_counters(&total, &lc);
  }
  if (counters) *counters = total;
  return 0;
}

int pto_render_tile(const PtoScene* scene, const pto_accel* a, const PtoSettings* s,
                    uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                    uint32_t sample_begin, uint32_t sample_count, uint64_t frame_index,
                    int threads, int shadow_mode, void* film_sum, uint32_t* film_count, PtoCounters* counters) {
  if (!scene || !a || !s || s->width == 0 || s->height == 0) return -1;
  const int sup = check_supported(scene);
  if (sup) return sup;
  if (shadow_mode == PTO_SHADOW_ANYHIT && (any_transmissive(scene) || a->sdf_count)) shadow_mode = PTO_SHADOW_CLOSEST;
  ctx_t x = {scene, a, s, s->max_depth > 1u ? s->max_depth : 1u, shadow_mode};
  const camera_t cam = make_camera(scene, s);
  real* film = (real*)film_sum;
  const uint32_t xa = x0 < s->width ? x0 : s->width, ya = y0 < s->height ? y0 : s->height;
  const uint32_t xb = (uint64_t)xa + w < s->width ? xa + w : s->width;
  const uint32_t yb = (uint64_t)ya + h < s->height ? ya + h : s->height;
  const int64_t tw = (int64_t)(xb - xa), npx = tw * (int64_t)(yb - ya);
  PtoCounters total;
  memset(&total, 0, sizeof(total));
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  for (uint32_t si = 0; si < sample_count; ++si) {
    const uint32_t sample = sample_begin + si;
#pragma omp parallel
    {
      PtoCounters lc;
      memset(&lc, 0, sizeof(lc));
#pragma omp for schedule(dynamic, 64)
      for (int64_t i = 0; i < npx; ++i) {
        const uint32_t px = xa + (uint32_t)(i % tw), py = ya + (uint32_t)(i / tw);
        const uint32_t pixel = py * s->width + px;
        v3 v;
        uint64_t rays = 0;
        const int ok = shade_pixel(&x, &cam, pixel, sample, frame_index, &v, &rays, &lc, NULL);
        /* shade_pixel drops a non-finite sample before FilmBuffer::add_sample and before it adds the
         * sample's rays to ray_count (Render.cpp:270-281) */
        if (!ok) continue;
        lc.rays += rays;
        lc.samples++;
        real* f = film + 3u * (uint64_t)pixel;
        f[0] = f[0] + v.x;
        f[1] = f[1] + v.y;
        f[2] = f[2] + v.z;
        film_count[pixel] += 1u;
      }
#pragma omp critical
      add_counters(&total, &lc);
    }
  }
  if (counters) *counters = total;
  return 0;
}

/* TiledCpuPathTracer's schedule (TiledCpuPathTracer.cpp:279-323, :663-773): for each sample, the
 * image is cut into bands of band_rows rows and each band is one job rendered by one thread in
 * row-major order (ScalarCpuPathTracer::render_tile on the band), jobs taken dynamically by the
 * workers. Bands band_first, band_first + band_stride, ... below the image height are rendered;
 * samples [sample_begin, sample_begin + sample_count) are added to the film in order, so the film
 * equals pto_render_tile's over the same rows. Used to time the CPU baseline. */
int pto_render_bands(const PtoScene* scene, const pto_accel* a, const PtoSettings* s, uint32_t band_rows,
                     uint32_t band_first, uint32_t band_stride, uint32_t sample_begin, uint32_t sample_count,
                     uint64_t frame_index, int threads, int shadow_mode, void* film_sum, uint32_t* film_count,
                     PtoCounters* counters) {
  if (!scene || !a || !s || s->width == 0 || s->height == 0 || band_rows == 0 || band_stride == 0) return -1;
  const int sup = check_supported(scene);
  if (sup) return sup;
  if (shadow_mode == PTO_SHADOW_ANYHIT && (any_transmissive(scene) || a->sdf_count)) shadow_mode = PTO_SHADOW_CLOSEST;
  ctx_t x = {scene, a, s, s->max_depth > 1u ? s->max_depth : 1u, shadow_mode};
  const camera_t cam = make_camera(scene, s);
  real* film = (real*)film_sum;
  const uint32_t n_bands = (s->height + band_rows - 1u) / band_rows;
  const int64_t jobs = band_first < n_bands ? (int64_t)((n_bands - 1u - band_first) / band_stride + 1u) : 0;
  PtoCounters total;
  memset(&total, 0, sizeof(total));
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  for (uint32_t si = 0; si < sample_count; ++si) {
    const uint32_t sample = sample_begin + si;
#pragma omp parallel
    {
      PtoCounters lc;
      memset(&lc, 0, sizeof(lc));
#pragma omp for schedule(dynamic, 1)
      for (int64_t j = 0; j < jobs; ++j) {
        const uint32_t y0 = (band_first + (uint32_t)j * band_stride) * band_rows;
        const uint32_t y1 = y0 + band_rows < s->height ? y0 + band_rows : s->height;
        for (uint32_t py = y0; py < y1; ++py)
          for (uint32_t px = 0; px < s->width; ++px) {
            const uint32_t pixel = py * s->width + px;
            v3 v;
            uint64_t rays = 0;
            const int ok = shade_pixel(&x, &cam, pixel, sample, frame_index, &v, &rays, &lc, NULL);
            if (!ok) continue;
            lc.rays += rays;
            lc.samples++;
            real* f = film + 3u * (uint64_t)pixel;
            f[0] = f[0] + v.x;
            f[1] = f[1] + v.y;
            f[2] = f[2] + v.z;
            film_count[pixel] += 1u;
          }
      }
#pragma omp critical
      add_counters(&total, &lc);
    }
  }
  if (counters) *counters = total;
  return 0;
}

/* Path-replay adjoint for one sample of each listed pixel. delta: 3 floats per listed pixel,
 * dLoss / d(the path's radiance) (= dLoss / d film sum at that pixel). The forward radiance is
 * recomputed first (nothing is stored between the passes). Gradients are accumulated in double:
 * g_albedo, g_emission (3 per material), g_light (3 per light colour), g_rough (1 per material,
 * the roughness parameter; may be NULL); out_rgb (may be NULL)
 * receives the forward radiance. Paths with non-finite radiance contribute nothing. */
int pto_adjoint_pixels(const PtoScene* scene, const pto_accel* a, const PtoSettings* s,
                       const uint32_t* pixels, uint32_t n, uint32_t sample_index, uint64_t frame_index,
                       int threads, int shadow_mode, const float* delta, double* g_albedo, double* g_emission,
                       double* g_light, double* g_rough, double* g_tex, void* out_rgb) {
  if (!scene || !a || !s || s->width == 0 || s->height == 0) return -1;
  const int sup = check_supported(scene);
  if (sup) return sup;
  if (s->rr_start_depth > PTO_ADJ_HIST) return -4;
  if (shadow_mode == PTO_SHADOW_ANYHIT && (any_transmissive(scene) || a->sdf_count)) shadow_mode = PTO_SHADOW_CLOSEST;
  ctx_t x = {scene, a, s, s->max_depth > 1u ? s->max_depth : 1u, shadow_mode};
  const camera_t cam = make_camera(scene, s);
  const uint32_t nm = scene->material_count, nl = scene->light_count;
  real* rgb = (real*)out_rgb;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel
  {
    double* ga = (double*)calloc(3u * (nm ? nm : 1u), sizeof(double));
    double* ge = (double*)calloc(3u * (nm ? nm : 1u), sizeof(double));
    double* gl = (double*)calloc(3u * (nl ? nl : 1u), sizeof(double));
    double* gr = (double*)calloc(nm ? nm : 1u, sizeof(double));
    const size_t ntx = g_tex ? 3u * (size_t)a->texel_count : 0u;
    double* gt = ntx ? (double*)calloc(ntx, sizeof(double)) : NULL;
    PtoCounters lc;
    memset(&lc, 0, sizeof(lc));
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      v3 L;
      uint64_t rays = 0;
      const uint64_t total_px = (uint64_t)s->width * s->height;
      if ((uint64_t)pixels[i] >= total_px) continue;
      const int ok = shade_pixel(&x, &cam, pixels[i], sample_index, frame_index, &L, &rays, &lc, NULL);
      if (rgb) { rgb[3 * i] = L.x; rgb[3 * i + 1] = L.y; rgb[3 * i + 2] = L.z; }
      if (!ok) continue;
      adj_t adj;
      memset(&adj, 0, sizeof(adj));
      adj.L_total = L;
      for (int ch = 0; ch < 3; ++ch) adj.delta[ch] = (double)delta[3 * i + ch];
      adj.g_albedo = ga;
      adj.g_emission = ge;
      adj.g_light = gl;
      adj.g_rough = gr;
      adj.g_tex = gt;
      adj.sc = scene;
      v3 L2;
      shade_pixel(&x, &cam, pixels[i], sample_index, frame_index, &L2, &rays, &lc, &adj);
    }
#pragma omp critical
    {
      for (uint32_t k = 0; k < 3u * nm; ++k) { g_albedo[k] += ga[k]; g_emission[k] += ge[k]; }
      for (uint32_t k = 0; k < 3u * nl; ++k) g_light[k] += gl[k];
      if (g_rough)
        for (uint32_t k = 0; k < nm; ++k) g_rough[k] += gr[k];
      for (size_t k = 0; k < ntx; ++k) g_tex[k] += gt[k];
    }
    free(gt);
    free(gr);
    free(ga);
    free(ge);
    free(gl);
  }
  return 0;
}

/* ------------------------------------------------------------------ film resolve
 * FilmBuffer::resolve_hdr (FilmBuffer.cpp:160-214) and ApplyFilmResolve (FilmBuffer.cpp:255-332) with
 * WhiteBalanceScale / ColorTemperatureToRgb (FilmBuffer.cpp:30-49, :216-224). The colour-temperature fit
 * is the published Tanner Helland approximation (coefficients / 255), evaluated with libm as the
 * reference's host code does; the per-pixel gamma power uses this build's powf (portable or libm). */
#if PTO_LIBM || PTO_REAL_IS_DOUBLE
#define RESOLVE_POW(x, y) powf((x), (y))
#else
#define RESOLVE_POW(x, y) pm_powf((x), (y))
#endif

/* std::max / std::min / std::clamp on floats, with their comparison semantics. */
static inline float fmaxf_std(float a, float b) { return (a < b) ? b : a; }
static inline float fminf_std(float a, float b) { return (b < a) ? b : a; }
static inline float fclampf(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

void pto_color_temperature_rgb(float kelvin, float out[3]) {
  const float t = fclampf(kelvin, 1000.0f, 40000.0f) / 100.0f;
  float r = 1.0f, g = 1.0f, b = 1.0f;
  if (t <= 66.0f) {
    g = fclampf(0.39008158f * logf(fmaxf_std(1.0f, t)) - 0.63184144f, 0.0f, 1.0f);
    b = t <= 19.0f ? 0.0f : fclampf(0.5432068f * logf(fmaxf_std(1.0f, t - 10.0f)) - 1.1962541f, 0.0f, 1.0f);
  } else {
    r = fclampf(1.2929362f * powf(t - 60.0f, -0.13320476f), 0.0f, 1.0f);
    g = fclampf(1.1298909f * powf(t - 60.0f, -0.075514846f), 0.0f, 1.0f);
  }
  out[0] = r;
  out[1] = g;
  out[2] = b;
}

void pto_white_balance_scale(float kelvin, float out[3]) {
  float d65[3], tgt[3];
  pto_color_temperature_rgb(6500.0f, d65);
  pto_color_temperature_rgb(kelvin, tgt);
  for (int k = 0; k < 3; ++k) out[k] = d65[k] / fmaxf_std(1.0e-3f, tgt[k]);
}

void pto_resolve_hdr(const float* sum3, const uint32_t* count, uint64_t pixels, float* rgb) {
  for (uint64_t i = 0; i < pixels; ++i) {
    const float inv = 1.0f / (float)(count[i] > 1u ? count[i] : 1u);
    for (int k = 0; k < 3; ++k) rgb[3 * i + k] = sum3[3 * i + k] * inv;
  }
}

static float resolve_tone(float x, uint32_t mode) {
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

/* static_cast<uint8_t>(float) as x86-64 code evaluates it: truncation to int32 (INT_MIN for NaN and
 * out-of-range values), then the low byte. */
static uint8_t to_byte(float v) {
  const int32_t i = (v > -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : INT32_MIN;
  return (uint8_t)(uint32_t)i;
}

void pto_resolve_ldr(const float* sum3, const uint32_t* count, uint64_t pixels, const PtoResolve* rs,
                     uint8_t* rgba) {
  for (uint64_t i = 0; i < pixels; ++i) {
    const float inv = 1.0f / (float)(count[i] > 1u ? count[i] : 1u);
    float c[3];
    for (int k = 0; k < 3; ++k) {
      float v = sum3[3 * i + k] * inv * rs->exposure * rs->white_balance[k];
      if (!isfinite(v)) v = 0.0f;
      v = resolve_tone(v, rs->tone_map);
      if (rs->output_transform == 0) v = RESOLVE_POW(fmaxf_std(0.0f, v), rs->inv_gamma);
      if (rs->clamp_output) v = fminf_std(1.0f, fmaxf_std(0.0f, v));
      c[k] = v;
    }
    for (int k = 0; k < 3; ++k) rgba[4 * i + k] = to_byte(c[k] * 255.0f + 0.5f);
    rgba[4 * i + 3] = 255u;
  }
}
