// pt_bvh.cpp — flattening and BVH construction on the host.
//
// Both builders reproduce the reference's split decisions exactly (same float
// expressions, same stable partition, same node numbering), so the device
// traverses the very tree the reference's algorithm would build.
#include "pt_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace ptb {
namespace {

constexpr float kEps = 1e-4f;

inline float smin(float a, float b) { return (b < a) ? b : a; }  // std::min
inline float smax(float a, float b) { return (a < b) ? b : a; }  // std::max
inline float min3(float a, float b, float c) { float m = a; if (b < m) m = b; if (c < m) m = c; return m; }
inline float max3(float a, float b, float c) { float m = a; if (m < b) m = b; if (m < c) m = c; return m; }

struct V3 { float x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 muls(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 load3(const float* p) { return {p[0], p[1], p[2]}; }

// normalize_vec3, CpuBvhAccelerator.cpp:76-82
inline V3 normalize_accel(V3 v) {
  const float len_sq = dot(v, v);
  if (len_sq <= kEps * kEps) return {0.0f, 1.0f, 0.0f};
  return muls(v, 1.0f / std::sqrt(len_sq));
}

// rotate_quat / transform_instance_vertex, CpuBvhAccelerator.cpp:48-74
V3 rotate_quat(V3 value, const float* q) {
  const float len_sq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  if (len_sq <= kEps * kEps) return value;
  const float inv_len = 1.0f / std::sqrt(len_sq);
  const V3 qv{q[0] * inv_len, q[1] * inv_len, q[2] * inv_len};
  const float qw = q[3] * inv_len;
  const V3 t = muls(cross(qv, value), 2.0f);
  return add(add(value, muls(t, qw)), cross(qv, t));
}

V3 transform_vertex(V3 local, const PtInstance& inst) {
  const V3 scaled{local.x * inst.scale[0], local.y * inst.scale[1], local.z * inst.scale[2]};
  return add(rotate_quat(scaled, inst.rotation), load3(inst.translation));
}

struct Ref {
  uint32_t orig;
  float c[3], mn[3], mx[3];
};

float bvh_sa(const float* mn, const float* mx) {
  float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
  if (dx < 0) dx = 0;
  if (dy < 0) dy = 0;
  if (dz < 0) dz = 0;
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct SahBuilder {
  std::vector<Ref>& refs;
  std::vector<Ref> tmp;
  Bvh& bvh;
  uint32_t leaf, buckets;

  uint32_t build(uint32_t start, uint32_t count, uint32_t depth) {
    bvh.max_depth = std::max(bvh.max_depth, depth);
    const uint32_t ni = static_cast<uint32_t>(bvh.nodes.size());
    bvh.nodes.push_back(Node{});
    float bmn[3] = {1e30f, 1e30f, 1e30f}, bmx[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = start; i < start + count; ++i)
      for (int k = 0; k < 3; ++k) {
        bmn[k] = smin(bmn[k], refs[i].mn[k]);
        bmx[k] = smax(bmx[k], refs[i].mx[k]);
      }
    {
      Node& node = bvh.nodes[ni];
      for (int k = 0; k < 3; ++k) {
        node.mn[k] = bmn[k] - kEps;
        node.mx[k] = bmx[k] + kEps;
      }
      if (count <= leaf) {
        node.left = node.right = -1;
        node.first = static_cast<int32_t>(bvh.order.size());
        node.count = static_cast<int32_t>(count);
        for (uint32_t i = start; i < start + count; ++i) bvh.order.push_back(refs[i].orig);
        ++bvh.leaf_count;
        return ni;
      }
    }
    const int B = static_cast<int>(std::clamp(buckets, 2u, 16u));
    const float node_sa = smax(1e-9f, bvh_sa(bmn, bmx));
    int best_axis = 0;
    float best_split = 0.0f, best_cost = 1e30f;
    bool found = false;
    for (int axis = 0; axis < 3; ++axis) {
      float cmin = 1e30f, cmax = -1e30f;
      for (uint32_t i = start; i < start + count; ++i) {
        cmin = smin(cmin, refs[i].c[axis]);
        cmax = smax(cmax, refs[i].c[axis]);
      }
      const float extent = cmax - cmin;
      if (extent < 1e-6f) continue;
      float bmn_b[16][3], bmx_b[16][3];
      uint32_t bn[16];
      for (int b = 0; b < B; ++b) {
        bn[b] = 0;
        for (int k = 0; k < 3; ++k) { bmn_b[b][k] = 1e30f; bmx_b[b][k] = -1e30f; }
      }
      const float fB = static_cast<float>(B);
      for (uint32_t i = start; i < start + count; ++i) {
        // the reference converts this to int unchecked; a non-finite vertex makes it NaN or infinite,
        // where the conversion is undefined: such primitives go to bucket 0 (or the last, for +inf)
        float x = (refs[i].c[axis] - cmin) / extent * fB;
        if (!(x >= 0.0f)) x = 0.0f;
        int b = x < fB ? static_cast<int>(x) : B - 1;
        if (B - 1 < b) b = B - 1;
        bn[b]++;
        for (int k = 0; k < 3; ++k) {
          bmn_b[b][k] = smin(bmn_b[b][k], refs[i].mn[k]);
          bmx_b[b][k] = smax(bmx_b[b][k], refs[i].mx[k]);
        }
      }
      for (int s = 1; s < B; ++s) {
        float lmn[3] = {1e30f, 1e30f, 1e30f}, lmx[3] = {-1e30f, -1e30f, -1e30f};
        float rmn[3] = {1e30f, 1e30f, 1e30f}, rmx[3] = {-1e30f, -1e30f, -1e30f};
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
        const float cost = (bvh_sa(lmn, lmx) * static_cast<float>(lc) + bvh_sa(rmn, rmx) * static_cast<float>(rc)) / node_sa;
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_split = cmin + static_cast<float>(s) / fB * extent;
          found = true;
        }
      }
    }
    uint32_t mid;
    if (found) {
      uint32_t nl = 0, nr = 0;
      for (uint32_t i = start; i < start + count; ++i) {
        if (refs[i].c[best_axis] < best_split) refs[start + nl++] = refs[i];
        else tmp[nr++] = refs[i];
      }
      std::copy(tmp.begin(), tmp.begin() + nr, refs.begin() + start + nl);
      mid = start + nl;
      if (mid <= start || mid >= start + count) mid = start + count / 2u;
    } else {
      mid = start + count / 2u;
    }
    const uint32_t l = build(start, mid - start, depth + 1);
    const uint32_t r = build(mid, start + count - mid, depth + 1);
    Node& node = bvh.nodes[ni];
    node.left = static_cast<int32_t>(l);
    node.right = static_cast<int32_t>(r);
    node.first = -1;
    node.count = 0;
    return ni;
  }
};

struct MidBuilder {
  const std::vector<float>& boxes;  // 6 per primitive (padded)
  std::vector<uint32_t> tmp;
  Bvh& bvh;
  uint32_t leaf;
  int32_t alloc = 1;

  void build(int32_t ni, uint32_t* idx, uint32_t count, uint32_t depth) {
    bvh.max_depth = std::max(bvh.max_depth, depth);
    Node b;
    const float fmax = std::numeric_limits<float>::max();
    for (int k = 0; k < 3; ++k) { b.mn[k] = fmax; b.mx[k] = -fmax; }
    for (uint32_t i = 0; i < count; ++i) {
      const float* a = &boxes[6u * idx[i]];
      for (int k = 0; k < 3; ++k) { b.mn[k] = smin(b.mn[k], a[k]); b.mx[k] = smax(b.mx[k], a[3 + k]); }
    }
    Node& node = bvh.nodes[static_cast<size_t>(ni)];
    std::memcpy(node.mn, b.mn, sizeof(b.mn));
    std::memcpy(node.mx, b.mx, sizeof(b.mx));
    if (count <= leaf) {
      node.left = node.right = -1;
      node.first = static_cast<int32_t>(idx - bvh.order.data());
      node.count = static_cast<int32_t>(count);
      ++bvh.leaf_count;
      return;
    }
    float cmn[3], cmx[3];
    for (int k = 0; k < 3; ++k) { cmn[k] = fmax; cmx[k] = -fmax; }
    for (uint32_t i = 0; i < count; ++i) {
      const float* a = &boxes[6u * idx[i]];
      for (int k = 0; k < 3; ++k) {
        const float c = 0.5f * (a[k] + a[3 + k]);
        cmn[k] = smin(cmn[k], c);
        cmx[k] = smax(cmx[k], c);
      }
    }
    int axis = 0;
    float best = cmx[0] - cmn[0];
    for (int k = 1; k < 3; ++k) {
      const float e = cmx[k] - cmn[k];
      if (e > best) { best = e; axis = k; }
    }
    const float split = 0.5f * (cmn[axis] + cmx[axis]);
    uint32_t nl = 0, nr = 0;
    for (uint32_t i = 0; i < count; ++i) {
      const float* a = &boxes[6u * idx[i]];
      const float c = 0.5f * (a[axis] + a[3 + axis]);
      if (c < split) idx[nl++] = idx[i];
      else tmp[nr++] = idx[i];
    }
    std::copy(tmp.begin(), tmp.begin() + nr, idx + nl);
    const uint32_t safe_left = (nl == 0 || nl == count) ? count / 2 : nl;
    const int32_t li = alloc++;
    const int32_t ri = alloc++;
    node.left = li;
    node.right = ri;
    node.first = -1;
    node.count = 0;
    build(li, idx, safe_left, depth + 1);
    build(ri, idx + safe_left, count - safe_left, depth + 1);
  }
};

}  // namespace

// One triangle of an instance (CpuBvhAccelerator.cpp:497-560's primitive record). Returns false, and
// writes nothing, for a triangle whose indices are out of range (the reference skips it).
static bool make_prim(const PtSceneView& sc, const PtInstance& inst, uint32_t tri, const float* texcoords, Prim& p,
               float* r9) {
  const bool use_local = (inst.flags & 1u) != 0u && inst.local_vertex_count > 0u && inst.local_index_count > 0u;
  V3 a{0, 0, 0}, b{0, 0, 0}, c{0, 0, 0};
  float uv[6] = {0, 0, 0, 0, 0, 0};
  bool loaded = false;
  if (use_local && sc.local_indices && sc.local_vertices) {
    const uint32_t local_base = inst.local_first_index + tri * 3u;
    const uint32_t local_end = inst.local_first_vertex + inst.local_vertex_count;
    if (static_cast<uint64_t>(local_base) + 2u < sc.local_index_count) {
      const uint32_t f = inst.local_first_vertex;
      const uint32_t l0 = f + sc.local_indices[local_base], l1 = f + sc.local_indices[local_base + 1u],
                     l2 = f + sc.local_indices[local_base + 2u];
      if (l0 < local_end && l1 < local_end && l2 < local_end && l0 < sc.local_vertex_count &&
          l1 < sc.local_vertex_count && l2 < sc.local_vertex_count) {
        a = transform_vertex(load3(sc.local_vertices + 3u * l0), inst);
        b = transform_vertex(load3(sc.local_vertices + 3u * l1), inst);
        c = transform_vertex(load3(sc.local_vertices + 3u * l2), inst);
        loaded = true;
      }
    }
  }
  if (!loaded) {
    const uint64_t base = (static_cast<uint64_t>(inst.first_triangle) + tri) * 3u;
    if (base + 2u >= sc.index_count) return false;
    const uint32_t i0 = sc.indices[base], i1 = sc.indices[base + 1], i2 = sc.indices[base + 2];
    if (i0 >= sc.vertex_count || i1 >= sc.vertex_count || i2 >= sc.vertex_count) return false;
    a = load3(sc.vertices + 3u * i0);
    b = load3(sc.vertices + 3u * i1);
    c = load3(sc.vertices + 3u * i2);
    if (texcoords) {
      const uint32_t vi[3] = {i0, i1, i2};
      for (int k = 0; k < 3; ++k) {
        uv[2 * k] = texcoords[2u * vi[k]];
        uv[2 * k + 1] = texcoords[2u * vi[k] + 1u];
      }
    }
  }
  p.material = inst.material_index;
  const V3 e1 = sub(b, a), e2 = sub(c, a);
  const V3 n = normalize_accel(cross(e1, e2));
  p.v0[0] = a.x; p.v0[1] = a.y; p.v0[2] = a.z;
  p.e1[0] = e1.x; p.e1[1] = e1.y; p.e1[2] = e1.z;
  p.e2[0] = e2.x; p.e2[1] = e2.y; p.e2[2] = e2.z;
  p.n[0] = n.x; p.n[1] = n.y; p.n[2] = n.z;
  std::memcpy(p.uv, uv, sizeof(uv));
  const float v9[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
  std::memcpy(r9, v9, sizeof(v9));
  return true;
}

// Every (instance, triangle) slot is made in parallel into its own place; slots the reference skips
// are then squeezed out in order, so the primitives are numbered as the sequential loop numbers them.
bool flatten(const PtSceneView& sc, HostVec<Prim>& prims, HostVec<float>& raw, std::string& err,
             const float* texcoords, std::vector<uint32_t>* inst_first, unsigned threads, HostVec<uint32_t>* materials,
             bool* bounded) {
  std::vector<uint64_t> first(static_cast<size_t>(sc.instance_count) + 1u, 0u);
  for (uint32_t i = 0; i < sc.instance_count; ++i) first[i + 1u] = first[i] + sc.instances[i].triangle_count;
  const uint64_t cap = first.back();
  if (cap >= kMaxPrimitives) {
    err = "scene has more triangles than the device BVH encoding supports";
    return false;
  }
  prims.resize(cap);
  raw.resize(9u * cap);
  if (materials) materials->resize(cap);
  HostVec<uint8_t> valid(cap);
  std::atomic<bool> unbounded{false};
  parallel_for(cap, threads, [&](size_t lo, size_t hi) {
    // the instance holding slot lo: the last one starting at or before it (empty instances skipped)
    size_t ii = static_cast<size_t>(std::upper_bound(first.begin(), first.end(), lo) - first.begin()) - 1u;
    bool big = false;
    for (size_t g = lo; g < hi; ++g) {
      while (g >= first[ii + 1u]) ++ii;
      const bool ok = make_prim(sc, sc.instances[ii], static_cast<uint32_t>(g - first[ii]), texcoords, prims[g],
                                &raw[9u * g]);
      valid[g] = ok ? 1u : 0u;
      if (!ok) continue;
      if (materials) (*materials)[g] = prims[g].material;
      for (int k = 0; k < 9; ++k) big = big || !(std::fabs(raw[9u * g + k]) <= 1e30f);
    }
    if (big) unbounded.store(true, std::memory_order_relaxed);
  });
  if (bounded) *bounded = !unbounded.load();
  size_t w = 0;
  if (inst_first) inst_first->clear();
  for (uint32_t ii = 0; ii < sc.instance_count; ++ii) {
    if (inst_first) inst_first->push_back(static_cast<uint32_t>(w));
    for (uint64_t g = first[ii]; g < first[ii + 1u]; ++g) {
      if (!valid[g]) continue;
      if (w != g) {
        prims[w] = prims[g];
        std::memcpy(&raw[9u * w], &raw[9u * g], 9 * sizeof(float));
        if (materials) (*materials)[w] = (*materials)[g];
      }
      ++w;
    }
  }
  prims.resize(w);
  raw.resize(9u * w);
  if (materials) materials->resize(w);
  if (inst_first) inst_first->push_back(static_cast<uint32_t>(w));
  return true;
}

void build_sah(Bvh& bvh, const HostVec<float>& raw, uint32_t leaf_size, uint32_t buckets) {
  const uint32_t n = static_cast<uint32_t>(bvh.prims.size());
  bvh.nodes.clear();
  bvh.order.clear();
  bvh.max_depth = 0;
  bvh.leaf_count = 0;
  if (n == 0) return;
  std::vector<Ref> refs(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* t = &raw[9u * i];
    Ref& r = refs[i];
    r.orig = i;
    for (int k = 0; k < 3; ++k) {
      r.mn[k] = min3(t[k], t[3 + k], t[6 + k]);
      r.mx[k] = max3(t[k], t[3 + k], t[6 + k]);
      r.c[k] = (r.mn[k] + r.mx[k]) * 0.5f;
    }
  }
  bvh.nodes.reserve(2u * n - 1u);
  bvh.order.reserve(n);
  SahBuilder b{refs, std::vector<Ref>(n), bvh, leaf_size ? leaf_size : 4u, buckets ? buckets : 8u};
  b.build(0, n, 0);
}

void build_midpoint(Bvh& bvh, const HostVec<float>& raw, uint32_t leaf_size) {
  const uint32_t n = static_cast<uint32_t>(bvh.prims.size());
  bvh.nodes.clear();
  bvh.order.clear();
  bvh.max_depth = 0;
  bvh.leaf_count = 0;
  if (n == 0) return;
  std::vector<float> boxes(6u * n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* t = &raw[9u * i];
    for (int k = 0; k < 3; ++k) {
      boxes[6u * i + k] = min3(t[k], t[3 + k], t[6 + k]) - kEps;
      boxes[6u * i + 3 + k] = max3(t[k], t[3 + k], t[6 + k]) + kEps;
    }
  }
  bvh.nodes.assign(2u * n - 1u, Node{});
  bvh.order.resize(n);
  for (uint32_t i = 0; i < n; ++i) bvh.order[i] = i;
  MidBuilder b{boxes, std::vector<uint32_t>(n), bvh, leaf_size ? leaf_size : 4u};
  b.build(0, bvh.order.data(), n, 0);
  bvh.nodes.resize(static_cast<size_t>(b.alloc));
}

void pack_device_layout(const Bvh& bvh, DeviceLayout& out, bool with_uv, unsigned threads) {
  out.nodes.clear();
  out.tris.clear();
  out.prims.clear();
  out.prim_uv.clear();
  out.prim_tri.clear();
  if (with_uv) {
    out.prim_uv.resize(6u * bvh.prims.size());
    parallel_for(bvh.prims.size(), threads, [&](size_t a, size_t b) {
      for (size_t i = a; i < b; ++i) std::memcpy(&out.prim_uv[6u * i], bvh.prims[i].uv, 6 * sizeof(float));
    });
    out.prim_tri.resize(bvh.prims.size());
    parallel_for(bvh.order.size(), threads, [&](size_t a, size_t b) {
      for (size_t i = a; i < b; ++i) out.prim_tri[bvh.order[i]] = static_cast<uint32_t>(i);
    });
  }
  out.inner_count = 0;
  const size_t nprim = bvh.prims.size();
  out.prims.resize(4u * nprim);
  parallel_for(nprim, threads, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const Prim& p = bvh.prims[i];
      out.prims[4 * i + 0] = p.n[0];
      out.prims[4 * i + 1] = p.n[1];
      out.prims[4 * i + 2] = p.n[2];
      std::memcpy(&out.prims[4 * i + 3], &p.material, 4);
    }
  });
  out.tris.resize(12u * bvh.order.size());
  parallel_for(bvh.order.size(), threads, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const uint32_t pid = bvh.order[i];
      const Prim& p = bvh.prims[pid];
      float* t = &out.tris[12 * i];
      t[0] = p.v0[0]; t[1] = p.v0[1]; t[2] = p.v0[2];
      std::memcpy(&t[3], &pid, 4);
      t[4] = p.e1[0]; t[5] = p.e1[1]; t[6] = p.e1[2]; t[7] = 0.0f;
      t[8] = p.e2[0]; t[9] = p.e2[1]; t[10] = p.e2[2]; t[11] = 0.0f;
    }
  });
  if (bvh.nodes.empty()) {
    out.root_ref = 0;
    return;
  }
  // Number inner nodes in their original (pre-order) order.
  std::vector<int32_t> remap(bvh.nodes.size(), -1);
  uint32_t inner = 0;
  for (size_t i = 0; i < bvh.nodes.size(); ++i)
    if (bvh.nodes[i].count == 0) remap[i] = static_cast<int32_t>(inner++);
  auto ref_of = [&](int32_t ni) -> int32_t {
    const Node& nd = bvh.nodes[static_cast<size_t>(ni)];
    if (nd.count > 0) {
      const uint32_t code = (static_cast<uint32_t>(nd.first) << kLeafCountBits) | static_cast<uint32_t>(nd.count);
      return ~static_cast<int32_t>(code);
    }
    return remap[static_cast<size_t>(ni)];
  };
  out.inner_count = inner;
  out.nodes.resize(16u * inner);
  parallel_for(bvh.nodes.size(), threads, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const Node& nd = bvh.nodes[i];
      if (nd.count > 0) continue;
      const Node& c0 = bvh.nodes[static_cast<size_t>(nd.left)];
      const Node& c1 = bvh.nodes[static_cast<size_t>(nd.right)];
      float* g = &out.nodes[16u * static_cast<size_t>(remap[i])];
      g[0] = c0.mn[0]; g[1] = c0.mn[1]; g[2] = c0.mn[2]; g[3] = c0.mx[0];
      g[4] = c0.mx[1]; g[5] = c0.mx[2]; g[6] = c1.mn[0]; g[7] = c1.mn[1];
      g[8] = c1.mn[2]; g[9] = c1.mx[0]; g[10] = c1.mx[1]; g[11] = c1.mx[2];
      const int32_t r0 = ref_of(nd.left), r1 = ref_of(nd.right), z = 0;
      std::memcpy(&g[12], &r0, 4);
      std::memcpy(&g[13], &r1, 4);
      std::memcpy(&g[14], &z, 4);
      std::memcpy(&g[15], &z, 4);
    }
  });
  const Node& root = bvh.nodes[0];
  for (int k = 0; k < 3; ++k) {
    out.root_box[k] = root.mn[k];
    out.root_box[3 + k] = root.mx[k];
  }
  out.root_ref = ref_of(0);
}

}  // namespace ptb
