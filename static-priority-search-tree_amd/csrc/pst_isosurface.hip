// Isosurface extraction over active metacells: the consumer of the metacell
// IDs an isovalue search returns. The reference stops at those IDs and
// copies them back "for on-device marching cubes"
// (testers/pst-tester-functions.tu:230-252).
//
// Method: marching tetrahedra. Each voxel (cube of 8 grid vertices, corner k
// at offset (k & 1, k >> 1 & 1, k >> 2)) is cut into the 6 tetrahedra
// {0, e_a, e_a + e_b, 7} of the axis orders (a, b, c) (Freudenthal / Kuhn);
// neighbouring voxels cut their shared faces along the same diagonal, so the
// surface has no cracks. A tetrahedron with one vertex on the other side of
// the isovalue q from the rest gives one triangle, with two on each side a
// quad (two triangles). A vertex is "below" when its value is < q. A surface
// vertex on grid edge (u, v) is u + (q - f(u)) / (f(v) - f(u)) (v - u) with u
// the endpoint of lower linear vertex index, so every voxel sharing the edge
// computes the same bits. Triangles are wound so that their normal points
// from below to above q.
//
// Metacell (a, b, c) of size (mx, my, mz) holds the voxels with base vertex
// x in [a mx, min((a + 1) mx, px - 1)) (likewise in y and z); its ID is
// a + (b + c gy) gx, as formMetacellTags numbers them. Count pass: one wave
// per metacell, a lane per voxel, the wave's total per metacell. Write pass:
// the same walk; for each tetrahedron index the lanes' triangles are staged
// in LDS and stored as one contiguous run after the metacell's offset
// (order: 64-voxel chunk, tetrahedron, voxel). Positions are grid
// coordinates (x, y, z) as float, 9 floats per triangle.
#include <cstring>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kIsoWave = 64;
constexpr int kIsoWavesPerCu = 24;  // measured on 1.6M metacells: 24 beats 16 and 32

// interpolation precision: float for float volumes, double otherwise
template <typename T>
using Acc = typename std::conditional<std::is_same<T, float>::value, float, double>::type;

// The six tetrahedra of a voxel as corner indices (corner k: x = k & 1,
// y = k >> 1 & 1, z = k >> 2): 0 -> e_a -> e_a + e_b -> 7.
__constant__ uint8_t kTets[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7},
                                    {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};

struct Voxel {
  uint32_t x, y, z;
};

// Triangles of one tetrahedron: 0, 1 or 2.
__device__ inline int tet_tris(uint32_t below_mask) {
  const int nb = __popc(below_mask);
  return (nb == 0 || nb == 4) ? 0 : (nb == 2 ? 2 : 1);
}

template <typename T>
__device__ inline int voxel_count(const T f[8], T q) {
  int n = 0;
#pragma unroll
  for (int t = 0; t < 6; t++) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) m |= (f[kTets[t][k]] < q ? 1u : 0u) << k;
    n += tet_tris(m);
  }
  return n;
}

template <typename T>
__device__ inline void load_corners(const T* __restrict__ vol, uint32_t px, uint32_t py, const Voxel& v, T f[8]) {
  const size_t base = v.x + (static_cast<size_t>(v.y) + static_cast<size_t>(v.z) * py) * px;
#pragma unroll
  for (int k = 0; k < 8; k++)
    f[k] = vol[base + (k & 1) + (static_cast<size_t>((k >> 1) & 1) + static_cast<size_t>(k >> 2) * py) * px];
}

// Surface point on the edge between corners i and j of voxel v.
template <typename T>
__device__ inline float3 edge_point(const Voxel& v, const T f[8], int i, int j, T q) {
  // corner index order is linear vertex index order: interpolate from the lower one
  const int u = i < j ? i : j, w = i < j ? j : i;
  using A = Acc<T>;
  const A t = (static_cast<A>(q) - static_cast<A>(f[u])) / (static_cast<A>(f[w]) - static_cast<A>(f[u]));
  const float ux = static_cast<float>(v.x + (u & 1)), uy = static_cast<float>(v.y + ((u >> 1) & 1)),
              uz = static_cast<float>(v.z + (u >> 2));
  const float tf = static_cast<float>(t);
  return make_float3(ux + tf * static_cast<float>((w & 1) - (u & 1)),
                     uy + tf * static_cast<float>(((w >> 1) & 1) - ((u >> 1) & 1)),
                     uz + tf * static_cast<float>((w >> 2) - (u >> 2)));
}

__device__ inline float3 corner_pos(const Voxel& v, int k) {
  return make_float3(static_cast<float>(v.x + (k & 1)), static_cast<float>(v.y + ((k >> 1) & 1)),
                     static_cast<float>(v.z + (k >> 2)));
}

// Writes triangle (a, b, c), wound so that its normal points along `up`.
__device__ inline void put_tri(float* __restrict__ out, float3 a, float3 b, float3 c, float3 up) {
  const float3 e1 = make_float3(b.x - a.x, b.y - a.y, b.z - a.z), e2 = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
  const float3 n = make_float3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
  if (n.x * up.x + n.y * up.y + n.z * up.z < 0.0f) {
    const float3 s = b;
    b = c;
    c = s;
  }
  out[0] = a.x, out[1] = a.y, out[2] = a.z;
  out[3] = b.x, out[4] = b.y, out[5] = b.z;
  out[6] = c.x, out[7] = c.y, out[8] = c.z;
}

// Triangles of tetrahedron t of voxel v (0, 1 or 2), written at `out`.
template <typename T>
__device__ inline void tet_write(const Voxel& v, const T f[8], T q, int t, float* out) {
  int below[4], above[4], nb = 0, na = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int c = kTets[t][k];
    if (f[c] < q) below[nb++] = c;
    else above[na++] = c;
  }
  if (nb == 0 || na == 0) return;
  // a direction from below to above q across this tetrahedron
  const float3 pb = corner_pos(v, below[0]), pa = corner_pos(v, above[0]);
  const float3 up = make_float3(pa.x - pb.x, pa.y - pb.y, pa.z - pb.z);
  if (nb == 1 || na == 1) {
    const int s = nb == 1 ? below[0] : above[0];
    const int* o = nb == 1 ? above : below;
    put_tri(out, edge_point(v, f, s, o[0], q), edge_point(v, f, s, o[1], q), edge_point(v, f, s, o[2], q), up);
  } else {
    const float3 p0 = edge_point(v, f, below[0], above[0], q), p1 = edge_point(v, f, below[0], above[1], q);
    const float3 p2 = edge_point(v, f, below[1], above[1], q), p3 = edge_point(v, f, below[1], above[0], q);
    put_tri(out, p0, p1, p2, up);
    put_tri(out + 9, p0, p2, p3, up);
  }
}

template <typename T>
__device__ inline int tet_count(const T f[8], T q, int t) {
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) m |= (f[kTets[t][k]] < q ? 1u : 0u) << k;
  return tet_tris(m);
}

struct IsoGrid {
  uint32_t px, py, pz;  // vertices per axis
  uint32_t mx, my, mz;  // voxels per metacell edge
  uint32_t gx, gy, gz;  // metacells per axis
};

// Voxel j of metacell `id` (false when past the metacell's voxels).
__device__ inline bool metacell_voxel(const IsoGrid& g, uint64_t id, uint32_t j, Voxel& v, uint32_t& nvox) {
  const uint32_t a = static_cast<uint32_t>(id % g.gx);
  const uint64_t r = id / g.gx;
  const uint32_t b = static_cast<uint32_t>(r % g.gy), c = static_cast<uint32_t>(r / g.gy);
  const uint32_t x0 = a * g.mx, y0 = b * g.my, z0 = c * g.mz;
  const uint32_t nx = min(g.mx, g.px - 1 - x0), ny = min(g.my, g.py - 1 - y0), nz = min(g.mz, g.pz - 1 - z0);
  nvox = nx * ny * nz;
  if (j >= nvox) return false;
  v.x = x0 + j % nx;
  v.y = y0 + (j / nx) % ny;
  v.z = z0 + j / (nx * ny);
  return true;
}

template <typename T, typename I>
__global__ __launch_bounds__(kIsoWave) void k_iso_count(const T* __restrict__ vol, IsoGrid g,
                                                         const I* __restrict__ ids, uint64_t nids, T q,
                                                         unsigned long long* __restrict__ counts,
                                                         unsigned long long* __restrict__ err) {
  const uint32_t lane = threadIdx.x;
  const uint64_t ncells = static_cast<uint64_t>(g.gx) * g.gy * g.gz;
  for (uint64_t i = blockIdx.x; i < nids; i += gridDim.x) {
    const uint64_t id = static_cast<uint64_t>(ids[i]);
    unsigned long long n = 0;
    if (id >= ncells) {
      if (lane == 0) atomicOr(err, 1ull);
    } else {
      uint32_t nvox = 1;
      for (uint32_t j0 = 0; j0 < nvox; j0 += kIsoWave) {
        Voxel v;
        if (metacell_voxel(g, id, j0 + lane, v, nvox)) {
          T f[8];
          load_corners(vol, g.px, g.py, v, f);
          n += voxel_count(f, q);
        }
      }
    }
    for (int o = kIsoWave / 2; o > 0; o >>= 1) n += __shfl_xor(n, o, kIsoWave);
    if (lane == 0) counts[i] = n;
  }
}

// One wave per metacell; per chunk of 64 voxels and per tetrahedron index,
// the lanes' triangles (at most 128) are staged in LDS at their prefix
// positions and written out as one contiguous run of floats, so the stores
// are coalesced. Order within a metacell: voxel chunk, tetrahedron, voxel.
template <typename T, typename I>
__global__ __launch_bounds__(kIsoWave) void k_iso_write(const T* __restrict__ vol, IsoGrid g,
                                                         const I* __restrict__ ids, uint64_t nids, T q,
                                                         const unsigned long long* __restrict__ offsets,
                                                         unsigned long long capacity, float* __restrict__ out) {
  __shared__ float stage[2 * kIsoWave * 9];
  const uint32_t lane = threadIdx.x;
  const uint64_t ncells = static_cast<uint64_t>(g.gx) * g.gy * g.gz;
  for (uint64_t i = blockIdx.x; i < nids; i += gridDim.x) {
    const uint64_t id = static_cast<uint64_t>(ids[i]);
    if (id >= ncells) continue;
    unsigned long long base = offsets[i];  // triangles written so far, as an output index
    uint32_t nvox = 1;
    for (uint32_t j0 = 0; j0 < nvox; j0 += kIsoWave) {
      Voxel v;
      T f[8];
      const bool have = metacell_voxel(g, id, j0 + lane, v, nvox);
      if (have) load_corners(vol, g.px, g.py, v, f);
      for (int t = 0; t < 6; t++) {
        const uint32_t n = have ? static_cast<uint32_t>(tet_count(f, q, t)) : 0u;
        uint32_t incl = n;
        for (int o = 1; o < kIsoWave; o <<= 1) {
          const uint32_t u = __shfl_up(incl, o, kIsoWave);
          if (static_cast<int>(lane) >= o) incl += u;
        }
        const uint32_t total = __shfl(incl, kIsoWave - 1, kIsoWave);
        if (total == 0) continue;
        if (n) tet_write(v, f, q, t, stage + 9 * (incl - n));
        __syncthreads();
        for (uint32_t k = lane; k < 9 * total; k += kIsoWave) {
          const unsigned long long tri = base + k / 9;
          if (tri < capacity) out[9 * base + k] = stage[k];
        }
        __syncthreads();
        base += total;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Marching cubes: one surface patch per voxel from a 256-case table over the
// 8 corner signs (corner k below when f < q). The table is generated at
// compile time from one rule, so no case list is transcribed:
//  - on each face of the cube, walked counter-clockwise as seen from outside,
//    every maximal run of below corners is cut off by one segment, from the
//    edge where the walk leaves the run to the edge where it entered it; a
//    face with two diagonal below corners thus gets two segments, each
//    cutting off one corner (the below corners stay apart), and the choice
//    depends on the face's corner signs alone, so the two voxels that share
//    a face cut it the same way and the surface has no cracks;
//  - every crossing edge starts one segment and ends one (the two faces of
//    an edge walk it in opposite directions), so the segments form closed
//    loops; each loop becomes a triangle fan, wound so that the normal
//    points from below to above q.
// Edges are numbered by axis: x edges 0-3 (corner pairs 0-1, 2-3, 4-5,
// 6-7), y edges 4-7 (0-2, 1-3, 4-6, 5-7), z edges 8-11 (0-4, 1-5, 2-6, 3-7).
// Surface vertices interpolate from the edge endpoint of lower linear
// index, as for the tetrahedra, so shared edges give identical vertices.
// ---------------------------------------------------------------------------
constexpr int kMcMaxTri = 5;
struct McCase {
  uint8_t ntri;
  uint8_t e[3 * kMcMaxTri];  // edge triples of the triangles
};
struct McTable {
  McCase c[256];
  int maxtri;
};
constexpr uint8_t kMcEdgeLo[12] = {0, 2, 4, 6, 0, 1, 4, 5, 0, 1, 2, 3};
constexpr uint8_t kMcEdgeHi[12] = {1, 3, 5, 7, 2, 3, 6, 7, 4, 5, 6, 7};

constexpr int mc_edge_of(int a, int b) {
  for (int e = 0; e < 12; e++)
    if ((kMcEdgeLo[e] == a && kMcEdgeHi[e] == b) || (kMcEdgeLo[e] == b && kMcEdgeHi[e] == a)) return e;
  return -1;
}

constexpr McTable make_mc_table() {
  McTable T{};
  // faces: axis bit, side, and corner offsets u, v with u x v along the
  // outward normal (x = 1, y = 2, z = 4)
  const int ax[6] = {1, 1, 2, 2, 4, 4};
  const int side[6] = {0, 1, 0, 1, 0, 1};
  const int uu[6] = {4, 2, 1, 4, 2, 1};
  const int vv[6] = {2, 4, 4, 1, 1, 2};
  for (int cfg = 0; cfg < 256; cfg++) {
    int next[12] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    for (int f = 0; f < 6; f++) {
      const int b = side[f] ? ax[f] : 0;
      const int c[4] = {b, b | uu[f], b | uu[f] | vv[f], b | vv[f]};
      bool in[4] = {};
      for (int i = 0; i < 4; i++) in[i] = (cfg >> c[i]) & 1;
      for (int i = 0; i < 4; i++) {
        if (!in[i] || in[(i + 1) % 4]) continue;  // the walk leaves a run at corner i
        int j = i;
        while (in[(j + 3) % 4]) j = (j + 3) % 4;  // back to the run's first corner
        next[mc_edge_of(c[i], c[(i + 1) % 4])] = mc_edge_of(c[(j + 3) % 4], c[j]);
      }
    }
    bool seen[12] = {};
    int nt = 0;
    for (int e0 = 0; e0 < 12; e0++) {
      if (next[e0] < 0 || seen[e0]) continue;
      int cyc[12] = {};
      int len = 0;
      for (int e = e0; !seen[e]; e = next[e]) {
        seen[e] = true;
        cyc[len++] = e;
      }
      for (int k = 1; k + 1 < len; k++) {
        if (nt < kMcMaxTri) {
          T.c[cfg].e[3 * nt] = static_cast<uint8_t>(cyc[0]);
          T.c[cfg].e[3 * nt + 1] = static_cast<uint8_t>(cyc[k + 1]);
          T.c[cfg].e[3 * nt + 2] = static_cast<uint8_t>(cyc[k]);
        }
        nt++;
      }
    }
    T.c[cfg].ntri = static_cast<uint8_t>(nt);
    if (nt > T.maxtri) T.maxtri = nt;
  }
  return T;
}

constexpr McTable kMcHost = make_mc_table();
static_assert(kMcHost.maxtri == kMcMaxTri, "marching cubes table capacity");
__constant__ McTable kMc = make_mc_table();

template <typename T>
__device__ inline uint32_t mc_config(const T f[8], T q) {
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) m |= (f[k] < q ? 1u : 0u) << k;
  return m;
}

template <typename T, typename I>
__global__ __launch_bounds__(kIsoWave) void k_mc_count(const T* __restrict__ vol, IsoGrid g,
                                                        const I* __restrict__ ids, uint64_t nids, T q,
                                                        unsigned long long* __restrict__ counts,
                                                        unsigned long long* __restrict__ err) {
  const uint32_t lane = threadIdx.x;
  const uint64_t ncells = static_cast<uint64_t>(g.gx) * g.gy * g.gz;
  for (uint64_t i = blockIdx.x; i < nids; i += gridDim.x) {
    const uint64_t id = static_cast<uint64_t>(ids[i]);
    unsigned long long n = 0;
    if (id >= ncells) {
      if (lane == 0) atomicOr(err, 1ull);
    } else {
      uint32_t nvox = 1;
      for (uint32_t j0 = 0; j0 < nvox; j0 += kIsoWave) {
        Voxel v;
        if (metacell_voxel(g, id, j0 + lane, v, nvox)) {
          T f[8];
          load_corners(vol, g.px, g.py, v, f);
          n += kMc.c[mc_config(f, q)].ntri;
        }
      }
    }
    for (int o = kIsoWave / 2; o > 0; o >>= 1) n += __shfl_xor(n, o, kIsoWave);
    if (lane == 0) counts[i] = n;
  }
}

// One wave per metacell; per chunk of 64 voxels the lanes' triangles are
// staged in LDS at their prefix positions (voxel order) and written out as
// one contiguous run. Order within a metacell: voxel, then the voxel's
// table order.
template <typename T, typename I>
__global__ __launch_bounds__(kIsoWave) void k_mc_write(const T* __restrict__ vol, IsoGrid g,
                                                        const I* __restrict__ ids, uint64_t nids, T q,
                                                        const unsigned long long* __restrict__ offsets,
                                                        unsigned long long capacity, float* __restrict__ out) {
  __shared__ float stage[kIsoWave * kMcMaxTri * 9];
  const uint32_t lane = threadIdx.x;
  const uint64_t ncells = static_cast<uint64_t>(g.gx) * g.gy * g.gz;
  for (uint64_t i = blockIdx.x; i < nids; i += gridDim.x) {
    const uint64_t id = static_cast<uint64_t>(ids[i]);
    if (id >= ncells) continue;
    unsigned long long base = offsets[i];
    uint32_t nvox = 1;
    for (uint32_t j0 = 0; j0 < nvox; j0 += kIsoWave) {
      Voxel v;
      T f[8];
      const bool have = metacell_voxel(g, id, j0 + lane, v, nvox);
      uint32_t cfg = 0, n = 0;
      if (have) {
        load_corners(vol, g.px, g.py, v, f);
        cfg = mc_config(f, q);
        n = kMc.c[cfg].ntri;
      }
      uint32_t incl = n;
      for (int o = 1; o < kIsoWave; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, kIsoWave);
        if (static_cast<int>(lane) >= o) incl += u;
      }
      const uint32_t total = __shfl(incl, kIsoWave - 1, kIsoWave);
      if (total == 0) continue;
      float* my = stage + 9 * (incl - n);
      for (uint32_t t = 0; t < n; t++) {
        const uint8_t* e = kMc.c[cfg].e + 3 * t;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const float3 p = edge_point(v, f, kMcEdgeLo[e[k]], kMcEdgeHi[e[k]], q);
          my[9 * t + 3 * k] = p.x;
          my[9 * t + 3 * k + 1] = p.y;
          my[9 * t + 3 * k + 2] = p.z;
        }
      }
      __syncthreads();
      for (uint32_t k = lane; k < 9 * total; k += kIsoWave) {
        const unsigned long long tri = base + k / 9;
        if (tri < capacity) out[9 * base + k] = stage[k];
      }
      __syncthreads();
      base += total;
    }
  }
}

template <typename F>
Status with_t(int dt, F&& f) {
  switch (dt) {
    case F32: return f(float{});
    case F64: return f(double{});
    case I32: return f(int32_t{});
    case I64: return f(int64_t{});
  }
  return Status{1, "Error: unsupported data type code " + std::to_string(dt)};
}
template <typename F>
Status with_idx(int idt, F&& f) {
  switch (idt) {
    case F32: case I32: case U32: return f(uint32_t{});
    case F64: case I64: case U64: return f(uint64_t{});
  }
  return Status{1, "Error: metacell IDs must be of an ID type with 4 or 8 bytes"};
}

Status iso_grid(const uint32_t pt_dims[3], const uint32_t mc_dims[3], IsoGrid& g) {
  for (int d = 0; d < 3; d++)
    if (pt_dims[d] < 2 || mc_dims[d] == 0)
      return Status{1, "Error: need at least 2 vertices and metacells of >= 1 voxel per axis"};
  g.px = pt_dims[0], g.py = pt_dims[1], g.pz = pt_dims[2];
  g.mx = mc_dims[0], g.my = mc_dims[1], g.mz = mc_dims[2];
  g.gx = (g.px - 1 + g.mx - 1) / g.mx;
  g.gy = (g.py - 1 + g.my - 1) / g.my;
  g.gz = (g.pz - 1 + g.mz - 1) / g.mz;
  return Status{};
}

unsigned iso_grid_blocks(size_t nids) {
  int dev = 0, cus = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t cap = static_cast<size_t>(cus > 0 ? cus : 1) * kIsoWavesPerCu;
  return static_cast<unsigned>(std::max<size_t>(1, std::min(nids, cap)));
}

}  // namespace

Status mc_count(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                const uint32_t mc_dims[3], const void* ids_d, size_t nids, const void* isovalue,
                unsigned long long* counts_d, unsigned long long* err_d, hipStream_t stream) {
  IsoGrid g;
  Status st = iso_grid(pt_dims, mc_dims, g);
  if (!st.ok() || nids == 0) return st;
  return with_t(dtype, [&](auto t) {
    return with_idx(id_dtype, [&](auto i) {
      using T = decltype(t);
      using I = decltype(i);
      T q;
      std::memcpy(&q, isovalue, sizeof(T));
      k_mc_count<T, I><<<iso_grid_blocks(nids), kIsoWave, 0, stream>>>(
          static_cast<const T*>(vertices_d), g, static_cast<const I*>(ids_d), nids, q, counts_d, err_d);
      return hip_status(hipGetLastError(), "Error in launching the marching cubes count");
    });
  });
}

Status mc_write(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                const uint32_t mc_dims[3], const void* ids_d, size_t nids, const void* isovalue,
                const unsigned long long* offsets_d, size_t capacity, float* tri_d, hipStream_t stream) {
  IsoGrid g;
  Status st = iso_grid(pt_dims, mc_dims, g);
  if (!st.ok() || nids == 0) return st;
  return with_t(dtype, [&](auto t) {
    return with_idx(id_dtype, [&](auto i) {
      using T = decltype(t);
      using I = decltype(i);
      T q;
      std::memcpy(&q, isovalue, sizeof(T));
      k_mc_write<T, I><<<iso_grid_blocks(nids), kIsoWave, 0, stream>>>(
          static_cast<const T*>(vertices_d), g, static_cast<const I*>(ids_d), nids, q, offsets_d, capacity, tri_d);
      return hip_status(hipGetLastError(), "Error in launching the marching cubes write");
    });
  });
}

Status iso_count(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                 const uint32_t mc_dims[3], const void* ids_d, size_t nids, const void* isovalue,
                 unsigned long long* counts_d, unsigned long long* err_d, hipStream_t stream) {
  IsoGrid g;
  Status st = iso_grid(pt_dims, mc_dims, g);
  if (!st.ok() || nids == 0) return st;
  return with_t(dtype, [&](auto t) {
    return with_idx(id_dtype, [&](auto i) {
      using T = decltype(t);
      using I = decltype(i);
      T q;
      std::memcpy(&q, isovalue, sizeof(T));
      k_iso_count<T, I><<<iso_grid_blocks(nids), kIsoWave, 0, stream>>>(
          static_cast<const T*>(vertices_d), g, static_cast<const I*>(ids_d), nids, q, counts_d, err_d);
      return hip_status(hipGetLastError(), "Error in launching the isosurface count");
    });
  });
}

Status iso_write(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                 const uint32_t mc_dims[3], const void* ids_d, size_t nids, const void* isovalue,
                 const unsigned long long* offsets_d, size_t capacity, float* tri_d, hipStream_t stream) {
  IsoGrid g;
  Status st = iso_grid(pt_dims, mc_dims, g);
  if (!st.ok() || nids == 0) return st;
  return with_t(dtype, [&](auto t) {
    return with_idx(id_dtype, [&](auto i) {
      using T = decltype(t);
      using I = decltype(i);
      T q;
      std::memcpy(&q, isovalue, sizeof(T));
      k_iso_write<T, I><<<iso_grid_blocks(nids), kIsoWave, 0, stream>>>(
          static_cast<const T*>(vertices_d), g, static_cast<const I*>(ids_d), nids, q, offsets_d, capacity, tri_d);
      return hip_status(hipGetLastError(), "Error in launching the isosurface write");
    });
  });
}

}  // namespace pst
