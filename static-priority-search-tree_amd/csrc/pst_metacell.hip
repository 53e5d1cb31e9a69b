// Metacell formation for isosurface queries, and a seeded synthetic volume.
//
// formMetacellTags (helpers/isosurface-data-processing.h:66-137, .tu:7-127,
// getVoxelMinMax .tu:179-212) turns a vertex grid into one interval per
// metacell: metacell (a, b, c) of size (mx, my, mz) voxels covers the
// vertices x in [a*mx, min((a+1)*mx, px-1)] (and likewise in y, z); its
// interval is (min, max) over them and its ID is a + (b + c*gy)*gx
// (helpers/linearise-id.h:11-19), with g = ceil((p-1)/m) metacells per axis
// (calcNumMetacells, isosurface-data-processing.h:21-36).
//
// Here a workgroup owns a block of vertex columns (a thread per column), one
// z slab of metacells (mz + 1 vertex planes) and a run of metacell rows, and
// sweeps down y: for each vertex row it reduces the slab's planes in
// registers, folds the row into the current metacell row, and carries the
// boundary row (the last of one metacell row, the first of the next) instead
// of reading it again; at the end of each metacell row it reduces every
// metacell's mx + 1 columns from LDS. The volume is read about 1 + 1/mz
// times (the shared boundary planes of neighbouring slabs); the bound is HBM
// bandwidth. Metacells of 256 or more voxels along x use k_metacells, one
// workgroup per metacell row with an LDS column tile.
#include <cfloat>
#include <climits>

#include "pst_internal.hpp"

namespace pst {

namespace {

constexpr int kMcBlock = 256;
constexpr uint32_t kMcCols = 4096;  // vertex columns per LDS tile

template <typename T>
__device__ inline T vmin(T a, T b) {
  if constexpr (std::is_floating_point<T>::value) return a < b ? a : (b < a ? b : a);
  else return a < b ? a : b;
}
template <typename T>
__device__ inline T vmax(T a, T b) {
  if constexpr (std::is_floating_point<T>::value) return a > b ? a : (b > a ? b : a);
  else return a > b ? a : b;
}

template <typename T, typename I>
__global__ __launch_bounds__(kMcBlock) void k_metacells(const T* __restrict__ vol, uint32_t px, uint32_t py,
                                                        uint32_t pz, uint32_t mx, uint32_t my, uint32_t mz,
                                                        uint32_t gx, uint32_t gy, Point<T, I>* __restrict__ out) {
  __shared__ T cmin[kMcCols];
  __shared__ T cmax[kMcCols];
  const uint32_t b = blockIdx.x, c = blockIdx.y;
  const uint32_t y0 = b * my, y1 = min((b + 1) * my, py - 1);  // inclusive vertex ranges
  const uint32_t z0 = c * mz, z1 = min((c + 1) * mz, pz - 1);
  const uint32_t cells_per_tile = (kMcCols - 1) / mx;
  for (uint32_t a0 = 0; a0 < gx; a0 += cells_per_tile) {
    const uint32_t a1 = min(gx, a0 + cells_per_tile);
    const uint32_t x0 = a0 * mx, x1 = min(a1 * mx, px - 1);
    for (uint32_t x = x0 + threadIdx.x; x <= x1; x += kMcBlock) {
      T mn = vol[x + (static_cast<size_t>(y0) + static_cast<size_t>(z0) * py) * px];
      T mxv = mn;
      for (uint32_t z = z0; z <= z1; z++) {
        const T* plane = vol + static_cast<size_t>(z) * py * px;
        for (uint32_t y = y0; y <= y1; y++) {
          const T v = plane[static_cast<size_t>(y) * px + x];
          mn = vmin(mn, v);
          mxv = vmax(mxv, v);
        }
      }
      cmin[x - x0] = mn;
      cmax[x - x0] = mxv;
    }
    __syncthreads();
    for (uint32_t a = a0 + threadIdx.x; a < a1; a += kMcBlock) {
      const uint32_t xs = a * mx - x0, xe = min((a + 1) * mx, px - 1) - x0;
      T mn = cmin[xs], mxv = cmax[xs];
      for (uint32_t x = xs + 1; x <= xe; x++) {
        mn = vmin(mn, cmin[x]);
        mxv = vmax(mxv, cmax[x]);
      }
      const uint64_t id = a + (static_cast<uint64_t>(b) + static_cast<uint64_t>(c) * gy) * gx;
      Point<T, I> r;
      r.dim1_val = mn;
      r.dim2_val = mxv;
      if constexpr (has_id_v<I>) r.id = static_cast<I>(id);
      out[id] = r;
    }
    __syncthreads();
  }
}

constexpr int kSweepBlock = 256;
constexpr uint32_t kSweepRows = 16;  // metacell rows per workgroup
constexpr uint32_t kRows = 4;        // vertex rows loaded together
constexpr uint32_t kPlanes = 5;      // vertex planes loaded together

template <typename T, typename I>
__global__ __launch_bounds__(kSweepBlock) void k_metacells_sweep(const T* __restrict__ vol, uint32_t px, uint32_t py,
                                                                 uint32_t pz, uint32_t mx, uint32_t my, uint32_t mz,
                                                                 uint32_t gx, uint32_t gy,
                                                                 Point<T, I>* __restrict__ out) {
  __shared__ T cmin[kSweepBlock];
  __shared__ T cmax[kSweepBlock];
  const uint32_t cpb = (kSweepBlock - 1) / mx;  // metacells per block along x
  const uint32_t a0 = blockIdx.x * cpb;
  const uint32_t c = blockIdx.y;
  const uint32_t b0 = blockIdx.z * kSweepRows, b1 = min(gy, b0 + kSweepRows);
  const uint32_t tid = threadIdx.x;
  const uint32_t x = a0 * mx + tid;
  const uint32_t xlast = min(min(gx, a0 + cpb) * mx, px - 1);  // last vertex column of the block
  const bool col = x <= xlast;
  const uint32_t z0 = c * mz, z1 = min((c + 1) * mz, pz - 1);
  const size_t plane = static_cast<size_t>(py) * px;
  const uint32_t nz = z1 - z0;  // planes of the slab, less one
  const T* colp = vol + x + static_cast<size_t>(z0) * plane;
  // min and max over the slab's planes of vertex rows y..y+kRows-1 (rows
  // past `ylast` and planes past the slab repeat the last one: min and max
  // do not change), every load issued before the first use
  auto rows = [&](uint32_t y, uint32_t ylast, T lo[kRows], T hi[kRows]) {
    for (uint32_t zb = 0; zb <= nz; zb += kPlanes) {
      T v[kRows][kPlanes];
#pragma unroll
      for (uint32_t r = 0; r < kRows; r++)
#pragma unroll
        for (uint32_t k = 0; k < kPlanes; k++)
          v[r][k] = colp[static_cast<size_t>(min(y + r, ylast)) * px + static_cast<size_t>(min(zb + k, nz)) * plane];
#pragma unroll
      for (uint32_t r = 0; r < kRows; r++)
#pragma unroll
        for (uint32_t k = 0; k < kPlanes; k++) {
          lo[r] = (zb == 0 && k == 0) ? v[r][k] : vmin(lo[r], v[r][k]);
          hi[r] = (zb == 0 && k == 0) ? v[r][k] : vmax(hi[r], v[r][k]);
        }
    }
  };
  T clo{}, chi{};  // the carried boundary row
  bool carry = false;
  for (uint32_t b = b0; b < b1; b++) {
    const uint32_t y0 = b * my, y1 = min((b + 1) * my, py - 1);
    T mn{}, mxv{};
    if (col) {
      if (carry) {
        mn = clo;
        mxv = chi;
      } else {
        T lo[kRows], hi[kRows];
        rows(y0, y0, lo, hi);
        mn = lo[0];
        mxv = hi[0];
      }
      for (uint32_t y = y0 + 1; y <= y1; y += kRows) {
        T lo[kRows], hi[kRows];
        rows(y, y1, lo, hi);
#pragma unroll
        for (uint32_t r = 0; r < kRows; r++) {
          mn = vmin(mn, lo[r]);
          mxv = vmax(mxv, hi[r]);
        }
        clo = lo[kRows - 1];  // row min(y + kRows - 1, y1): y1 on the last pass
        chi = hi[kRows - 1];
      }
      cmin[tid] = mn;
      cmax[tid] = mxv;
    }
    carry = true;  // row y1 starts the next metacell row
    __syncthreads();
    const uint32_t a = a0 + tid;
    if (tid < cpb && a < gx) {
      const uint32_t xs = tid * mx, xe = min((a + 1) * mx, px - 1) - a0 * mx;
      T lo = cmin[xs], hi = cmax[xs];
      for (uint32_t k = xs + 1; k <= xe; k++) {
        lo = vmin(lo, cmin[k]);
        hi = vmax(hi, cmax[k]);
      }
      const uint64_t id = a + (static_cast<uint64_t>(b) + static_cast<uint64_t>(c) * gy) * gx;
      Point<T, I> r;
      r.dim1_val = lo;
      r.dim2_val = hi;
      if constexpr (has_id_v<I>) r.id = static_cast<I>(id);
      out[id] = r;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Synthetic volume: value(i, j, k) = sum_g amp_g exp(-|p - c_g|^2 / (2 s_g^2))
// + noise_std * N(0, 1), p = (i, j, k) / (dims - 1). The normal draw is a
// Box-Muller transform of two uniforms from SplitMix64 over (seed, vertex
// index), so any vertex's value depends only on the seed and its index.
// ---------------------------------------------------------------------------
constexpr int kMaxGauss = 256;

__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gauss_volume(float* __restrict__ out, uint32_t px, uint32_t py, uint32_t pz,
                                                      const float4* __restrict__ g, uint32_t ng, float noise_std,
                                                      uint64_t seed) {
  __shared__ float4 sg[kMaxGauss];  // (cx, cy, cz, -1 / (2 s^2))
  __shared__ float sa[kMaxGauss];   // amplitudes
  for (uint32_t t = threadIdx.x; t < ng; t += blockDim.x) {
    sg[t] = g[2 * t];
    sa[t] = g[2 * t + 1].x;
  }
  __syncthreads();
  const uint64_t nvox = static_cast<uint64_t>(px) * py * pz;
  const float sx = px > 1 ? 1.0f / (px - 1) : 0.0f, sy = py > 1 ? 1.0f / (py - 1) : 0.0f,
              sz = pz > 1 ? 1.0f / (pz - 1) : 0.0f;
  for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < nvox;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t i = static_cast<uint32_t>(v % px);
    const uint64_t r = v / px;
    const uint32_t j = static_cast<uint32_t>(r % py), k = static_cast<uint32_t>(r / py);
    const float x = i * sx, y = j * sy, z = k * sz;
    float acc = 0.0f;
    for (uint32_t t = 0; t < ng; t++) {
      const float4 q = sg[t];
      const float dx = x - q.x, dy = y - q.y, dz = z - q.z;
      acc += sa[t] * __expf((dx * dx + dy * dy + dz * dz) * q.w);
    }
    if (noise_std != 0.0f) {
      const uint64_t h1 = splitmix64(seed * 0x100000001B3ull + 2 * v);
      const uint64_t h2 = splitmix64(seed * 0x100000001B3ull + 2 * v + 1);
      const float u1 = ((h1 >> 40) + 1) * (1.0f / 16777217.0f);  // (0, 1]
      const float u2 = (h2 >> 40) * (1.0f / 16777216.0f);
      acc += noise_std * sqrtf(-2.0f * __logf(u1)) * __cosf(6.28318530718f * u2);
    }
    out[v] = acc;
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
Status with_id(int idt, F&& f) {
  switch (idt) {
    case NO_ID: return f(NoId{});
    case F32: return f(float{});
    case F64: return f(double{});
    case I32: return f(int32_t{});
    case I64: return f(int64_t{});
    case U32: return f(uint32_t{});
    case U64: return f(uint64_t{});
  }
  return Status{1, "Error: unsupported ID type code " + std::to_string(idt)};
}

}  // namespace

Status form_metacells(int dtype, int id_dtype, const void* vertices_d, const uint32_t pt_dims[3],
                      const uint32_t mc_dims[3], void* out_pts_d, hipStream_t stream) {
  uint32_t g[3];
  for (int d = 0; d < 3; d++) {
    if (pt_dims[d] < 2 || mc_dims[d] == 0) return Status{1, "Error: need at least 2 vertices and metacells of >= 1 voxel per axis"};
    g[d] = (pt_dims[d] - 1 + mc_dims[d] - 1) / mc_dims[d];
  }
  if (mc_dims[0] >= kMcCols) return Status{1, "Error: metacells hold fewer than 4096 voxels along x"};
  if (g[2] > 65535u) return Status{1, "Error: metacell grid too large along z"};
  return with_t(dtype, [&](auto t) {
    return with_id(id_dtype, [&](auto i) {
      using T = decltype(t);
      using I = decltype(i);
      if (mc_dims[0] < static_cast<uint32_t>(kSweepBlock)) {
        const uint32_t cpb = (kSweepBlock - 1) / mc_dims[0];
        const dim3 grid((g[0] + cpb - 1) / cpb, g[2], (g[1] + kSweepRows - 1) / kSweepRows);
        k_metacells_sweep<T, I><<<grid, kSweepBlock, 0, stream>>>(
            static_cast<const T*>(vertices_d), pt_dims[0], pt_dims[1], pt_dims[2], mc_dims[0], mc_dims[1],
            mc_dims[2], g[0], g[1], static_cast<Point<T, I>*>(out_pts_d));
      } else {
        k_metacells<T, I><<<dim3(g[1], g[2]), kMcBlock, 0, stream>>>(
            static_cast<const T*>(vertices_d), pt_dims[0], pt_dims[1], pt_dims[2], mc_dims[0], mc_dims[1],
            mc_dims[2], g[0], g[1], static_cast<Point<T, I>*>(out_pts_d));
      }
      return hip_status(hipGetLastError(), "Error in launching metacell formation");
    });
  });
}

Status gauss_volume(float* out_d, const uint32_t dims[3], const float* params_d, uint32_t ngauss, float noise_std,
                    uint64_t seed, hipStream_t stream) {
  if (ngauss > static_cast<uint32_t>(kMaxGauss)) return Status{1, "Error: at most 256 Gaussians"};
  const uint64_t n = static_cast<uint64_t>(dims[0]) * dims[1] * dims[2];
  const unsigned blocks = static_cast<unsigned>(std::min<uint64_t>((n + 255) / 256, 1u << 20));
  k_gauss_volume<<<blocks, 256, 0, stream>>>(out_d, dims[0], dims[1], dims[2],
                                             reinterpret_cast<const float4*>(params_d), ngauss, noise_std, seed);
  return hip_status(hipGetLastError(), "Error in launching volume synthesis");
}

}  // namespace pst
