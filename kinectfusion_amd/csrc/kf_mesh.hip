// Marching cubes on the device with the reference's exact output order
// (MarchingCubes::extract + MeshBuilder, src/kinectfusion/marching_cubes.cpp:
// 315-507; the reference runs it on the host only).
//
// The reference walks cells x-fastest, emits each case's triangles as
// (e0, e2, e1) and numbers a welded vertex the first time its edge is met.
// The first cell to meet an edge is the lowest-index *active* cell sharing
// it (a crossed edge is used by every valid cell around it), so:
//   1. classify   -- cube index per cell (0 when a corner is unobserved,
//                    non-finite or below min_weight);
//   2. counts     -- per 4096-cell chunk: triangles, and vertices the cell
//                    owns (edges for which it is the lowest active cell);
//   3. scan       -- exclusive scan of the chunk counts (one workgroup);
//   4. offsets    -- in-chunk scans give every active cell its first
//                    triangle and first owned vertex;
//   5. emit       -- owned vertices (position, central-difference normal,
//                    trilinear colour) and triangles, whose foreign vertices
//                    are resolved by recomputing the owner's emit order;
//   6. fill       -- vertices without a gradient normal get the normalised
//                    sum of their triangles' face normals, summed in triangle
//                    order (MeshBuilder::fill_missing_normals).
// Triangle order, vertex numbering, positions, normals and colours are
// therefore identical to the CPU mesher's.
#include <cstring>
#include <vector>

#include "kf_sampler.hpp"

namespace kf {
namespace {

#include "mc_cases.inc"

constexpr int kChunk = 4096;  // cells per chunk (256 threads x 16)

__constant__ int8_t c_tri[256][16];
__constant__ uint8_t c_ntri[256];
// corner offsets (marching_cubes.cpp:315-324) and edges {near, far, axis}
// (:330-348)
__constant__ int8_t c_corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                      {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__constant__ int8_t c_edge[12][3] = {{0, 1, 0}, {1, 2, 1}, {3, 2, 0}, {0, 3, 1}, {4, 5, 0}, {5, 6, 1},
                                     {7, 6, 0}, {4, 7, 1}, {0, 4, 2}, {1, 5, 2}, {2, 6, 2}, {3, 7, 2}};

struct Cells {
  int nx, ny, nz;  // voxels
  int cx, cy, cz;  // cells per axis (voxels - 1)
  long long n;     // cx * cy * cz
  __device__ inline long long flat(int x, int y, int z) const {
    return (static_cast<long long>(z) * cy + y) * cx + x;
  }
  __device__ inline bool inside(int x, int y, int z) const {
    return x >= 0 && y >= 0 && z >= 0 && x < cx && y < cy && z < cz;
  }
};

// Traversal orders of MarchingCubes::extract (marching_cubes.cpp:478-507).
// A cell's rank is its position in the traversal; rank() returns -1 for a
// cell the traversal never visits.
struct DenseOrder {  // cells x-fastest
  Cells c;
  __host__ __device__ inline long long n() const { return c.n; }
  __device__ inline bool coords(long long r, int& x, int& y, int& z) const {
    if (r >= c.n) return false;
    x = static_cast<int>(r % c.cx);
    y = static_cast<int>((r / c.cx) % c.cy);
    z = static_cast<int>(r / (static_cast<long long>(c.cx) * c.cy));
    return true;
  }
  __device__ inline long long rank(int x, int y, int z) const { return c.flat(x, y, z); }
};

// Sparse: allocated blocks in flat order, each block's cells x-fastest,
// clipped to the cell grid.  keys: the allocated block keys, sorted.
struct BlockOrder {
  Cells c;
  const uint32_t* keys;
  int nkeys;
  int bx, by;  // blocks per axis over the voxel grid
  __host__ __device__ inline long long n() const { return static_cast<long long>(nkeys) * kBlockVoxels; }
  __device__ inline bool coords(long long r, int& x, int& y, int& z) const {
    if (r >= n()) return false;
    const uint32_t key = keys[r >> 9];
    const int i = static_cast<int>(r & 511);
    const int kx = static_cast<int>(key % static_cast<uint32_t>(bx));
    const int ky = static_cast<int>((key / static_cast<uint32_t>(bx)) % static_cast<uint32_t>(by));
    const int kz = static_cast<int>(key / static_cast<uint32_t>(bx * by));
    x = kx * 8 + (i & 7);
    y = ky * 8 + ((i >> 3) & 7);
    z = kz * 8 + (i >> 6);
    return x < c.cx && y < c.cy && z < c.cz;
  }
  __device__ inline long long rank(int x, int y, int z) const {
    const uint32_t key = static_cast<uint32_t>(((z >> 3) * by + (y >> 3)) * bx + (x >> 3));
    int lo = 0, hi = nkeys - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const uint32_t k = keys[mid];
      if (k == key) return static_cast<long long>(mid) * kBlockVoxels + intra_of(x, y, z);
      if (k < key) lo = mid + 1;
      else hi = mid - 1;
    }
    return -1;
  }
};

__device__ inline bool active(uint8_t code) { return code != 0 && code != 255; }

// Local edge of a cell whose near corner sits at offset (ox, oy, oz) on axis a.
__device__ inline int local_edge(int ox, int oy, int oz, int axis) {
  for (int e = 0; e < 12; ++e) {
    const int nc = c_edge[e][0];
    if (c_edge[e][2] == axis && c_corner[nc][0] == ox && c_corner[nc][1] == oy && c_corner[nc][2] == oz) return e;
  }
  return -1;
}

// The (up to) four cells sharing the edge from voxel (vx, vy, vz) along
// `axis` that the traversal visits, as ranks in increasing order.
template <typename Order>
__device__ inline int edge_cells(const Order& o, int vx, int vy, int vz, int axis, long long (&r)[4]) {
  int n = 0;
  for (int k = 1; k >= 0; --k)
    for (int j = 1; j >= 0; --j) {
      int x = vx, y = vy, z = vz;
      if (axis == 0) {
        y -= j;
        z -= k;
      } else if (axis == 1) {
        x -= j;
        z -= k;
      } else {
        x -= j;
        y -= k;
      }
      if (!o.c.inside(x, y, z)) continue;
      const long long ci = o.rank(x, y, z);
      if (ci < 0) continue;
      int p = n++;
      while (p > 0 && r[p - 1] > ci) {
        r[p] = r[p - 1];
        --p;
      }
      r[p] = ci;
    }
  return n;
}

// The first active cell in traversal order sharing the edge; -1 when none
// (cannot happen for an edge some active cell uses).
template <typename Order>
__device__ inline long long owner_of(const Order& o, const uint8_t* __restrict__ code, int vx, int vy, int vz,
                                     int axis) {
  long long r[4];
  const int n = edge_cells(o, vx, vy, vz, axis, r);
  for (int i = 0; i < n; ++i)
    if (active(code[r[i]])) return r[i];
  return -1;
}

template <typename Order>
__device__ inline bool owns(const Order& o, const uint8_t* __restrict__ code, long long ci, int x, int y, int z,
                            int e) {
  const int nc = c_edge[e][0];
  return owner_of(o, code, x + c_corner[nc][0], y + c_corner[nc][1], z + c_corner[nc][2], c_edge[e][2]) == ci;
}

// Emit-order edge of triangle slot s (0, 1, 2) -> table entries (0, 2, 1).
__device__ inline int emit_edge(uint8_t cube, int t, int s) {
  return c_tri[cube][3 * t + (s == 0 ? 0 : (s == 1 ? 2 : 1))];
}

// Number of vertices cell ci owns; if `target` >= 0 returns instead the rank
// of local edge `target` among them.
template <typename Order>
__device__ int owned_rank(const Order& o, const uint8_t* __restrict__ code, long long ci, int x, int y, int z,
                          uint8_t cube, int target) {
  unsigned seen = 0;
  int r = 0;
  const int nt = c_ntri[cube];
  for (int t = 0; t < nt; ++t)
    for (int s = 0; s < 3; ++s) {
      const int e = emit_edge(cube, t, s);
      if (seen & (1U << e)) continue;
      seen |= 1U << e;
      if (owns(o, code, ci, x, y, z, e)) {
        if (e == target) return r;
        ++r;
      }
    }
  return target >= 0 ? -1 : r;
}

template <int Kind, typename Store, typename Order>
__global__ void __launch_bounds__(256) classify_kernel(Sampler<Kind, Store> vol, Order o, float min_weight,
                                                        uint8_t* __restrict__ code) {
  using IO = VoxelIO<Kind>;
  const long long n = o.n();
  for (long long ci = blockIdx.x * 256LL + threadIdx.x; ci < n; ci += static_cast<long long>(gridDim.x) * 256) {
    int x, y, z;
    if (!o.coords(ci, x, y, z)) {
      code[ci] = 0;
      continue;
    }
    unsigned cube = 0;
    bool valid = true;
    SlotCache cache;
    for (int k = 0; k < 8 && valid; ++k) {
      float t, w;
      bool obs, fin;
      IO::unpack(vol.s.voxel_cached(vol.g, x + c_corner[k][0], y + c_corner[k][1], z + c_corner[k][2], cache), t, w,
                 obs, fin);
      if (!obs || !fin || w < min_weight) valid = false;
      else if (t < 0.0F) cube |= 1U << k;
    }
    code[ci] = valid ? static_cast<uint8_t>(cube) : 0;
  }
}

template <typename Order>
__device__ inline void cell_counts(const Order& o, const uint8_t* __restrict__ code, long long ci, unsigned& nt,
                                   unsigned& nv) {
  nt = 0;
  nv = 0;
  if (ci >= o.n()) return;
  const uint8_t cube = code[ci];
  if (!active(cube)) return;
  int x, y, z;
  if (!o.coords(ci, x, y, z)) return;
  nt = c_ntri[cube];
  nv = static_cast<unsigned>(owned_rank(o, code, ci, x, y, z, cube, -1));
}

// Exclusive scan of two counters across the 256 threads of a workgroup.
__device__ inline void block_scan2(unsigned a, unsigned b, unsigned& ea, unsigned& eb, unsigned& ta, unsigned& tb) {
  __shared__ unsigned sa[4], sb[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned ia = a, ib = b;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned va = __shfl_up(ia, off), vb = __shfl_up(ib, off);
    if (lane >= off) {
      ia += va;
      ib += vb;
    }
  }
  if (lane == 63) {
    sa[wave] = ia;
    sb[wave] = ib;
  }
  __syncthreads();
  unsigned pa = 0, pb = 0;
  for (int w = 0; w < wave; ++w) {
    pa += sa[w];
    pb += sb[w];
  }
  ea = pa + ia - a;
  eb = pb + ib - b;
  ta = sa[0] + sa[1] + sa[2] + sa[3];
  tb = sb[0] + sb[1] + sb[2] + sb[3];
  __syncthreads();
}

template <typename Order>
__global__ void __launch_bounds__(256) chunk_count_kernel(Order o, const uint8_t* __restrict__ code,
                                                           unsigned* __restrict__ chunk_t,
                                                           unsigned* __restrict__ chunk_v) {
  unsigned st = 0, sv = 0;
  const long long base = static_cast<long long>(blockIdx.x) * kChunk;
  for (int j = 0; j < kChunk / 256; ++j) {
    unsigned nt, nv;
    cell_counts(o, code, base + j * 256 + threadIdx.x, nt, nv);
    st += nt;
    sv += nv;
  }
  unsigned ea, eb, ta, tb;
  block_scan2(st, sv, ea, eb, ta, tb);
  if (threadIdx.x == 0) {
    chunk_t[blockIdx.x] = ta;
    chunk_v[blockIdx.x] = tb;
  }
}

// Single workgroup of 256 threads: exclusive scan in place; totals[0..1].
__global__ void __launch_bounds__(256) chunk_scan_kernel(unsigned* __restrict__ chunk_t,
                                                          unsigned* __restrict__ chunk_v, long long n,
                                                          unsigned long long* __restrict__ totals) {
  const long long per = (n + 255) / 256;
  const long long lo = threadIdx.x * per, hi = lo + per < n ? lo + per : n;
  unsigned st = 0, sv = 0;
  for (long long i = lo; i < hi; ++i) {
    st += chunk_t[i];
    sv += chunk_v[i];
  }
  unsigned ea, eb, ta, tb;
  block_scan2(st, sv, ea, eb, ta, tb);
  for (long long i = lo; i < hi; ++i) {
    const unsigned t = chunk_t[i], v = chunk_v[i];
    chunk_t[i] = ea;
    chunk_v[i] = eb;
    ea += t;
    eb += v;
  }
  if (threadIdx.x == 0) {
    totals[0] = ta;
    totals[1] = tb;
  }
}

template <typename Order>
__global__ void __launch_bounds__(256) cell_offset_kernel(Order o, const uint8_t* __restrict__ code,
                                                           const unsigned* __restrict__ chunk_t,
                                                           const unsigned* __restrict__ chunk_v,
                                                           unsigned* __restrict__ tri_off,
                                                           unsigned* __restrict__ vert_off) {
  unsigned rt = chunk_t[blockIdx.x], rv = chunk_v[blockIdx.x];
  const long long base = static_cast<long long>(blockIdx.x) * kChunk;
  for (int j = 0; j < kChunk / 256; ++j) {
    const long long ci = base + j * 256 + threadIdx.x;
    unsigned nt, nv;
    cell_counts(o, code, ci, nt, nv);
    unsigned ea, eb, ta, tb;
    block_scan2(nt, nv, ea, eb, ta, tb);
    if (nt != 0) {
      tri_off[ci] = rt + ea;
      vert_off[ci] = rv + eb;
    }
    rt += ta;
    rv += tb;
  }
}

template <int Kind, typename Store, typename Order>
__global__ void __launch_bounds__(256) emit_kernel(Sampler<Kind, Store> vol, Order o,
                                                    const uint8_t* __restrict__ code,
                                                    const unsigned* __restrict__ tri_off,
                                                    const unsigned* __restrict__ vert_off,
                                                    float* __restrict__ pos, float* __restrict__ nrm,
                                                    float* __restrict__ col, uint32_t* __restrict__ tris,
                                                    unsigned long long* __restrict__ vkey,
                                                    unsigned* __restrict__ fix_list,
                                                    unsigned* __restrict__ fix_count, unsigned nvert) {
  using IO = VoxelIO<Kind>;
  const long long n = o.n();
  for (long long ci = blockIdx.x * 256LL + threadIdx.x; ci < n; ci += static_cast<long long>(gridDim.x) * 256) {
    const uint8_t cube = code[ci];
    if (!active(cube)) continue;
    int x, y, z;
    if (!o.coords(ci, x, y, z)) continue;
    const int nt = c_ntri[cube];
    // owned vertices, numbered in order of first appearance
    unsigned seen = 0, rank = 0;
    unsigned ids[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < nt; ++t)
      for (int s = 0; s < 3; ++s) {
        const int e = emit_edge(cube, t, s);
        if (seen & (1U << e)) continue;
        seen |= 1U << e;
        const int ne = c_edge[e][0], fe = c_edge[e][1], axis = c_edge[e][2];
        const int vx = x + c_corner[ne][0], vy = y + c_corner[ne][1], vz = z + c_corner[ne][2];
        const long long own = owner_of(o, code, vx, vy, vz, axis);
        if (own == ci) {
          const unsigned vid = vert_off[ci] + rank++;
          if (vid >= nvert) continue;  // never taken for a consistent classification
          ids[e] = vid;
          const int fx = x + c_corner[fe][0], fy = y + c_corner[fe][1], fz = z + c_corner[fe][2];
          float nv, fv, w;
          bool obs, fin;
          SlotCache cache;
          IO::unpack(vol.s.voxel_cached(vol.g, vx, vy, vz, cache), nv, w, obs, fin);
          IO::unpack(vol.s.voxel_cached(vol.g, fx, fy, fz, cache), fv, w, obs, fin);
          const V3 np = vol.g.center(vx, vy, vz);
          const V3 fp = vol.g.center(fx, fy, fz);
          const float frac = -nv / (fv - nv);
          const V3 p = add3(np, scale3(sub3(fp, np), frac));
          V3 nn;
          if (!vol.normal(p, false, nn)) nn = mk3(0.0F, 0.0F, 0.0F);
          V3 cc;
          if (!vol.color(p, cc)) cc = mk3(255.0F, 255.0F, 255.0F);
          pos[3 * static_cast<size_t>(vid)] = p.x;
          pos[3 * static_cast<size_t>(vid) + 1] = p.y;
          pos[3 * static_cast<size_t>(vid) + 2] = p.z;
          nrm[3 * static_cast<size_t>(vid)] = nn.x;
          nrm[3 * static_cast<size_t>(vid) + 1] = nn.y;
          nrm[3 * static_cast<size_t>(vid) + 2] = nn.z;
          col[3 * static_cast<size_t>(vid)] = cc.x;
          col[3 * static_cast<size_t>(vid) + 1] = cc.y;
          col[3 * static_cast<size_t>(vid) + 2] = cc.z;
          vkey[vid] = static_cast<unsigned long long>(vol.g.flat(vx, vy, vz)) * 3ULL + axis;
          if (nn.x == 0.0F && nn.y == 0.0F && nn.z == 0.0F) fix_list[atomicAdd(fix_count, 1U)] = vid;
        } else {
          // foreign vertex: its rank in the owner's emit order
          int ox = 0, oy = 0, oz = 0;
          const bool known = own >= 0 && o.coords(own, ox, oy, oz);
          const int le = known ? local_edge(vx - ox, vy - oy, vz - oz, axis) : -1;
          const int r = le >= 0 ? owned_rank(o, code, own, ox, oy, oz, code[own], le) : -1;
          const unsigned id = r >= 0 ? vert_off[own] + static_cast<unsigned>(r) : 0U;
          ids[e] = (r >= 0 && id < nvert) ? id : 0U;  // never taken for a consistent classification
        }
      }
    const unsigned t0 = tri_off[ci];
    for (int t = 0; t < nt; ++t)
      for (int s = 0; s < 3; ++s) tris[3 * static_cast<size_t>(t0 + t) + s] = ids[emit_edge(cube, t, s)];
  }
}

template <typename Order>
__global__ void __launch_bounds__(256) fill_normals_kernel(Order o, const uint8_t* __restrict__ code,
                                                            const unsigned* __restrict__ tri_off,
                                                            const uint32_t* __restrict__ tris,
                                                            const float* __restrict__ pos,
                                                            const unsigned long long* __restrict__ vkey,
                                                            const unsigned* __restrict__ fix_list,
                                                            const unsigned* __restrict__ fix_count,
                                                            float* __restrict__ nrm, unsigned nvert) {
  const unsigned n = *fix_count;
  const Cells& c = o.c;
  for (unsigned i = blockIdx.x * 256U + threadIdx.x; i < n; i += gridDim.x * 256U) {
    const unsigned vid = fix_list[i];
    const unsigned long long key = vkey[vid];
    const int axis = static_cast<int>(key % 3ULL);
    const unsigned long long vflat = key / 3ULL;
    const int vx = static_cast<int>(vflat % c.nx), vy = static_cast<int>((vflat / c.nx) % c.ny);
    const int vz = static_cast<int>(vflat / (static_cast<unsigned long long>(c.nx) * c.ny));
    // the vertex's triangles in triangle (traversal) order
    long long r[4];
    const int nc = edge_cells(o, vx, vy, vz, axis, r);
    V3 sum = mk3(0.0F, 0.0F, 0.0F);
    for (int q = 0; q < nc; ++q) {
      const long long ci = r[q];
      const uint8_t cube = code[ci];
      if (!active(cube)) continue;
      const unsigned t0 = tri_off[ci];
      for (int t = 0; t < c_ntri[cube]; ++t) {
        const uint32_t* tr = tris + 3 * static_cast<size_t>(t0 + t);
        if (tr[0] != vid && tr[1] != vid && tr[2] != vid) continue;
        if (tr[0] >= nvert || tr[1] >= nvert || tr[2] >= nvert) continue;
        const V3 p0 = mk3(pos[3 * tr[0]], pos[3 * tr[0] + 1], pos[3 * tr[0] + 2]);
        const V3 p1 = mk3(pos[3 * tr[1]], pos[3 * tr[1] + 1], pos[3 * tr[1] + 2]);
        const V3 p2 = mk3(pos[3 * tr[2]], pos[3 * tr[2] + 1], pos[3 * tr[2] + 2]);
        sum = add3(sum, cross3(sub3(p1, p0), sub3(p2, p0)));
      }
    }
    const V3 nn = normalize3(sum);
    nrm[3 * static_cast<size_t>(vid)] = nn.x;
    nrm[3 * static_cast<size_t>(vid) + 1] = nn.y;
    nrm[3 * static_cast<size_t>(vid) + 2] = nn.z;
  }
}

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { KF_HIP(hipMalloc(&p, bytes < 1 ? 1 : bytes)); }
  ~DevBuf() {
    if (p != nullptr) (void)hipFree(p);
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

void upload_tables(hipStream_t stream) {
  int8_t tri[256][16];
  uint8_t ntri[256];
  for (int k = 0; k < 256; ++k) {
    const char* s = kTriCases[k];
    const int n = static_cast<int>(std::strlen(s));
    for (int i = 0; i < 16; ++i) tri[k][i] = -1;
    for (int i = 0; i < n; ++i) tri[k][i] = static_cast<int8_t>(s[i] <= '9' ? s[i] - '0' : s[i] - 'a' + 10);
    ntri[k] = static_cast<uint8_t>(n / 3);
  }
  KF_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tri), tri, sizeof(tri), 0, hipMemcpyHostToDevice, stream));
  KF_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_ntri), ntri, sizeof(ntri), 0, hipMemcpyHostToDevice, stream));
  KF_HIP(hipStreamSynchronize(stream));
}

__global__ void zero_kernel(unsigned* p, int n) {
  if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0U;
}

template <int Kind>
void extract_t(kf_context* ctx, float min_weight, kf_mesh* mesh);

}  // namespace
}  // namespace kf

struct kf_mesh {
  std::vector<float> positions, normals, colors;
  std::vector<uint32_t> triangles;
};

namespace kf {
namespace {

template <int Kind, typename Store, typename Order>
void run_mc(kf_context* ctx, const Sampler<Kind, Store>& vol, const Order& o, float min_weight, kf_mesh* mesh) {
  const long long ncell = o.n();
  if (ncell <= 0) return;
  upload_tables(ctx->stream);
  const long long nchunks = (ncell + kChunk - 1) / kChunk;
  DevBuf code(static_cast<size_t>(ncell)), tri_off(sizeof(unsigned) * ncell), vert_off(sizeof(unsigned) * ncell);
  DevBuf chunk_t(sizeof(unsigned) * nchunks), chunk_v(sizeof(unsigned) * nchunks), totals(16), fixc(16);
  const int grid = 4096;
  hipLaunchKernelGGL((classify_kernel<Kind, Store, Order>), dim3(grid), dim3(256), 0, ctx->stream, vol, o,
                     min_weight, code.as<uint8_t>());
  hipLaunchKernelGGL(chunk_count_kernel<Order>, dim3(static_cast<unsigned>(nchunks)), dim3(256), 0, ctx->stream, o,
                     code.as<uint8_t>(), chunk_t.as<unsigned>(), chunk_v.as<unsigned>());
  hipLaunchKernelGGL(chunk_scan_kernel, dim3(1), dim3(256), 0, ctx->stream, chunk_t.as<unsigned>(),
                     chunk_v.as<unsigned>(), nchunks, totals.as<unsigned long long>());
  KF_HIP(hipGetLastError());
  unsigned long long tot[2] = {0, 0};
  KF_HIP(hipMemcpyAsync(tot, totals.p, sizeof(tot), hipMemcpyDeviceToHost, ctx->stream));
  KF_HIP(hipStreamSynchronize(ctx->stream));
  const unsigned long long ntri = tot[0], nvert = tot[1];
  if (ntri >= (1ULL << 31) || nvert >= (1ULL << 31)) fail(KF_ERR_OUT_OF_MEMORY, "mesh exceeds 2^31 elements");
  if (ntri == 0) return;
  hipLaunchKernelGGL(cell_offset_kernel<Order>, dim3(static_cast<unsigned>(nchunks)), dim3(256), 0, ctx->stream, o,
                     code.as<uint8_t>(), chunk_t.as<unsigned>(), chunk_v.as<unsigned>(), tri_off.as<unsigned>(),
                     vert_off.as<unsigned>());
  DevBuf pos(12 * nvert), nrm(12 * nvert), col(12 * nvert), tris(12 * ntri), vkey(8 * nvert), fix(4 * nvert);
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, ctx->stream, fixc.as<unsigned>(), 4);
  hipLaunchKernelGGL((emit_kernel<Kind, Store, Order>), dim3(grid), dim3(256), 0, ctx->stream, vol, o,
                     code.as<uint8_t>(), tri_off.as<unsigned>(), vert_off.as<unsigned>(), pos.as<float>(),
                     nrm.as<float>(), col.as<float>(), tris.as<uint32_t>(), vkey.as<unsigned long long>(),
                     fix.as<unsigned>(), fixc.as<unsigned>(), static_cast<unsigned>(nvert));
  hipLaunchKernelGGL(fill_normals_kernel<Order>, dim3(256), dim3(256), 0, ctx->stream, o, code.as<uint8_t>(),
                     tri_off.as<unsigned>(), tris.as<uint32_t>(), pos.as<float>(), vkey.as<unsigned long long>(),
                     fix.as<unsigned>(), fixc.as<unsigned>(), nrm.as<float>(), static_cast<unsigned>(nvert));
  KF_HIP(hipGetLastError());
  mesh->positions.resize(3 * nvert);
  mesh->normals.resize(3 * nvert);
  mesh->colors.resize(3 * nvert);
  mesh->triangles.resize(3 * ntri);
  KF_HIP(hipMemcpyAsync(mesh->positions.data(), pos.p, 12 * nvert, hipMemcpyDeviceToHost, ctx->stream));
  KF_HIP(hipMemcpyAsync(mesh->normals.data(), nrm.p, 12 * nvert, hipMemcpyDeviceToHost, ctx->stream));
  KF_HIP(hipMemcpyAsync(mesh->colors.data(), col.p, 12 * nvert, hipMemcpyDeviceToHost, ctx->stream));
  KF_HIP(hipMemcpyAsync(mesh->triangles.data(), tris.p, 12 * ntri, hipMemcpyDeviceToHost, ctx->stream));
  KF_HIP(hipStreamSynchronize(ctx->stream));
}

template <int Kind>
void extract_t(kf_context* ctx, float min_weight, kf_mesh* mesh) {
  using Raw = typename VoxelIO<Kind>::Raw;
  const Grid& g = ctx->grid;
  Cells c{g.nx, g.ny, g.nz, g.nx - 1, g.ny - 1, g.nz - 1, 0};
  c.n = static_cast<long long>(c.cx) * c.cy * c.cz;
  if (c.n <= 0) return;
  if (!hashed(ctx)) {
    Sampler<Kind> vol{DenseStore<Kind>{static_cast<const Raw*>(ctx->voxels), ctx->colors}, g, 1.0F / g.voxel};
    run_mc(ctx, vol, DenseOrder{c}, min_weight, mesh);
    return;
  }
  // sparse traversal: the allocated blocks in flat order
  int64_t nkeys = 0;
  if (kf_hashed_blocks(ctx, nullptr, 0, &nkeys) != KF_OK) fail(KF_ERR_DEVICE, ctx->error);
  if (nkeys == 0) return;
  std::vector<uint32_t> keys(static_cast<size_t>(nkeys));
  if (kf_hashed_blocks(ctx, keys.data(), nkeys, &nkeys) != KF_OK) fail(KF_ERR_DEVICE, ctx->error);
  DevBuf dkeys(sizeof(uint32_t) * keys.size());
  KF_HIP(hipMemcpyAsync(dkeys.p, keys.data(), sizeof(uint32_t) * keys.size(), hipMemcpyHostToDevice, ctx->stream));
  const BlockMap bm = block_map(g);
  BlockOrder o{c, dkeys.as<uint32_t>(), static_cast<int>(nkeys), bm.bx, bm.by};
  Sampler<Kind, HashedStore<Kind>> vol{HashedStore<Kind>{hash_view(ctx), static_cast<const Raw*>(ctx->voxels), ctx->colors},
                                       g, 1.0F / g.voxel};
  run_mc(ctx, vol, o, min_weight, mesh);
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" {

kf_status kf_mesh_extract(kf_context* ctx, float min_weight, kf_mesh** out) {
  if (ctx == nullptr || out == nullptr) return KF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto* mesh = new (std::nothrow) kf_mesh();
  if (mesh == nullptr) return KF_ERR_OUT_OF_MEMORY;
  try {
    switch (ctx->cfg.voxel_store) {
      case KF_VOXEL_QUANTIZED:
        extract_t<KF_VOXEL_QUANTIZED>(ctx, min_weight, mesh);
        break;
      case KF_VOXEL_BF16:
        extract_t<KF_VOXEL_BF16>(ctx, min_weight, mesh);
        break;
      default:
        extract_t<KF_VOXEL_FLOAT>(ctx, min_weight, mesh);
        break;
    }
  } catch (const Failure& f) {
    ctx->error = f.message;
    delete mesh;
    return f.status;
  } catch (...) {
    ctx->error = "mesh extraction failed";
    delete mesh;
    return KF_ERR_DEVICE;
  }
  *out = mesh;
  return KF_OK;
}

kf_status kf_mesh_size(const kf_mesh* mesh, int64_t* vertices, int64_t* triangles) {
  if (mesh == nullptr) return KF_ERR_INVALID_ARGUMENT;
  if (vertices != nullptr) *vertices = static_cast<int64_t>(mesh->positions.size() / 3);
  if (triangles != nullptr) *triangles = static_cast<int64_t>(mesh->triangles.size() / 3);
  return KF_OK;
}

kf_status kf_mesh_copy(const kf_mesh* mesh, float* positions, float* normals, float* colors, uint32_t* triangles) {
  if (mesh == nullptr) return KF_ERR_INVALID_ARGUMENT;
  if (positions != nullptr) std::memcpy(positions, mesh->positions.data(), mesh->positions.size() * sizeof(float));
  if (normals != nullptr) std::memcpy(normals, mesh->normals.data(), mesh->normals.size() * sizeof(float));
  if (colors != nullptr) std::memcpy(colors, mesh->colors.data(), mesh->colors.size() * sizeof(float));
  if (triangles != nullptr)
    std::memcpy(triangles, mesh->triangles.data(), mesh->triangles.size() * sizeof(uint32_t));
  return KF_OK;
}

void kf_mesh_free(kf_mesh* mesh) { delete mesh; }

}  // extern "C"
