// Projective point-to-plane ICP on the device.
//
// Reference: CorrespondenceSearch::match, IcpCorrespondence::from_match and
// IcpNormalEquations::accumulate (include/kinectfusion/icp_correspondence.hpp:
// 44-294); the whole-level device loop (src/kinectfusion/cuda/
// icp_correspondence.cu:216-343) with its host epilogue
// (src/kinectfusion/icp_optimizer.cpp:82-169).
//
// Two kernels per Gauss-Newton / LM trial:
//  * every thread matches its pixels (five at a time, their memory reads in
//    flight together) and keeps the 29 per-correspondence terms (21 J^T J,
//    6 J^T r, distance, r^2) in FP32 registers;
//  * a wave reduces them with an xor butterfly, the four waves of a
//    workgroup combine in a fixed order through LDS, and each workgroup
//    writes one partial row -- no float atomics, so the sums do not depend
//    on scheduling;
//  * a one-workgroup kernel adds the partial rows in a fixed order with
//    Kahan-compensated FP32 (a kernel boundary, not a device-scope fence in
//    every sweep workgroup, makes the rows visible: such a fence writes back
//    the L2 and costs more than the sweep itself);
//  * one thread of it converts the sums to FP64 and runs the trial: FP64
//    pivoted LDL^T solve, caps, Rodrigues increment and pose composition in
//    FP64, convergence and failure latches.
// Once a level's status latches (converged or failed) the remaining trial
// kernels return at once, so a converged level costs no further sweeps.  The
// stability spectrum (FP64 Jacobi) is checked once per level, by the level's
// last trial, as in the reference's device loop.  A pass that usually exits
// at once runs each level in one launch instead (icp_level_kernel).
#include <mutex>

#include "kf_internal.hpp"

// Partial rows of the allocation-free reduction (kf_icp_reduce_static): one
// row of 32 floats per sweep workgroup, at most 256 workgroups (32 KiB).
// Each device holds its own copy for the life of the process; host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (static_scratch_rows below), never through this name.
static __device__ __align__(256) float scratch[256 * 32];

namespace kf {
namespace {

constexpr int kSums = 29;
constexpr int kPartialStride = 32;
constexpr int kMaxBlocks = 256;  // partial rows staged in 32 KB of LDS

struct Maps {
  const float4* __restrict__ lv;
  const float4* __restrict__ ln;
  const float4* __restrict__ mv;
  const float4* __restrict__ mn;
  int lw, lh, mw, mh;
  Intr k;
};

// match() (icp_correspondence.hpp:44-110) in two halves so that a thread
// can keep the memory reads of several pixels in flight together.
// Front: read the live vertex and normal, transform, project into the model
// maps; returns false when rejected before the model read.
__device__ inline bool match_front(const Maps& m, const Rigid& cam, const Rigid& mod, V3 lv, V3 ln, V3& src,
                                   size_t& mi) {
  if (!finite3(lv) || !finite3(ln)) return false;
  src = xform(cam, lv);
  const V3 sim = xform(mod, src);
  if (!finite3(sim) || sim.z <= 0.0F) return false;
  float u, v;
  project(m.k, sim, u, v);
  int col, row;
  if (!round_index(u, m.mw, col) || !round_index(v, m.mh, row)) return false;
  mi = static_cast<size_t>(row) * m.mw + col;
  return true;
}

// Back: with the model vertex and normal read, the distance and normal
// tests and the point-to-plane row (J, residual).
__device__ inline bool match_back(const Rigid& cam, float max_dist, float min_dot, V3 ln, V3 src, V3 mv, V3 mn,
                                  float* J, float& res, float& dist) {
  if (!finite3(mv) || !finite3(mn)) return false;
  const V3 sn = normalize3(rot(cam, ln));
  dist = norm3(sub3(src, mv));
  if (dist > max_dist) return false;
  if (dot3(sn, mn) < min_dot) return false;
  const V3 r3 = cross3(src, mn);
  J[0] = r3.x;
  J[1] = r3.y;
  J[2] = r3.z;
  J[3] = mn.x;
  J[4] = mn.y;
  J[5] = mn.z;
  res = dot3(mn, sub3(mv, src));
  return true;
}

// Pixels a thread matches per pass (grid-stride order): their live reads,
// then their model reads, are issued together; the terms are added in pixel
// order.
constexpr int kPixBatch = 5;

// Per-workgroup sweep: sums of block `blk`'s pixels (of `nblk` sweeping
// blocks) into partial[blk].
__device__ inline void sweep_block(const Maps& m, const Rigid& cam, const Rigid& mod, float max_dist,
                                   float min_dot, float* __restrict__ partial, float (*lds)[kSums + 1],
                                   int blk, int nblk) {
  float s[kSums];
#pragma unroll
  for (int i = 0; i < kSums; ++i) s[i] = 0.0F;
  unsigned int cnt = 0;
  const int n = m.lw * m.lh;
  const int stride = nblk * static_cast<int>(blockDim.x);
  for (int base = blk * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x); base < n;
       base += kPixBatch * stride) {
    V3 lv[kPixBatch], ln[kPixBatch];
    bool ok[kPixBatch];
#pragma unroll
    for (int j = 0; j < kPixBatch; ++j) {
      const int p = base + j * stride;
      ok[j] = p < n;
      lv[j] = mk3(0.0F, 0.0F, 0.0F);
      ln[j] = lv[j];
      if (ok[j]) {
        lv[j] = ld3(m.lv[p]);
        ln[j] = ld3(m.ln[p]);
      }
    }
    V3 src[kPixBatch], mv[kPixBatch], mn[kPixBatch];
    size_t mi[kPixBatch];
#pragma unroll
    for (int j = 0; j < kPixBatch; ++j) {
      src[j] = mk3(0.0F, 0.0F, 0.0F);
      mi[j] = 0;
      ok[j] = ok[j] && match_front(m, cam, mod, lv[j], ln[j], src[j], mi[j]);
    }
#pragma unroll
    for (int j = 0; j < kPixBatch; ++j) {
      mv[j] = mk3(0.0F, 0.0F, 0.0F);
      mn[j] = mv[j];
      if (ok[j]) {
        mv[j] = ld3(m.mv[mi[j]]);
        mn[j] = ld3(m.mn[mi[j]]);
      }
    }
#pragma unroll
    for (int j = 0; j < kPixBatch; ++j) {
      float J[6], res, dist;
      if (!ok[j] || !match_back(cam, max_dist, min_dot, ln[j], src[j], mv[j], mn[j], J, res, dist)) continue;
      int e = 0;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = r; c < 6; ++c, ++e) s[e] += J[r] * J[c];
#pragma unroll
      for (int r = 0; r < 6; ++r) s[21 + r] += J[r] * res;
      s[27] += dist;
      s[28] += res * res;
      ++cnt;
    }
  }
  // wave butterfly
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int i = 0; i < kSums; ++i) s[i] += __shfl_xor(s[i], off);
    cnt += __shfl_xor(cnt, off);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < kSums; ++i) lds[wave][i] = s[i];
    lds[wave][kSums] = __uint_as_float(cnt);
  }
  __syncthreads();
  if (threadIdx.x < kSums) {
    const int i = threadIdx.x;
    partial[blk * kPartialStride + i] = (lds[0][i] + lds[1][i]) + (lds[2][i] + lds[3][i]);
  } else if (threadIdx.x == kSums) {
    const unsigned int c = __float_as_uint(lds[0][kSums]) + __float_as_uint(lds[1][kSums]) +
                           __float_as_uint(lds[2][kSums]) + __float_as_uint(lds[3][kSums]);
    partial[blk * kPartialStride + kSums] = __uint_as_float(c);
  }
}

// Kahan-compensated FP32 sum of the partial rows.  The rows are first
// staged in LDS with coalesced loads.  Column c (29 sums and the count) is
// split over 8 threads: part q adds rows q, q + 8, ... with Kahan, then
// thread c adds the 8 part sums (net of their compensations) in order
// q = 0..7.  The order is fixed, so the result does not depend on scheduling.
__device__ inline void combine_partials(const float* __restrict__ partial, int nblocks, float* out) {
  __shared__ __align__(16) float rows[kMaxBlocks * kPartialStride];
  __shared__ float part_sum[8][kPartialStride];
  __shared__ float part_comp[8][kPartialStride];
  __shared__ unsigned part_cnt[8];
  {
    // every row in flight at once (8 float4 per thread), not one round trip
    // per loop iteration
    constexpr int kQuads = kMaxBlocks * kPartialStride / 4 / 256;
    const int quads = nblocks * kPartialStride / 4;
    const float4* src = reinterpret_cast<const float4*>(partial);
    float4 v[kQuads];
#pragma unroll
    for (int j = 0; j < kQuads; ++j) {
      const int i = static_cast<int>(threadIdx.x) + j * 256;
      v[j] = i < quads ? src[i] : make_float4(0.0F, 0.0F, 0.0F, 0.0F);
    }
#pragma unroll
    for (int j = 0; j < kQuads; ++j) {
      const int i = static_cast<int>(threadIdx.x) + j * 256;
      if (i < quads) reinterpret_cast<float4*>(rows)[i] = v[j];
    }
  }
  __syncthreads();
  const int col = threadIdx.x & 31, q = threadIdx.x >> 5;  // 256 threads: 32 columns x 8 parts
  float sum = 0.0F, comp = 0.0F;
  unsigned cnt = 0;
  if (col < kSums) {
    for (int b = q; b < nblocks; b += 8) {
      const float y = rows[b * kPartialStride + col] - comp;
      const float t = sum + y;
      comp = (t - sum) - y;
      sum = t;
    }
  } else if (col == kSums) {
    for (int b = q; b < nblocks; b += 8) cnt += __float_as_uint(rows[b * kPartialStride + kSums]);
  }
  part_sum[q][col] = sum;
  part_comp[q][col] = comp;
  if (col == kSums) part_cnt[q] = cnt;
  __syncthreads();
  if (threadIdx.x < kSums) {
    float s2 = 0.0F, c2 = 0.0F;
    for (int k = 0; k < 8; ++k) {
      const float y = (part_sum[k][threadIdx.x] - part_comp[k][threadIdx.x]) - c2;
      const float t = s2 + y;
      c2 = (t - s2) - y;
      s2 = t;
    }
    out[threadIdx.x] = s2;
  } else if (threadIdx.x == kSums) {
    unsigned c = 0;
    for (int k = 0; k < 8; ++k) c += part_cnt[k];
    out[kSums] = __uint_as_float(c);
  }
}

// FP64 LDL^T with symmetric diagonal pivoting (largest remaining |d| first),
// the factorisation Eigen::LDLT uses in the reference host solve
// (icp_optimizer.cpp:391): a pivot with |d| <= DBL_MIN contributes a zero
// solution component, so a damped semidefinite system still solves; the
// solve fails only on a non-finite solution.
__device__ inline void swap_d(double& a, double& b) {
  const double t = a;
  a = b;
  b = t;
}

__device__ bool solve6(const double* eq, int damping_mode, double scale, double* x) {
  // Every index below is a compile-time constant after unrolling (the pivot
  // swap is predicated), so the factorisation lives in registers.
  double a[6][6];
  {
    int e = 0;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = r; c < 6; ++c, ++e) {
        a[r][c] = eq[e];
        a[c][r] = eq[e];
      }
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    if (damping_mode == 1) a[r][r] += scale;
    else if (damping_mode == 2) a[r][r] += scale * a[r][r];
  }
  int perm[6] = {0, 1, 2, 3, 4, 5};
  double d[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int piv = k;
    double best = fabs(a[k][k]);
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const double v = fabs(a[i][i]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      if (piv == i) {
#pragma unroll
        for (int c = 0; c < 6; ++c) swap_d(a[k][c], a[i][c]);
#pragma unroll
        for (int r = 0; r < 6; ++r) swap_d(a[r][k], a[r][i]);
        const int tp = perm[k];
        perm[k] = perm[i];
        perm[i] = tp;
      }
    }
    d[k] = a[k][k];
    const bool zero = fabs(d[k]) <= 2.2250738585072014e-308;
#pragma unroll
    for (int i = k + 1; i < 6; ++i) a[i][k] = zero ? 0.0 : a[i][k] / d[k];
#pragma unroll
    for (int i = k + 1; i < 6; ++i)
#pragma unroll
      for (int jj = k + 1; jj <= i; ++jj) {
        a[i][jj] -= a[i][k] * d[k] * a[jj][k];
        a[jj][i] = a[i][jj];
      }
  }
  double y[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = eq[21 + perm[i]];
#pragma unroll
    for (int q = 0; q < i; ++q) s -= a[i][q] * y[q];
    y[i] = s;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) y[i] = fabs(d[i]) <= 2.2250738585072014e-308 ? 0.0 : y[i] / d[i];
  double z[6];
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int q = i + 1; q < 6; ++q) s -= a[q][i] * z[q];
    z[i] = s;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int jj = 0; jj < 6; ++jj)
      if (perm[i] == jj) x[jj] = z[i];
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 6; ++i) ok = ok && isfinite(x[i]);
  return ok;
}

// pose <- exp([w]) pose in FP64 (Rodrigues; left-multiplied world-frame
// increment, icp_optimizer.cpp:286).
__device__ void compose_increment(const double* x, double* pose) {
  const double wx = x[0], wy = x[1], wz = x[2];
  const double angle = sqrt(wx * wx + wy * wy + wz * wz);
  double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (angle > 1.0e-12) {
    const double ux = wx / angle, uy = wy / angle, uz = wz / angle;
    double s, c;
    sincos(angle, &s, &c);
    const double oc = 1.0 - c;
    R[0] = c + ux * ux * oc;
    R[1] = ux * uy * oc - uz * s;
    R[2] = ux * uz * oc + uy * s;
    R[3] = uy * ux * oc + uz * s;
    R[4] = c + uy * uy * oc;
    R[5] = uy * uz * oc - ux * s;
    R[6] = uz * ux * oc - uy * s;
    R[7] = uz * uy * oc + ux * s;
    R[8] = c + uz * uz * oc;
  }
  double out[12];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c)
      out[r * 4 + c] = R[r * 3 + 0] * pose[0 * 4 + c] + R[r * 3 + 1] * pose[1 * 4 + c] +
                       R[r * 3 + 2] * pose[2 * 4 + c];
    out[r * 4 + 3] = R[r * 3 + 0] * pose[3] + R[r * 3 + 1] * pose[7] + R[r * 3 + 2] * pose[11] + x[3 + r];
  }
  for (int i = 0; i < 12; ++i) pose[i] = out[i];
}

__device__ void sym_eig6(const double* eq, double* ev);

// check_system_stability (icp_optimizer.cpp:319-347) on FP64 sums.
__device__ bool stable_from_eigenvalues(const double* ev, const DeviceIcpParams& p, double* min_eig, double* cond);

__device__ bool stable_system(const double* eq, long long count, const DeviceIcpParams& p,
                              double* min_eig, double* cond) {
  if (count < p.min_corr) return false;
  double ev[6];
  sym_eig6(eq, ev);
  return stable_from_eigenvalues(ev, p, min_eig, cond);
}

__device__ bool stable_from_eigenvalues(const double* ev, const DeviceIcpParams& p, double* min_eig, double* cond) {
  double mn = ev[0], mx = ev[0];
  for (int i = 0; i < 6; ++i) {
    if (!isfinite(ev[i])) return false;
    mn = fmin(mn, ev[i]);
    mx = fmax(mx, ev[i]);
  }
  if (min_eig != nullptr) *min_eig = mn;
  if (!(mn > 0.0 && mx > 0.0)) return false;
  if (cond != nullptr) *cond = mx / mn;
  return mn >= static_cast<double>(p.min_eig) && mx / mn <= static_cast<double>(p.max_cond);
}

// One trial on the combined sums (solve_update_kernel semantics).  A failing
// trial reports kUnconstrainedSystem when the undamped system is unstable,
// as the reference host loop vetoes before solving (icp_optimizer.cpp:263-270).
__device__ void run_trial(IcpState* st, const DeviceIcpParams& p, const float* sums32) {
  double sums[kSums];
  for (int i = 0; i < kSums; ++i) sums[i] = static_cast<double>(sums32[i]);
  const long long count = static_cast<long long>(__float_as_uint(sums32[kSums]));
  if (count < p.min_corr) {
    for (int i = 0; i < kSums; ++i) st->eq[i] = sums[i];
    st->eq_count = count;
    st->status = 2;
    return;
  }
  if (st->trials == 0) st->lambda = p.lambda0;
  const bool damping = p.damping_mode != 0;
  const bool adaptive = damping && p.adaptive != 0;
  bool accepted = true;
  if (adaptive) {
    const double cost = sums[28] / static_cast<double>(count);
    accepted = st->trials == 0 || cost <= st->best_cost;
    if (accepted) {
      st->best_cost = cost;
      for (int i = 0; i < kSums; ++i) st->best_eq[i] = sums[i];
      st->best_count = count;
      for (int i = 0; i < 16; ++i) st->best_pose[i] = st->pose[i];
      st->lambda = fmax(st->lambda * p.dec, static_cast<double>(p.lmin));
    } else {
      ++st->rejected;
      for (int i = 0; i < 16; ++i) st->pose[i] = st->best_pose[i];
      st->lambda = fmin(st->lambda * p.inc, static_cast<double>(p.lmax));
    }
    for (int i = 0; i < kSums; ++i) st->eq[i] = st->best_eq[i];
    st->eq_count = st->best_count;
  } else {
    for (int i = 0; i < kSums; ++i) st->eq[i] = sums[i];
    st->eq_count = count;
  }
  ++st->trials;
  const int attempts = damping ? 8 : 1;
  double x[6] = {0, 0, 0, 0, 0, 0};
  double scale = st->lambda, used = st->lambda, ut = 0.0, ang = 0.0;
  bool solved = false;
  for (int a = 0; a < attempts; ++a) {
    double xs[6];
    if (solve6(st->eq, p.damping_mode, scale, xs)) {
      solved = true;
      used = scale;
      for (int i = 0; i < 6; ++i) x[i] = xs[i];
      ang = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
      ut = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
      if (!damping || (ut <= p.max_t && ang <= p.max_r)) break;
    }
    scale = fmin(scale * p.inc, static_cast<double>(p.lmax));
  }
  if (!solved) {
    st->status = (!damping && !stable_system(st->eq, st->eq_count, p, nullptr, nullptr)) ? 5 : 3;
    return;
  }
  st->lambda = used;
  st->update_t = ut;
  st->update_r = ang;
  if (ut > p.max_t || ang > p.max_r) {
    st->status = (!damping && !stable_system(st->eq, st->eq_count, p, nullptr, nullptr)) ? 5 : 4;
    return;
  }
  compose_increment(x, st->pose);
  if (accepted && ut < p.min_t && ang < p.min_r) st->status = 1;
}

__device__ void finish_level_block(IcpState* st, const DeviceIcpParams& p, int level, float check_rms);

// One trial's correspondence sweep: one partial row per workgroup.
__global__ void __launch_bounds__(256) icp_trial_sweep_kernel(Maps m, DeviceIcpParams p, IcpState* st,
                                                               float* __restrict__ partial) {
  __shared__ float lds[4][kSums + 1];
  if (st->status != 0 || st->frame_failed) return;  // latched: nothing to do
  // the trial's pose, and the level's model pose inverse (kept in float by
  // the level's preparation)
  const Rigid cam = rigid_from_pose64(st->pose), mod = rigid_from12(st->model_inv);
  sweep_block(m, cam, mod, p.max_dist, p.min_dot, partial, lds, static_cast<int>(blockIdx.x),
                static_cast<int>(gridDim.x));
}

// Ticket protocol: returns true in the workgroup that finished last, after
// which the partial rows of every workgroup are visible (device-scope
// release before the ticket, acquire after it).
__device__ inline bool last_block(unsigned int* ticket) {
  __shared__ bool is_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) is_last = atomicAdd(ticket, 1U) == gridDim.x - 1U;
  __syncthreads();
  if (is_last) __threadfence();
  return is_last;
}

// A whole level in one launch, for a pass whose levels usually exit at once
// (the learned warm start's fallback, when the learned run was kept): the
// workgroups run the level's trials in step.  Each trial: every workgroup
// sweeps its pixels into its partial row; the last to finish (ticket
// protocol) reduces them, solves, and publishes the trial by advancing an
// epoch word with a device-scope release; the others wait for the epoch
// with a device-scope acquire.  The grid is at most kMaxBlocks workgroups,
// all resident at once (checked on the host).  A wait that outlasts
// timeout_ticks of the 100 MHz wall clock (far beyond any trial by default)
// gives up: the workgroup marks the frame failed (it is then neither
// integrated nor reported tracked), counts the event in st->sync_timeouts
// for the host to report, and exits, so every wave ends.  The ticket a
// timed-out launch leaves short is reset when the next level begins.
// st->level_sync: [0] ticket, [1] epoch (only ever advanced: a launch reads
// its base before any workgroup can advance it).
__global__ void __launch_bounds__(256) icp_level_kernel(Maps m, DeviceIcpParams p, IcpState* st,
                                                        float* __restrict__ partial, long long timeout_ticks,
                                                        int iterations, int level) {
  __shared__ float lds[4][kSums + 1];
  __shared__ float combined[kSums + 1];
  __shared__ float xf[24];
  __shared__ int go, timed_out;
  if (st->frame_failed) return;
  unsigned int* sync = st->level_sync;
  const unsigned int base = __hip_atomic_load(&sync[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  for (int it = 0; it < iterations; ++it) {
    if (threadIdx.x == 0) {
      go = st->status == 0 ? 1 : 0;  // latched: the level is done
      if (go != 0) {
        for (int i = 0; i < 12; ++i) {
          xf[i] = static_cast<float>(st->pose[i]);
          xf[12 + i] = st->model_inv[i];
        }
      }
    }
    __syncthreads();
    if (go == 0) break;
    const Rigid cam = rigid_from12(xf), mod = rigid_from12(xf + 12);
    sweep_block(m, cam, mod, p.max_dist, p.min_dot, partial, lds, static_cast<int>(blockIdx.x),
                static_cast<int>(gridDim.x));
    const unsigned int target = base + static_cast<unsigned int>(it) + 1U;
    if (last_block(&sync[0])) {
      combine_partials(partial, gridDim.x, combined);
      __syncthreads();
      if (threadIdx.x <= kSums) st->sums32[threadIdx.x] = combined[threadIdx.x];
      if (threadIdx.x == 0) {
        sync[0] = 0U;
        run_trial(st, p, combined);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&sync[1], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
    } else {
      if (threadIdx.x == 0) {
        timed_out = 0;
        const long long t0 = wall_clock64();
        while (static_cast<int>(__hip_atomic_load(&sync[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) - target) <
               0) {
          if (wall_clock64() - t0 > timeout_ticks) {
            timed_out = 1;
            st->frame_failed = 1;
            st->failed_level = level;
            atomicAdd(&st->sync_timeouts, 1);
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      __syncthreads();
      if (timed_out != 0) return;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
  }
  if (blockIdx.x == 0) finish_level_block(st, p, level, 0.0F);
}

// The trial's reduction (one workgroup: the partial rows into st->sums32)
// and its FP64 solve and pose update (one thread).  A kernel of its own, so
// the partial rows of every sweep workgroup are visible without a
// device-scope fence in each of them.
__global__ void __launch_bounds__(256) icp_trial_solve_kernel(DeviceIcpParams p, IcpState* st,
                                                               const float* __restrict__ partial, int nblocks,
                                                               int finish_level, float check_rms) {
  __shared__ float combined[kSums + 1];
  if (st->status == 0 && st->frame_failed == 0) {
    combine_partials(partial, nblocks, combined);
    __syncthreads();
    if (threadIdx.x <= kSums) st->sums32[threadIdx.x] = combined[threadIdx.x];
    if (threadIdx.x == 0) run_trial(st, p, combined);
  }
  if (finish_level >= 0) {
    __syncthreads();
    finish_level_block(st, p, finish_level, check_rms);
  }
}

// Single sweep with explicit float transforms (kf_icp_reduce), then its
// reduction into st->sums32.
__global__ void __launch_bounds__(256) icp_sweep_kernel(Maps m, Rigid cam, Rigid mod, float max_dist,
                                                         float min_dot, float* __restrict__ partial) {
  __shared__ float lds[4][kSums + 1];
  sweep_block(m, cam, mod, max_dist, min_dot, partial, lds, static_cast<int>(blockIdx.x),
              static_cast<int>(gridDim.x));
}

__global__ void __launch_bounds__(256) icp_combine_kernel(IcpState* st, const float* __restrict__ partial,
                                                           int nblocks) {
  __shared__ float combined[kSums + 1];
  combine_partials(partial, nblocks, combined);
  __syncthreads();
  if (threadIdx.x <= kSums) st->sums32[threadIdx.x] = combined[threadIdx.x];
}

// The same reduction into caller memory: the 29 sums widened to FP64 and the
// correspondence count (kf_icp_reduce_static).
__global__ void __launch_bounds__(256) icp_combine_out_kernel(const float* __restrict__ partial, int nblocks,
                                                               double* __restrict__ sums,
                                                               int64_t* __restrict__ count) {
  __shared__ float combined[kSums + 1];
  combine_partials(partial, nblocks, combined);
  __syncthreads();
  if (threadIdx.x < kSums) sums[threadIdx.x] = static_cast<double>(combined[threadIdx.x]);
  else if (threadIdx.x == kSums) *count = static_cast<int64_t>(__float_as_uint(combined[kSums]));
}

// FP64 cyclic Jacobi eigenvalues of the symmetric 6x6 system.
__device__ void sym_eig6(const double* eq, double* ev) {
  double a[6][6];
  int e = 0;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c, ++e) {
      a[r][c] = eq[e];
      a[c][r] = eq[e];
    }
  // Sweeps stop once the off-diagonal mass is below (1e-16)^2 of the
  // diagonal's: the eigenvalues are then exact to ~1e-16 of the norm.
  for (int sweep = 0; sweep < 32; ++sweep) {
    double off = 0.0, diag = 0.0;
#pragma unroll
    for (int d = 0; d < 6; ++d) diag += a[d][d] * a[d][d];
#pragma unroll
    for (int pp = 0; pp < 6; ++pp)
#pragma unroll
      for (int q = pp + 1; q < 6; ++q) off += a[pp][q] * a[pp][q];
    if (off <= 1.0e-32 * diag) break;
#pragma unroll
    for (int pp = 0; pp < 6; ++pp)
#pragma unroll
      for (int q = pp + 1; q < 6; ++q) {
        if (a[pp][q] == 0.0) continue;
        const double theta = (a[q][q] - a[pp][pp]) / (2.0 * a[pp][q]);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0);
        const double s = t * c;
#pragma unroll
        for (int k2 = 0; k2 < 6; ++k2) {
          const double akp = a[k2][pp], akq = a[k2][q];
          a[k2][pp] = c * akp - s * akq;
          a[k2][q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k2 = 0; k2 < 6; ++k2) {
          const double apk = a[pp][k2], aqk = a[q][k2];
          a[pp][k2] = c * apk - s * aqk;
          a[q][k2] = s * apk + c * aqk;
        }
      }
  }
  for (int i = 0; i < 6; ++i) ev[i] = a[i][i];
}

// The same eigenvalues computed by one wave (the stability check at the end
// of a level), in two steps:
//  * lane 0 reduces the system to tridiagonal form by four Householder
//    reflections H = I - beta v v^T, each applied as A - v w^T - w v^T with
//    p = beta A v, K = beta (v . p) / 2, w = p - K v (every index a
//    compile-time constant, so the matrix stays in registers);
//  * 60 lanes multisect the six eigenvalues: the 10 lanes of eigenvalue k
//    probe 10 interior points of its interval (initially the Gershgorin
//    interval of the tridiagonal matrix), each counting the eigenvalues below
//    its point by the sign changes of the Sturm sequence p_0 = 1,
//    p_1 = d_0 - x, p_i = (d_{i-1} - x) p_{i-1} - e_{i-2}^2 p_{i-2}
//    (division-free; magnitudes stay far inside the FP64 range for these
//    systems), and keep the sub-interval where the count passes k.  Sixteen
//    rounds shrink the interval by 11^16, to the rounding of the
//    tridiagonal matrix's norm.
// Backward stable: the eigenvalues agree with a full-precision solver to
// ~1e-15 of the norm (a CPU emulation of these steps checks it against
// numpy on condition numbers up to 1e8).  A non-finite system gives NaN.
constexpr int kSectRounds = 16;
constexpr int kSectProbes = 10;

__device__ inline int sturm_count_below(const double (&d)[6], const double (&e2)[5], double x) {
  double p_prev = 1.0, p = d[0] - x;
  bool neg = p < 0.0 || p == 0.0;  // a zero takes the sign opposite to its predecessor (+1)
  int changes = neg ? 1 : 0;
#pragma unroll
  for (int i = 1; i < 6; ++i) {
    const double p_next = (d[i] - x) * p - e2[i - 1] * p_prev;
    const bool n_neg = p_next == 0.0 ? !neg : p_next < 0.0;
    changes += n_neg != neg ? 1 : 0;
    neg = n_neg;
    p_prev = p;
    p = p_next;
  }
  return changes;
}

__device__ void sym_eig6_wave(const double* eq, double* ev) {
  __shared__ double tri[11];  // d[6], e[5]
  const int lane = threadIdx.x & 63;
  if (lane == 0) {
    double a[6][6];
    {
      int idx = 0;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = r; c < 6; ++c, ++idx) {
          a[r][c] = eq[idx];
          a[c][r] = eq[idx];
        }
    }
    double d[6], e[5];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double n2 = 0.0;
#pragma unroll
      for (int i = k + 1; i < 6; ++i) n2 += a[i][k] * a[i][k];
      double alpha = sqrt(n2);
      if (a[k + 1][k] > 0.0) alpha = -alpha;
      double v[6];
#pragma unroll
      for (int i = k + 1; i < 6; ++i) v[i] = a[i][k];
      v[k + 1] -= alpha;
      double vn = 0.0;
#pragma unroll
      for (int i = k + 1; i < 6; ++i) vn += v[i] * v[i];
      e[k] = alpha;
      d[k] = a[k][k];
      if (vn > 0.0) {
        const double beta = 2.0 / vn;
        double pv[6], w[6];
        double kk = 0.0;
#pragma unroll
        for (int i = k + 1; i < 6; ++i) {
          double acc = 0.0;
#pragma unroll
          for (int j = k + 1; j < 6; ++j) acc += a[i][j] * v[j];
          pv[i] = beta * acc;
          kk += v[i] * pv[i];
        }
        kk *= 0.5 * beta;
#pragma unroll
        for (int i = k + 1; i < 6; ++i) w[i] = pv[i] - kk * v[i];
#pragma unroll
        for (int i = k + 1; i < 6; ++i)
#pragma unroll
          for (int j = k + 1; j < 6; ++j) a[i][j] -= v[i] * w[j] + w[i] * v[j];
      }
    }
    d[4] = a[4][4];
    d[5] = a[5][5];
    e[4] = a[5][4];
#pragma unroll
    for (int i = 0; i < 6; ++i) tri[i] = d[i];
#pragma unroll
    for (int i = 0; i < 5; ++i) tri[6 + i] = e[i];
  }
  // lane 0's LDS writes precede every lane's reads in the wave's program order
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double d[6], e2[5];
  bool finite = true;
  double lo = 1.0e308, hi = -1.0e308;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    d[i] = tri[i];
    finite = finite && isfinite(d[i]);
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const double ei = tri[6 + i];
    finite = finite && isfinite(ei);
    e2[i] = ei * ei;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {  // Gershgorin discs of the tridiagonal matrix
    const double r = (i > 0 ? fabs(tri[5 + i]) : 0.0) + (i < 5 ? fabs(tri[6 + i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
  }
  const double pad = (hi - lo) * 1.0e-12 + 1.0e-300;
  lo -= pad;
  hi += pad;
  const int k = lane / kSectProbes, j = lane % kSectProbes + 1;  // eigenvalue, probe (lanes < 60)
  for (int round = 0; round < kSectRounds; ++round) {
    const double step = (hi - lo) * (1.0 / (kSectProbes + 1));
    const bool above = k < 6 && sturm_count_below(d, e2, lo + step * j) > k;
    const unsigned long long votes = __ballot(above);
    const unsigned group = k < 6 ? static_cast<unsigned>((votes >> (kSectProbes * k)) & 0x3FFULL) : 0U;
    const int first = group != 0U ? __builtin_ctz(group) + 1 : kSectProbes + 1;  // first probe above
    const double nlo = first > 1 ? lo + step * (first - 1) : lo;
    const double nhi = first <= kSectProbes ? lo + step * first : hi;
    lo = nlo;
    hi = nhi;
  }
  if (k < 6 && j == 1) ev[k] = finite ? 0.5 * (lo + hi) : __builtin_nan("");
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Prepare a level: status cleared, model pose = running pose.
__global__ void __launch_bounds__(64) icp_begin_level_kernel(IcpState* st, int level) {
  if (threadIdx.x == 0) icp_begin_level(st, level);
}

// Stability check and result record of a level (icp_optimizer.cpp:108-169),
// by a whole workgroup (the Jacobi sweeps take its first wave), then, when
// check_rms > 0, the learned warm start's check of this level.
__device__ void finish_level_block(IcpState* st, const DeviceIcpParams& p, int level, float check_rms) {
  __shared__ double ev[6];
  __shared__ int skip, check;
  if (threadIdx.x == 0) {
    skip = st->frame_failed;
    check = (st->status == 0 || st->status == 1) && st->eq_count >= p.min_corr ? 1 : 0;
  }
  __syncthreads();
  if (skip == 0) {
    if (check != 0 && threadIdx.x < 64) sym_eig6_wave(st->eq, ev);  // the first wave
    if (threadIdx.x == 0) {
      kf_icp_result& r = st->results[level];
      int status = st->status;
      r.min_eigenvalue = 0.0;
      r.condition_number = 0.0;
      if (status == 0 || status == 1) {
        const bool stable = check != 0 && stable_from_eigenvalues(ev, p, &r.min_eigenvalue, &r.condition_number);
        if (!stable && p.damping_mode == 0) status = 5;
      }
      r.status = status;
      r.trials = st->trials;
      r.rejected = st->rejected;
      r.correspondences = st->eq_count;
      if (st->eq_count > 0) {
        r.mean_distance = st->eq[27] / static_cast<double>(st->eq_count);
        r.mean_squared_residual = st->eq[28] / static_cast<double>(st->eq_count);
      }
      r.update_translation = st->update_t;
      r.update_rotation = st->update_r;
      r.lambda = st->lambda;
      st->status = status;
      if (status >= 2) {
        // failure: the outcome keeps the level's initial pose
        for (int i = 0; i < 16; ++i) st->pose[i] = st->level_pose0[i];
        st->frame_failed = 1;
        st->failed_level = level;
      }
    }
  }
  if (check_rms > 0.0F && threadIdx.x == 0) learned_check(st, level, check_rms, p.learned_max_t, p.learned_max_r);
}

// A level with no trial: its finish alone.
__global__ void __launch_bounds__(64) icp_finish_level_kernel(IcpState* st, DeviceIcpParams p, int level,
                                                               float check_rms) {
  finish_level_block(st, p, level, check_rms);
}

Maps make_maps(const float4* lv_v, const float4* lv_n, int lw, int lh, const float4* md_v,
               const float4* md_n, int mw, int mh, Intr k) {
  return Maps{lv_v, lv_n, md_v, md_n, lw, lh, mw, mh, k};
}

int blocks_for(int n) { return std::max(1, std::min(kMaxBlocks, (n + 255) / 256)); }

}  // namespace

DeviceIcpParams to_device_params(const kf_icp_options& o) {
  DeviceIcpParams p{};
  p.min_corr = o.min_correspondences;
  p.max_dist = o.max_point_distance;
  p.min_dot = o.min_normal_dot;
  p.min_t = o.min_update_translation;
  p.min_r = o.min_update_rotation;
  p.max_t = o.max_update_translation;
  p.max_r = o.max_update_rotation;
  p.min_eig = o.min_system_eigenvalue;
  p.max_cond = o.max_condition_number;
  p.damping_mode = o.damping_mode;
  p.lambda0 = o.damping_lambda;
  p.adaptive = o.adaptive_damping;
  p.inc = o.lambda_increase;
  p.dec = o.lambda_decrease;
  p.lmin = o.min_lambda;
  p.lmax = o.max_lambda;
  return p;
}

void launch_icp_begin_level(kf_context* ctx, int level) {
  hipLaunchKernelGGL(icp_begin_level_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->icp, level);
  KF_HIP(hipGetLastError());
}

void launch_icp_level(kf_context* ctx, const float4* lv_v, const float4* lv_n, int lw, int lh,
                      const float4* md_v, const float4* md_n, int mw, int mh, Intr k, int iterations,
                      const DeviceIcpParams& p, int level, bool single_launch, float check_max_rms) {
  const Maps m = make_maps(lv_v, lv_n, lw, lh, md_v, md_n, mw, mh, k);
  const int nb = blocks_for(lw * lh);
  if (iterations <= 0) {
    hipLaunchKernelGGL(icp_finish_level_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->icp, p, level,
                       check_max_rms);
    KF_HIP(hipGetLastError());
    return;
  }
  if (single_launch) {
    if (check_max_rms > 0.0F) fail(KF_ERR_INVALID_ARGUMENT, "icp level: the learned check needs two-kernel trials");
    static int resident = -1;  // workgroups of icp_level_kernel resident at once on this device
    if (resident < 0) {
      int per_cu = 0, dev = 0;
      KF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, icp_level_kernel, 256, 0));
      KF_HIP(hipGetDevice(&dev));
      hipDeviceProp_t prop{};
      KF_HIP(hipGetDeviceProperties(&prop, dev));
      resident = per_cu * prop.multiProcessorCount;
    }
    if (nb > resident) fail(KF_ERR_DEVICE, "icp level: the grid does not fit on the device at once");
    hipLaunchKernelGGL(icp_level_kernel, dim3(nb), dim3(256), 0, ctx->stream, m, p, ctx->icp, ctx->partials,
                       ctx->sync_timeout_ticks, iterations, level);
    KF_HIP(hipGetLastError());
    return;
  }
  // the level's last trial also finishes the level (and runs the learned check)
  for (int it = 0; it < iterations; ++it) {
    const int finish = it + 1 == iterations ? level : -1;
    hipLaunchKernelGGL(icp_trial_sweep_kernel, dim3(nb), dim3(256), 0, ctx->stream, m, p, ctx->icp,
                       ctx->partials);
    hipLaunchKernelGGL(icp_trial_solve_kernel, dim3(1), dim3(256), 0, ctx->stream, p, ctx->icp, ctx->partials, nb,
                       finish, finish >= 0 ? check_max_rms : 0.0F);
  }
  KF_HIP(hipGetLastError());
}

void launch_icp_reduce_once(kf_context* ctx, const float4* lv_v, const float4* lv_n, int lw, int lh,
                            const float4* md_v, const float4* md_n, int mw, int mh, Intr k,
                            const float* camera12, const float* model12, float max_dist,
                            float min_dot) {
  const Maps m = make_maps(lv_v, lv_n, lw, lh, md_v, md_n, mw, mh, k);
  const int nb = blocks_for(lw * lh);
  hipLaunchKernelGGL(icp_sweep_kernel, dim3(nb), dim3(256), 0, ctx->stream, m,
                     rigid_from12(camera12), rigid_from12(model12), max_dist, min_dot, ctx->partials);
  hipLaunchKernelGGL(icp_combine_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->icp, ctx->partials, nb);
  KF_HIP(hipGetLastError());
}

static_assert(sizeof(scratch) == sizeof(float) * kMaxBlocks * kPartialStride,
              "the static partial rows hold one row per sweep workgroup");

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
float* static_scratch_rows() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static float* rows[kMaxDevices] = {};
  int dev = 0;
  KF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(KF_ERR_UNSUPPORTED, "kf_icp_reduce_static: device index out of range");
  const std::lock_guard<std::mutex> lock(mu);
  if (rows[dev] == nullptr) {
    void* p = nullptr;
    KF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) fail(KF_ERR_DEVICE, "kf_icp_reduce_static: no device address for the partial rows");
    rows[dev] = static_cast<float*>(p);
  }
  return rows[dev];
}

void launch_icp_reduce_static(hipStream_t stream, const float4* lv_v, const float4* lv_n, int lw, int lh,
                              const float4* md_v, const float4* md_n, int mw, int mh, Intr k,
                              const float* camera12, const float* model12, float max_dist, float min_dot,
                              double* sums, int64_t* count) {
  float* rows = static_scratch_rows();
  const Maps m = make_maps(lv_v, lv_n, lw, lh, md_v, md_n, mw, mh, k);
  const int nb = blocks_for(lw * lh);  // <= kMaxBlocks rows, every one written by the sweep
  hipLaunchKernelGGL(icp_sweep_kernel, dim3(nb), dim3(256), 0, stream, m, rigid_from12(camera12),
                     rigid_from12(model12), max_dist, min_dot, rows);
  hipLaunchKernelGGL(icp_combine_out_kernel, dim3(1), dim3(256), 0, stream, rows, nb, sums, count);
  KF_HIP(hipGetLastError());
}

}  // namespace kf
