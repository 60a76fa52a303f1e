// Ahead-of-time kernels of the linear solve on gfx950: BCSR SpMV, the fused
// BiCGSTAB vector updates, deterministic reductions, Jacobi (point or
// vertex-block) and
// level-scheduled block ILU0.  These replace dune-istl's sequential
// BiCGSTABSolver / SeqJac / SeqILU that the reference builds on every Newton
// iteration (model/make_step_operator.hh:102-146,
// solver/istl/factory/{iterative,preconditioner}.hh).
//
// Vectors are slot-major with B species per slot; matrices are BCSR with
// row-major B x B blocks.  Every reduction is two-stage with a fixed order
// (per-workgroup partial sums, then one workgroup sums the partials), so
// results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <atomic>

#include "kernels.hh"
#include "common.hh"
#include "../host/error.hh"

namespace dcopasi::k {

// ----------------------------------------------------------------- vector kernels
template <typename T>
__global__ void fill_k(T* p, T v, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// the padded species of a slot get a unit diagonal so their rows stay decoupled
__global__ void pad_identity_k(double* val, const int* diag_pos, const int* slot_b, int nslot, int B) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)nslot * B) return;
  int s = (int)(i / B), c = (int)(i % B);
  if (c >= slot_b[s]) val[(size_t)diag_pos[s] * B * B + c * B + c] = 1.0;
}

__global__ void zero_blocks_k(double* __restrict__ val, const int* __restrict__ pos, long n, int bb) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n * bb) val[(size_t)pos[i / bb] * bb + i % bb] = 0.0;
}

// plain streaming copy of known size: calibrates HBM traffic counters
__global__ void probe_copy_k(long n, const double4* __restrict__ a, double4* __restrict__ b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

// y = a*x + y
__global__ void axpy_k(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

// partial sums of x.y (or |x| / max|x|) per workgroup
__global__ void reduce_k(long n, const double* __restrict__ x, const double* __restrict__ y, int op,
                         double* __restrict__ part) {
  __shared__ double lds[16];
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double a = x[i];
    if (op == 0) v[0] += a * (y ? y[i] : a);
    else if (op == 1) v[0] += fabs(a);
    else v[0] = fmax(v[0], fabs(a));
  }
  if (op == 2) {
    // max reduction
    double m = v[0];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
    __syncthreads();
    if (threadIdx.x % kWave == 0) lds[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = 0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) r = fmax(r, lds[w]);
      part[blockIdx.x] = r;
    }
    return;
  }
  block_sum(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void reduce_final_k(int nparts, const double* __restrict__ part, int op, double* __restrict__ out) {
  __shared__ double lds[16];
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) v[0] = op == 2 ? fmax(v[0], part[i]) : v[0] + part[i];
  if (op == 2) {
    double m = v[0];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
    __syncthreads();
    if (threadIdx.x % kWave == 0) lds[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = 0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) r = fmax(r, lds[w]);
      *out = r;
    }
    return;
  }
  block_sum(v, lds);
  if (threadIdx.x == 0) *out = v[0];
}

// Finaliser: one workgroup of 1024 threads sums the partial sums in a fixed
// order and updates the scalars.  (Fusing this into the producers with a
// last-workgroup ticket needs an agent-scope release per workgroup, which on
// a multi-XCD part writes back L2 and costs far more than a launch.)
__global__ void __launch_bounds__(1024) bicg_fin_k(int nparts, const double* __restrict__ part, int stride, int stage,
                                                   BicgScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double red[32];
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    v[0] += part[(size_t)i * stride];
    if (stride == 2) v[1] += part[(size_t)i * stride + 1];
  }
  block_sum(v, red);
  if (threadIdx.x == 0) bicg_update(sc, stage, v[0], v[1]);
}

// Stages 2 and 3 in one launch: the s-update's partial norms (stride 1,
// na of them) and then the second operator's partial dots (stride 2, nb
// pairs), each summed as bicg_fin_k sums it.  Deferring the half-step test
// past the second operator launch costs that launch once per solve (in the
// iteration that converges at the half step) and saves a launch in every
// iteration.
__global__ void __launch_bounds__(1024) bicg_fin23_k(int na, const double* __restrict__ pa, int nb,
                                                     const double* __restrict__ pb, BicgScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double red[32];
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < na; i += blockDim.x) v[0] += pa[i];
  block_sum(v, red);
  __shared__ double a;
  if (threadIdx.x == 0) {
    a = v[0];
    bicg_update(sc, 2, a, v[1]);
  }
  __syncthreads();
  if (sc->done) return;
  double u[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    u[0] += pb[2 * (size_t)i];
    u[1] += pb[2 * (size_t)i + 1];
  }
  block_sum(u, red);
  if (threadIdx.x == 0) bicg_update(sc, 3, u[0], u[1]);
}

// ----------------------------------------------------------------- SpMV
// y = A x over BCSR, one workgroup per chunk of rows whose block-row entries
// (blocks x B) fit the LDS stage (body in spmv_entries_chunk).  Optionally
// accumulates partial dots y.w1 and y.w2 per chunk.
template <int B, typename T, typename V>
__global__ void __launch_bounds__(kThreads) spmv_k(int nchunk, const int* __restrict__ chunk_row,
                                                   const int* __restrict__ rp, const int* __restrict__ col,
                                                   const T* __restrict__ val, const V* __restrict__ x,
                                                   V* __restrict__ y, const V* __restrict__ w1,
                                                   const V* __restrict__ w2, double* __restrict__ part,
                                                   BicgScalars* __restrict__ sc, int stage) {
  if (sc && sc->done) return;
  __shared__ StageT<T> staged[kSpmvStage];
  __shared__ double red[32];
  const int c = xcd_chunk(blockIdx.x, gridDim.x);
  double acc[2] = {0.0, 0.0};
  if (c < nchunk) {
    const int r0 = chunk_row[c], r1 = chunk_row[c + 1];
    spmv_entries_chunk<B, T, V, StageT<T>>(r0, r1, rp[r0], rp[r1], rp, col, val, x, y, w1, w2, acc, staged);
  }
  if (part) {
    block_sum(acc, red);
    if (threadIdx.x == 0 && c < nchunk) {
      part[2 * c] = acc[0];
      part[2 * c + 1] = acc[1];
    }
  }
}

// Row-split variant: G consecutive lanes share one BCSR row; lane l of the
// group takes the row's entries e = l, l + G, ... (entry = block x
// block-row), so with B | G every lane accumulates one fixed species and a
// shuffle-xor tree over the lanes of equal species finishes the row.  No LDS,
// no barrier; each workgroup owns a contiguous, XCD-local range of rows.
template <int B, int G, typename T, typename V>
__global__ void __launch_bounds__(kThreads) spmv_rows_k(int nrows, int rows_per_wg, const int* __restrict__ rp,
                                                        const int* __restrict__ col, const T* __restrict__ val,
                                                        const V* __restrict__ x, V* __restrict__ y,
                                                        const V* __restrict__ w1, const V* __restrict__ w2,
                                                        double* __restrict__ part, const BicgScalars* __restrict__ sc) {
  if (sc && sc->done) return;
  static_assert(G % B == 0 && G <= kWave, "group size");
  __shared__ double red[32];
  const int c = xcd_chunk(blockIdx.x, gridDim.x);
  const int r_begin = c * rows_per_wg;
  const int r_end = min(nrows, r_begin + rows_per_wg);
  const int lg = threadIdx.x % G, gid = threadIdx.x / G;
  constexpr int NG = kThreads / G;
  const int i = lg % B;
  double acc2[2] = {0.0, 0.0};
  for (int r = r_begin + gid; r < r_end; r += NG) {
    const int k0 = rp[r], ne = (rp[r + 1] - k0) * B;
    double acc = 0.0;
    int e = lg;
    for (; e + G < ne; e += 2 * G) {
      const int kb0 = k0 + e / B, kb1 = k0 + (e + G) / B;
      const T* a0 = val + (size_t)kb0 * (B * B) + i * B;
      const T* a1 = val + (size_t)kb1 * (B * B) + i * B;
      T v0[B], v1[B];
#pragma unroll
      for (int j = 0; j < B; ++j) {
        v0[j] = __builtin_nontemporal_load(a0 + j);
        v1[j] = __builtin_nontemporal_load(a1 + j);
      }
      const V* x0 = x + (size_t)col[kb0] * B;
      const V* x1 = x + (size_t)col[kb1] * B;
#pragma unroll
      for (int j = 0; j < B; ++j) acc += (double)v0[j] * (double)x0[j] + (double)v1[j] * (double)x1[j];
    }
    if (e < ne) {
      const int kb = k0 + e / B;
      const T* a = val + (size_t)kb * (B * B) + i * B;
      const V* xv = x + (size_t)col[kb] * B;
#pragma unroll
      for (int j = 0; j < B; ++j) acc += (double)__builtin_nontemporal_load(a + j) * (double)xv[j];
    }
#pragma unroll
    for (int off = B; off < G; off <<= 1) acc += __shfl_xor(acc, off, G);
    if (lg < B) {
      const size_t yi = (size_t)r * B + lg;
      y[yi] = (V)acc;
      const double ys = (double)(V)acc;
      if (w1) acc2[0] += ys * (double)w1[yi];
      if (w2) acc2[1] += ys * (double)w2[yi];
    }
  }
  if (part) {
    block_sum(acc2, red);
    if (threadIdx.x == 0) {
      part[2 * c] = acc2[0];
      part[2 * c + 1] = acc2[1];
    }
  }
}

// Block variant of the LDS-staged SpMV (body in spmv_blocks_chunk).
template <int B, typename T, typename V>
__global__ void __launch_bounds__(kThreads) spmv_blk_k(int nchunk, const int* __restrict__ chunk_row,
                                                       const int* __restrict__ rp, const int* __restrict__ col,
                                                       const T* __restrict__ val, const V* __restrict__ x,
                                                       V* __restrict__ y, const V* __restrict__ w1,
                                                       const V* __restrict__ w2, double* __restrict__ part,
                                                       const BicgScalars* __restrict__ sc) {
  if (sc && sc->done) return;
  __shared__ StageT<T> staged[kSpmvStage];
  __shared__ double red[32];
  const int c = xcd_chunk(blockIdx.x, gridDim.x);
  double acc[2] = {0.0, 0.0};
  if (c < nchunk) {
    const int r0 = chunk_row[c], r1 = chunk_row[c + 1];
    spmv_blocks_chunk<B, T, V, StageT<T>>(r0, r1, rp[r0], rp[r1], rp, col, val, x, y, w1, w2, acc, staged);
  }
  if (part) {
    block_sum(acc, red);
    if (threadIdx.x == 0 && c < nchunk) {
      part[2 * c] = acc[0];
      part[2 * c + 1] = acc[1];
    }
  }
}

// ----------------------------------------------------------------- Jacobi
template <int B>
__global__ void block_inverse_k(int nslot, const double* __restrict__ val, const int* __restrict__ diag_pos,
                                double* __restrict__ dinv, int* __restrict__ status, int point) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  jacobi_inverse_dev<B>(val + (size_t)diag_pos[s] * B * B, dinv, s, status, point != 0);
}

// ----------------------------------------------------------------- BiCGSTAB
// Scalars live on the device (BicgScalars); every kernel returns at once when
// the solve has finished, so the host enqueues iterations without waiting.
__device__ __forceinline__ bool finished(const BicgScalars* sc) { return sc->done != 0; }

// The p- and s-updates take kVecSlots slots per thread (kThreads apart, so
// every pass of the workgroup is contiguous): more bytes in flight per
// thread, and a quarter of the s-update's workgroup partials.
constexpr int kVecSlots = 4;

// p = r + beta (p - omega v); ph = M^-1 p (Jacobi fused, else copy)
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) bicg_p_k(int nslot, const V* __restrict__ r, V* __restrict__ p,
                                                     const V* __restrict__ v, V* __restrict__ ph,
                                                     const double* __restrict__ dinv, int diag,
                                                     const BicgScalars* __restrict__ sc) {
  if (finished(sc)) return;
  const double beta = sc->beta, omega = sc->omega;
#pragma unroll
  for (int u = 0; u < kVecSlots; ++u) {
    const int s = (blockIdx.x * kVecSlots + u) * blockDim.x + threadIdx.x;
    if (s < nslot) {
      const size_t k0 = (size_t)s * B;
      V rv[B], po[B], vv[B], pv[B], o[B];
#pragma unroll
      for (int i = 0; i < B; ++i) rv[i] = r[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) po[i] = p[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) vv[i] = v[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) pv[i] = (V)((double)rv[i] + beta * ((double)po[i] - omega * (double)vv[i]));
#pragma unroll
      for (int i = 0; i < B; ++i) p[k0 + i] = pv[i];
      if (dinv) dinv_apply<B>(dinv, s, pv, o, diag != 0);
      else
#pragma unroll
        for (int i = 0; i < B; ++i) o[i] = pv[i];
#pragma unroll
      for (int i = 0; i < B; ++i) ph[k0 + i] = o[i];
    }
  }
}

// s = r - alpha v; sh = M^-1 s; partial ||s||^2
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) bicg_s_k(int nslot, const V* __restrict__ r, const V* __restrict__ v,
                                                     V* __restrict__ sv, V* __restrict__ sh,
                                                     const double* __restrict__ dinv, int diag,
                                                     const BicgScalars* __restrict__ sc, double* __restrict__ part) {
  if (finished(sc)) return;
  __shared__ double red[16];
  double acc[1] = {0.0};
  const double alpha = sc->alpha;
#pragma unroll
  for (int u = 0; u < kVecSlots; ++u) {
    const int s = (blockIdx.x * kVecSlots + u) * blockDim.x + threadIdx.x;
    if (s < nslot) {
      const size_t k0 = (size_t)s * B;
      V rv[B], vv[B], t[B], o[B];
#pragma unroll
      for (int i = 0; i < B; ++i) rv[i] = r[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) vv[i] = v[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) {
        t[i] = (V)((double)rv[i] - alpha * (double)vv[i]);
        acc[0] += (double)t[i] * (double)t[i];
      }
#pragma unroll
      for (int i = 0; i < B; ++i) sv[k0 + i] = t[i];
      if (dinv) dinv_apply<B>(dinv, s, t, o, diag != 0);
      else
#pragma unroll
        for (int i = 0; i < B; ++i) o[i] = t[i];
#pragma unroll
      for (int i = 0; i < B; ++i) sh[k0 + i] = o[i];
    }
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// partial (y.w1), (y.w2) per workgroup over a fixed grid (deterministic)
template <typename V>
__global__ void __launch_bounds__(kThreads) dot_parts_k(long n, const V* __restrict__ y, const V* __restrict__ w1,
                                                        const V* __restrict__ w2, double* __restrict__ part,
                                                        const BicgScalars* __restrict__ sc) {
  if (sc && finished(sc)) return;
  __shared__ double red[32];
  double acc[2] = {0.0, 0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double v = y[i];
    acc[0] += v * static_cast<double>(w1[i]);
    if (w2) acc[1] += v * static_cast<double>(w2[i]);
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// x += alpha ph + omega sh ; r = s - omega t ; partial (rh.r), (r.r).
// Each thread updates kXrSlots slots (kThreads apart, so every pass of the
// workgroup is contiguous): more bytes in flight per thread and a quarter of
// the workgroup reductions of one slot per thread.
constexpr int kXrSlots = 4;
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) bicg_xr_k(int nslot, V* __restrict__ x, const V* __restrict__ ph,
                                                      const V* __restrict__ sh, const V* __restrict__ sv,
                                                      const V* __restrict__ t, V* __restrict__ r,
                                                      const V* __restrict__ rh, const BicgScalars* __restrict__ sc,
                                                      double* __restrict__ part) {
  if (finished(sc)) return;
  __shared__ double red[32];
  double acc[2] = {0.0, 0.0};
  const bool half = sc->half != 0;
  const double alpha = sc->alpha, omega = sc->omega;
  // each vector's B values of a slot are read together (adjacent accesses
  // pair into 16-byte loads and stores)
  if (half) {
#pragma unroll
    for (int u = 0; u < kXrSlots; ++u) {
      const int s = (blockIdx.x * kXrSlots + u) * blockDim.x + threadIdx.x;
      if (s < nslot) {
        const size_t k0 = (size_t)s * B;
        V xv[B], pv[B];
#pragma unroll
        for (int i = 0; i < B; ++i) xv[i] = x[k0 + i];
#pragma unroll
        for (int i = 0; i < B; ++i) pv[i] = ph[k0 + i];
#pragma unroll
        for (int i = 0; i < B; ++i) x[k0 + i] = (V)((double)xv[i] + alpha * (double)pv[i]);
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < kXrSlots; ++u) {
      const int s = (blockIdx.x * kXrSlots + u) * blockDim.x + threadIdx.x;
      if (s < nslot) {
        const size_t k0 = (size_t)s * B;
        V xv[B], pv[B], hv[B], sw[B], tv[B], rv0[B];
#pragma unroll
        for (int i = 0; i < B; ++i) xv[i] = x[k0 + i];
#pragma unroll
        for (int i = 0; i < B; ++i) pv[i] = ph[k0 + i];
#pragma unroll
        for (int i = 0; i < B; ++i) hv[i] = sh[k0 + i];
#pragma unroll
        for (int i = 0; i < B; ++i) sw[i] = sv[k0 + i];
#pragma unroll
        for (int i = 0; i < B; ++i) tv[i] = t[k0 + i];
#pragma unroll
        for (int i = 0; i < B; ++i) rv0[i] = rh[k0 + i];
        V xn[B], rn[B];
#pragma unroll
        for (int i = 0; i < B; ++i) {
          xn[i] = (V)((double)xv[i] + alpha * (double)pv[i] + omega * (double)hv[i]);
          rn[i] = (V)((double)sw[i] - omega * (double)tv[i]);
          acc[0] += (double)rv0[i] * (double)rn[i];
          acc[1] += (double)rn[i] * (double)rn[i];
        }
#pragma unroll
        for (int i = 0; i < B; ++i) x[k0 + i] = xn[i];
#pragma unroll
        for (int i = 0; i < B; ++i) r[k0 + i] = rn[i];
      }
    }
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// x = 0, r = rh = b, p = v = 0, partial b.b
template <typename V, typename W>
__global__ void __launch_bounds__(kThreads) bicg_init_k(long n, const W* __restrict__ b, V* __restrict__ x,
                                                        V* __restrict__ r, V* __restrict__ rh, V* __restrict__ p,
                                                        V* __restrict__ v, double* __restrict__ part,
                                                        BicgScalars* __restrict__ sc) {
  __shared__ double red[16];
  double acc[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const V bi = (V)b[i];
    x[i] = (V)0;
    r[i] = bi;
    rh[i] = bi;
    p[i] = (V)0;
    v[i] = (V)0;
    acc[0] += (double)bi * (double)bi;
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// ----------------------------------------------------------------- block ILU0
template <int B>
__device__ __forceinline__ void mat_mul_sub(double* __restrict__ c, const double* a, const double* b) {
  // c -= a * b
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) s += a[i * B + k] * b[k * B + j];
      c[i * B + j] -= s;
    }
}

// dinv_i = (the updated diagonal block of row i)^-1, Gauss-Jordan with
// partial pivoting; a zero pivot sets status 4
template <int B>
__device__ __forceinline__ void ilu_invert_block(const double* d, double* __restrict__ o, int* __restrict__ status) {
  double a[B][B], inv[B][B];
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int c = 0; c < B; ++c) {
      a[r][c] = d[r * B + c];
      inv[r][c] = r == c ? 1.0 : 0.0;
    }
  for (int c = 0; c < B; ++c) {
    int piv = c;
    double best = fabs(a[c][c]);
    for (int r = c + 1; r < B; ++r)
      if (fabs(a[r][c]) > best) {
        best = fabs(a[r][c]);
        piv = r;
      }
    if (best == 0.0) {
      status[0] = 4;
      a[piv][c] = 1.0;
    }
    if (piv != c)
      for (int j = 0; j < B; ++j) {
        double tt = a[c][j];
        a[c][j] = a[piv][j];
        a[piv][j] = tt;
        tt = inv[c][j];
        inv[c][j] = inv[piv][j];
        inv[piv][j] = tt;
      }
    const double pv = 1.0 / a[c][c];
    for (int j = 0; j < B; ++j) {
      a[c][j] *= pv;
      inv[c][j] *= pv;
    }
    for (int r = 0; r < B; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int j = 0; j < B; ++j) {
        a[r][j] -= f * a[c][j];
        inv[r][j] -= f * inv[c][j];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int c = 0; c < B; ++c) o[r * B + c] = inv[r][c];
}

template <int B>
__device__ __forceinline__ void ilu_invert_diag(int i, const int* __restrict__ diag_pos, const double* __restrict__ lu,
                                                double* __restrict__ dinv, int* __restrict__ status) {
  ilu_invert_block<B>(lu + (size_t)diag_pos[i] * B * B, dinv + (size_t)i * B * B, status);
}

// Factor the rows of one level: L_ik <- A_ik U_kk^-1, A_ij -= L_ik U_kj,
// U_ii^-1 stored in dinv.  lu holds a copy of the matrix values.
template <int B>
__device__ __forceinline__ void ilu_factor_row(int i, const int* __restrict__ rp, const int* __restrict__ col,
                                               const int* __restrict__ diag_pos, double* __restrict__ lu,
                                               double* __restrict__ dinv, int* __restrict__ status) {
  for (int p = rp[i]; p < diag_pos[i]; ++p) {
    const int k = col[p];
    double* lik = lu + (size_t)p * B * B;
    double tmp[B * B];
    const double* dk = dinv + (size_t)k * B * B;
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < B; ++c) s += lik[a * B + c] * dk[c * B + b];
        tmp[a * B + b] = s;
      }
#pragma unroll
    for (int a = 0; a < B * B; ++a) lik[a] = tmp[a];
    // row i entries right of column k against row k's upper part
    int q = diag_pos[k] + 1, qe = rp[k + 1];
    int m = p + 1, me = rp[i + 1];
    while (q < qe && m < me) {
      const int cq = col[q], cm = col[m];
      if (cq == cm) {
        mat_mul_sub<B>(lu + (size_t)m * B * B, lik, lu + (size_t)q * B * B);
        ++q;
        ++m;
      } else if (cq < cm) {
        ++q;
      } else {
        ++m;
      }
    }
  }
  ilu_invert_diag<B>(i, diag_pos, lu, dinv, status);
}

template <int B>
__global__ void ilu_factor_k(int nrows, const int* __restrict__ rows, const int* __restrict__ rp,
                             const int* __restrict__ col, const int* __restrict__ diag_pos, double* __restrict__ lu,
                             double* __restrict__ dinv, int* __restrict__ status) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  ilu_factor_row<B>(rows[t], rp, col, diag_pos, lu, dinv, status);
}

// All levels in one workgroup: the rows of a level (rows[lev[l] .. lev[l+1]))
// in parallel, a barrier between levels.  A row reads only rows of earlier
// levels, which the barrier makes visible; for level widths a workgroup
// covers in a few passes this replaces one launch per level.
template <int B>
__global__ void __launch_bounds__(256) ilu_factor_sweep_k(int nlev, const int* __restrict__ lev,
                                                          const int* __restrict__ rows, const int* __restrict__ rp,
                                                          const int* __restrict__ col,
                                                          const int* __restrict__ diag_pos, double* __restrict__ lu,
                                                          double* __restrict__ dinv, int* __restrict__ status) {
  for (int l = 0; l < nlev; ++l) {
    const int e = lev[l + 1];
    for (int t = lev[l] + (int)threadIdx.x; t < e; t += blockDim.x)
      ilu_factor_row<B>(rows[t], rp, col, diag_pos, lu, dinv, status);
    __syncthreads();
  }
}

// The same sweep for rows of small blocks (B <= 2) with at most
// kIluRegLower lower blocks, kIluRegPairs update pairs, kIluRegRow blocks
// and kIluRegPush pushes, every operand at an address fixed by the row's
// position r in level order:
//   packed[r]: the row's blocks (gathered by ilu_pack_k before the sweep);
//   inbox[r]: slot a < nl: dinv_k of lower block a; slot kIluRegLower + u:
//   U_kq of update pair u, written there ("pushed") by row k when it was
//   factored;
//   fub[r], fm[r]: the pair ranges of the lower blocks and the pairs' row
//   blocks m (local index); fpush[r]: (source, destination) of this row's
//   pushes (source -1: its dinv, else one of its blocks).
// All of a row's loads are issued together and it computes in LDS (its
// row) and registers (its operands): no load waits for a store of the row.
// The arithmetic is that of ilu_factor_row: lower blocks in column order,
// each followed by its pairs in merge order.
template <int B>
__global__ void ilu_pack_k(int ns, IluSweep t, const double* __restrict__ val, double* __restrict__ packed) {
  constexpr int BB = B * B;
  const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long per = (long)kIluRegRow * BB;
  if (g >= (long)ns * per) return;
  const int r = (int)(g / per), b = (int)(g % per);
  const int4 f = t.frow[r];
  if (b < f.z * BB) packed[g] = val[(size_t)f.y * BB + b];
}

extern __shared__ double ilu_row_lds[];
template <int B>
__global__ void __launch_bounds__(256) ilu_factor_inbox_k(IluSweep t, const double* __restrict__ packed,
                                                          double* __restrict__ inbox, double* __restrict__ lu,
                                                          double* __restrict__ dinv, int* __restrict__ status) {
  constexpr int BB = B * B, ML = kIluRegLower, MU = kIluRegPairs, MR = kIluRegRow, MP = kIluRegPush;
  constexpr int IN = ML + MU;
  double* my = ilu_row_lds + (size_t)threadIdx.x * MR * BB;  // the thread's row
  for (int l = 0; l < t.nlo; ++l) {
    const int e = t.lo_lev[l + 1];
    for (int r = t.lo_lev[l] + (int)threadIdx.x; r < e; r += blockDim.x) {
      const int4 f = t.frow[r];  // row, first block, blocks, lower blocks
      const int nl = f.w;
      // every load of the row first: its blocks, its inbox, its tables
      // (16-byte loads where the layout allows: a single workgroup moves
      // every byte of the sweep through one CU, so the instruction count of
      // these per-lane gathers bounds a level)
      double rv[MR * BB], in[IN][BB];
      int ub[kIluFubStride], fm[MU];
      int2 push[MP];
      if constexpr (BB % 2 == 0) {
        const double2* pr = reinterpret_cast<const double2*>(packed + (size_t)r * MR * BB);
#pragma unroll
        for (int j = 0; j < MR * BB / 2; ++j) {
          const double2 v = 2 * j < f.z * BB ? pr[j] : make_double2(0.0, 0.0);
          rv[2 * j] = v.x;
          rv[2 * j + 1] = v.y;
        }
        const double2* pi = reinterpret_cast<const double2*>(inbox + (size_t)r * IN * BB);
#pragma unroll
        for (int j = 0; j < IN; ++j)
#pragma unroll
          for (int c = 0; c < BB / 2; ++c) {
            const double2 v = pi[j * BB / 2 + c];
            in[j][2 * c] = v.x;
            in[j][2 * c + 1] = v.y;
          }
      } else {
#pragma unroll
        for (int j = 0; j < MR * BB; ++j) rv[j] = j < f.z * BB ? packed[(size_t)r * MR * BB + j] : 0.0;
#pragma unroll
        for (int j = 0; j < IN; ++j)
#pragma unroll
          for (int c = 0; c < BB; ++c) in[j][c] = inbox[((size_t)r * IN + j) * BB + c];
      }
      {
        const int4* pu = reinterpret_cast<const int4*>(t.fub + (size_t)r * kIluFubStride);
#pragma unroll
        for (int a = 0; a < kIluFubStride / 4; ++a) {
          const int4 v = pu[a];
          ub[4 * a] = v.x;
          ub[4 * a + 1] = v.y;
          ub[4 * a + 2] = v.z;
          ub[4 * a + 3] = v.w;
        }
        const int4* pm = reinterpret_cast<const int4*>(t.fm + (size_t)r * MU);
#pragma unroll
        for (int u = 0; u < MU / 4; ++u) {
          const int4 v = pm[u];
          fm[4 * u] = v.x;
          fm[4 * u + 1] = v.y;
          fm[4 * u + 2] = v.z;
          fm[4 * u + 3] = v.w;
        }
        const int4* pp = reinterpret_cast<const int4*>(t.fpush + (size_t)r * MP);
#pragma unroll
        for (int j = 0; j < MP / 2; ++j) {
          const int4 v = pp[j];
          push[2 * j] = make_int2(v.x, v.y);
          push[2 * j + 1] = make_int2(v.z, v.w);
        }
      }
#pragma unroll
      for (int j = 0; j < MR * BB; ++j) my[j] = rv[j];
#pragma unroll
      for (int a = 0; a < ML; ++a) {
        if (a < nl) {
          double* lik = my + a * BB;
          double tmp[BB];
#pragma unroll
          for (int x = 0; x < B; ++x)
#pragma unroll
            for (int y = 0; y < B; ++y) {
              double sum = 0.0;
#pragma unroll
              for (int c = 0; c < B; ++c) sum += lik[x * B + c] * in[a][c * B + y];
              tmp[x * B + y] = sum;
            }
#pragma unroll
          for (int c = 0; c < BB; ++c) lik[c] = tmp[c];
#pragma unroll
          for (int u = 0; u < MU; ++u)
            if (u >= ub[a] && u < ub[a + 1]) mat_mul_sub<B>(my + fm[u] * BB, tmp, in[ML + u]);
        }
      }
      double inv[BB];
      ilu_invert_block<B>(my + nl * BB, inv, status);
      // stores: the row's blocks, its inverse, its pushes
      if constexpr (BB % 2 == 0) {
        double2* lr = reinterpret_cast<double2*>(lu + (size_t)f.y * BB);
        for (int j = 0; j < f.z * BB / 2; ++j) lr[j] = make_double2(my[2 * j], my[2 * j + 1]);
        double2* di = reinterpret_cast<double2*>(dinv + (size_t)f.x * BB);
#pragma unroll
        for (int c = 0; c < BB / 2; ++c) di[c] = make_double2(inv[2 * c], inv[2 * c + 1]);
#pragma unroll
        for (int k = 0; k < MP; ++k)
          if (push[k].y >= 0) {
            const double* v = push[k].x < 0 ? inv : my + push[k].x * BB;
            double2* d = reinterpret_cast<double2*>(inbox + (size_t)push[k].y);
#pragma unroll
            for (int c = 0; c < BB / 2; ++c) d[c] = make_double2(v[2 * c], v[2 * c + 1]);
          }
      } else {
        for (int j = 0; j < f.z * BB; ++j) lu[(size_t)f.y * BB + j] = my[j];
#pragma unroll
        for (int c = 0; c < BB; ++c) dinv[(size_t)f.x * BB + c] = inv[c];
#pragma unroll
        for (int k = 0; k < MP; ++k)
          if (push[k].y >= 0) {
            const double* v = push[k].x < 0 ? inv : my + push[k].x * BB;
#pragma unroll
            for (int c = 0; c < BB; ++c) inbox[(size_t)push[k].y + c] = v[c];
          }
      }
    }
    __syncthreads();
  }
}

// forward solve (unit lower): y_i = b_i - sum_{k<i} L_ik y_k
template <int B, typename V>
__device__ __forceinline__ void ilu_lower_row(int i, const int* __restrict__ rp, const int* __restrict__ col,
                                              const int* __restrict__ diag_pos, const double* __restrict__ lu,
                                              const V* __restrict__ b, V* __restrict__ y) {
  double s[B];
#pragma unroll
  for (int a = 0; a < B; ++a) s[a] = (double)b[(size_t)i * B + a];
  for (int p = rp[i]; p < diag_pos[i]; ++p) {
    const double* l = lu + (size_t)p * B * B;
    const V* yk = y + (size_t)col[p] * B;
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
      for (int c = 0; c < B; ++c) s[a] -= l[a * B + c] * (double)yk[c];
  }
#pragma unroll
  for (int a = 0; a < B; ++a) y[(size_t)i * B + a] = (V)s[a];
}

// backward solve: x_i = U_ii^-1 (y_i - sum_{k>i} U_ik x_k)
template <int B, typename V>
__device__ __forceinline__ void ilu_upper_row(int i, const int* __restrict__ rp, const int* __restrict__ col,
                                              const int* __restrict__ diag_pos, const double* __restrict__ lu,
                                              const double* __restrict__ dinv, const V* __restrict__ y,
                                              V* __restrict__ x) {
  double s[B];
#pragma unroll
  for (int a = 0; a < B; ++a) s[a] = (double)y[(size_t)i * B + a];
  for (int p = diag_pos[i] + 1; p < rp[i + 1]; ++p) {
    const double* u = lu + (size_t)p * B * B;
    const V* xk = x + (size_t)col[p] * B;
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
      for (int c = 0; c < B; ++c) s[a] -= u[a * B + c] * (double)xk[c];
  }
  const double* d = dinv + (size_t)i * B * B;
#pragma unroll
  for (int a = 0; a < B; ++a) {
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < B; ++c) v += d[a * B + c] * s[c];
    x[(size_t)i * B + a] = (V)v;
  }
}

template <int B, typename V>
__global__ void ilu_lower_k(int nrows, const int* __restrict__ rows, const int* __restrict__ rp,
                            const int* __restrict__ col, const int* __restrict__ diag_pos,
                            const double* __restrict__ lu, const V* __restrict__ b, V* __restrict__ y,
                            const BicgScalars* __restrict__ sc) {
  if (sc && finished(sc)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  ilu_lower_row<B, V>(rows[t], rp, col, diag_pos, lu, b, y);
}

template <int B, typename V>
__global__ void ilu_upper_k(int nrows, const int* __restrict__ rows, const int* __restrict__ rp,
                            const int* __restrict__ col, const int* __restrict__ diag_pos,
                            const double* __restrict__ lu, const double* __restrict__ dinv, const V* __restrict__ y,
                            V* __restrict__ x, const BicgScalars* __restrict__ sc) {
  if (sc && finished(sc)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  ilu_upper_row<B, V>(rows[t], rp, col, diag_pos, lu, dinv, y, x);
}

// both triangular solves, every level, in one workgroup (barriers between
// levels and between the two sweeps): out = U^-1 L^-1 in, tmp holds L^-1 in
template <int B, typename V>
__global__ void __launch_bounds__(256) ilu_apply_sweep_k(int nlo, const int* __restrict__ lo_lev,
                                                         const int* __restrict__ lo_rows, int nup,
                                                         const int* __restrict__ up_lev,
                                                         const int* __restrict__ up_rows, const int* __restrict__ rp,
                                                         const int* __restrict__ col,
                                                         const int* __restrict__ diag_pos,
                                                         const double* __restrict__ lu,
                                                         const double* __restrict__ dinv, const V* __restrict__ in,
                                                         V* __restrict__ tmp, V* __restrict__ out,
                                                         const BicgScalars* __restrict__ sc) {
  if (sc && finished(sc)) return;
  for (int l = 0; l < nlo; ++l) {
    const int e = lo_lev[l + 1];
    for (int t = lo_lev[l] + (int)threadIdx.x; t < e; t += blockDim.x)
      ilu_lower_row<B, V>(lo_rows[t], rp, col, diag_pos, lu, in, tmp);
    __syncthreads();
  }
  for (int l = 0; l < nup; ++l) {
    const int e = up_lev[l + 1];
    for (int t = up_lev[l] + (int)threadIdx.x; t < e; t += blockDim.x)
      ilu_upper_row<B, V>(up_rows[t], rp, col, diag_pos, lu, dinv, tmp, out);
    __syncthreads();
  }
}

// One sweep of ilu_apply_chunk_k: the levels in chunks (chunk[c] ..
// chunk[c+1]).  A chunk's blocks, source slots, row ids, entry offsets,
// right-hand sides and (backward sweep) diagonal inverses are first gathered
// into LDS with every load in flight at once; its levels then run from LDS
// alone, each row's result going to the ring slot (l mod ring) * width + r -
// lev[l] (value a of slot q at ring[a * ring * width + q]); at the chunk's
// end the results are written to `res`.  The sums run in the blocks' column
// order, as ilu_lower_row / ilu_upper_row sum them.
// Byte offsets of the chunk stage in LDS (kernel and launcher): blocks,
// diagonal inverses, right-hand sides (doubles), then source slots, entry
// offsets, row ids and level bounds (ints), then the ring (8-byte aligned).
// With reg_k = K > 0 every row has K entry places (padding has source -1)
// and the rows' values lie component-major: block entry a of place k of
// chunk row rr at [(k * B*B + a) * stage_rows + rr], source [k * stage_rows
// + rr], right-hand side and diagonal inverse [a * stage_rows + rr], so the
// rows of one level (consecutive rr, one per lane) read consecutive words.
struct IluChunkLayout {
  size_t d, in, src, ent, row, lev, ring, total;
  __host__ __device__ IluChunkLayout(const IluSweep& t, int B, int vbytes) {
    const size_t BB = (size_t)B * B;
    const size_t nent = t.reg_k > 0 ? (size_t)t.reg_k * t.stage_rows : (size_t)t.stage_ent;
    d = 8 * nent * BB;
    in = d + 8 * (size_t)t.stage_rows * BB;
    src = in + 8 * (size_t)t.stage_rows * B;
    ent = src + 4 * nent;
    row = ent + 4 * ((size_t)t.stage_rows + 1);
    lev = row + 4 * (size_t)t.stage_rows;
    ring = (lev + 4 * ((size_t)t.stage_levels + 1) + 7) & ~(size_t)7;
    total = ring + (size_t)t.ring * t.width * B * vbytes;
  }
};

// A thread's row of one level, read from the chunk stage (K > 0 below).
template <int B, int K, bool UPPER>
struct IluStagedRow {
  int rr = -1, o = 0;
  int src[K];
  double L[K][B * B], in[B], d[UPPER ? B * B : 1];
};

template <int B, typename V, bool UPPER, int K = 0>
__device__ __forceinline__ void ilu_chunk_sweep(const IluSweep& t, int nchunk, const int* __restrict__ chunk,
                                                const int* __restrict__ lev, const int* __restrict__ rows,
                                                const int* __restrict__ ent, const int* __restrict__ blk,
                                                const int* __restrict__ src, const double* __restrict__ lu,
                                                const double* __restrict__ dinv, const V* __restrict__ rhs,
                                                V* __restrict__ res, char* lds) {
  constexpr int BB = B * B;
  const IluChunkLayout lay(t, B, (int)sizeof(V));
  double* sL = reinterpret_cast<double*>(lds);                  // stage_ent * BB
  double* sD = reinterpret_cast<double*>(lds + lay.d);          // stage_rows * BB (backward sweep)
  double* sIn = reinterpret_cast<double*>(lds + lay.in);        // stage_rows * B
  int* sSrc = reinterpret_cast<int*>(lds + lay.src);            // stage_ent
  int* sEnt = reinterpret_cast<int*>(lds + lay.ent);            // stage_rows + 1
  int* sRow = reinterpret_cast<int*>(lds + lay.row);            // stage_rows
  int* sLev = reinterpret_cast<int*>(lds + lay.lev);            // stage_levels + 1
  V* ring = reinterpret_cast<V*>(lds + lay.ring);               // B * ring * width, component-major
  const int RS = t.ring * t.width;                              // ring slots
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int c = 0; c < nchunk; ++c) {
    const int L0 = chunk[c], L1 = chunk[c + 1];
    const int r0 = lev[L0], r1 = lev[L1];
    const int e0 = ent[r0], e1 = ent[r1];
    // the stage: KE entries (KR rows) per thread per round, every round's
    // index loads, then its value loads, issued together before any store
    constexpr int KE = BB <= 4 ? 8 : BB <= 16 ? 4 : 1, KR = 4;
    for (int l = L0 + tid; l <= L1; l += nt) sLev[l - L0] = lev[l];
    if constexpr (K == 0)
      for (int r = r0 + tid; r <= r1; r += nt) sEnt[r - r0] = ent[r] - e0;
    const int SR = t.stage_rows;
    for (int rb = r0; rb < r1; rb += nt * KR) {
      int ri[KR];
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        const int r = rb + tid + k * nt;
        ri[k] = r < r1 ? rows[r] : -1;
      }
      double vin[KR][B], vd[KR][UPPER ? BB : 1];
#pragma unroll
      for (int k = 0; k < KR; ++k)
        if (ri[k] >= 0) {
#pragma unroll
          for (int a = 0; a < B; ++a) vin[k][a] = (double)rhs[(size_t)ri[k] * B + a];
          if constexpr (UPPER)
#pragma unroll
            for (int a = 0; a < BB; ++a) vd[k][a] = dinv[(size_t)ri[k] * BB + a];
        }
#pragma unroll
      for (int k = 0; k < KR; ++k)
        if (ri[k] >= 0) {
          const int rr = rb + tid + k * nt - r0;
          sRow[rr] = ri[k];
#pragma unroll
          for (int a = 0; a < B; ++a) sIn[K > 0 ? a * SR + rr : rr * B + a] = vin[k][a];
          if constexpr (UPPER)
#pragma unroll
            for (int a = 0; a < BB; ++a) sD[K > 0 ? a * SR + rr : rr * BB + a] = vd[k][a];
        }
    }
    if constexpr (K > 0) {
      // the padded places (row-major in blk / src: row r's place k at
      // r * K + k), taken rr-fastest so that the LDS stores are consecutive
      const int nr = r1 - r0, nq = nr * K;
      for (int qb = 0; qb < nq; qb += nt * KE) {
        int bi[KE], si[KE], rq[KE], kq[KE];
#pragma unroll
        for (int k = 0; k < KE; ++k) {
          const int q = qb + tid + k * nt;
          kq[k] = q / (nr > 0 ? nr : 1);
          rq[k] = q - kq[k] * nr;
          const bool ok = q < nq;
          bi[k] = ok ? blk[(size_t)(r0 + rq[k]) * K + kq[k]] : -1;
          si[k] = ok ? src[(size_t)(r0 + rq[k]) * K + kq[k]] : -1;
          if (!ok) kq[k] = -1;
        }
        double v[KE][BB];
#pragma unroll
        for (int k = 0; k < KE; ++k)
          if (bi[k] >= 0)
#pragma unroll
            for (int a = 0; a < BB; ++a) v[k][a] = lu[(size_t)bi[k] * BB + a];
#pragma unroll
        for (int k = 0; k < KE; ++k)
          if (kq[k] >= 0) {
            sSrc[kq[k] * SR + rq[k]] = bi[k] >= 0 ? si[k] : -1;
            if (bi[k] >= 0)
#pragma unroll
              for (int a = 0; a < BB; ++a) sL[(size_t)(kq[k] * BB + a) * SR + rq[k]] = v[k][a];
          }
      }
    } else
    for (int eb = e0; eb < e1; eb += nt * KE) {
      int bi[KE], si[KE];
#pragma unroll
      for (int k = 0; k < KE; ++k) {
        const int e = eb + tid + k * nt;
        bi[k] = e < e1 ? blk[e] : -1;
        si[k] = e < e1 ? src[e] : 0;
      }
      double v[KE][BB];
#pragma unroll
      for (int k = 0; k < KE; ++k)
        if (bi[k] >= 0)
#pragma unroll
          for (int a = 0; a < BB; ++a) v[k][a] = lu[(size_t)bi[k] * BB + a];
#pragma unroll
      for (int k = 0; k < KE; ++k)
        if (bi[k] >= 0) {
          const int ee = eb + tid + k * nt - e0;
          sSrc[ee] = si[k];
#pragma unroll
          for (int a = 0; a < BB; ++a) sL[(size_t)ee * BB + a] = v[k][a];
        }
    }
    __syncthreads();
    if constexpr (K > 0) {
      // at most K blocks per row and one row per thread per level: the
      // thread's row of level l + 1 is read from the stage while level l
      // runs, so a level's critical path is its ring reads, its sums and
      // one barrier (the sums in the same order as below)
      using Row = IluStagedRow<B, K, UPPER>;
      auto fetch = [&](int l, Row& R) {
        R.rr = -1;
        if (l >= L1) return;
        const int b0 = sLev[l - L0], r = b0 + tid;
        if (r >= sLev[l + 1 - L0]) return;
        const int rr = r - r0;
        R.rr = rr;
        R.o = (l % t.ring) * t.width + (r - b0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          R.src[k] = sSrc[k * SR + rr];
#pragma unroll
          for (int a = 0; a < BB; ++a) R.L[k][a] = sL[(size_t)(k * BB + a) * SR + rr];
        }
#pragma unroll
        for (int a = 0; a < B; ++a) R.in[a] = sIn[a * SR + rr];
        if constexpr (UPPER)
#pragma unroll
          for (int a = 0; a < BB; ++a) R.d[a] = sD[(size_t)a * SR + rr];
      };
      Row cur, nxt;
      fetch(L0, cur);
      for (int l = L0; l < L1; ++l) {
        double yv[K][B];
        if (cur.rr >= 0) {
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (cur.src[k] >= 0)
#pragma unroll
              for (int cc = 0; cc < B; ++cc) yv[k][cc] = (double)ring[(size_t)cc * RS + cur.src[k]];
        }
        fetch(l + 1, nxt);
        if (cur.rr >= 0) {
          double sum[B];
#pragma unroll
          for (int a = 0; a < B; ++a) sum[a] = cur.in[a];
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (cur.src[k] >= 0)
#pragma unroll
              for (int a = 0; a < B; ++a)
#pragma unroll
                for (int cc = 0; cc < B; ++cc) sum[a] -= cur.L[k][a * B + cc] * yv[k][cc];
          V* o = ring + cur.o;
          if constexpr (UPPER) {
#pragma unroll
            for (int a = 0; a < B; ++a) {
              double v = 0.0;
#pragma unroll
              for (int cc = 0; cc < B; ++cc) v += cur.d[a * B + cc] * sum[cc];
              o[(size_t)a * RS] = (V)v;
            }
          } else {
#pragma unroll
            for (int a = 0; a < B; ++a) o[(size_t)a * RS] = (V)sum[a];
          }
        }
        __syncthreads();
        cur = nxt;
      }
    } else
    for (int l = L0; l < L1; ++l) {
      const int b0 = sLev[l - L0], b1 = sLev[l + 1 - L0];
      const int base = (l % t.ring) * t.width - b0;
      for (int r = b0 + tid; r < b1; r += nt) {
        const int rr = r - r0;
        double sum[B];
#pragma unroll
        for (int a = 0; a < B; ++a) sum[a] = sIn[rr * B + a];
        const int ee = sEnt[rr + 1];
        for (int e = sEnt[rr]; e < ee; ++e) {
          const double* L = sL + (size_t)e * BB;
          const V* yk = ring + sSrc[e];
#pragma unroll
          for (int a = 0; a < B; ++a)
#pragma unroll
            for (int cc = 0; cc < B; ++cc) sum[a] -= L[a * B + cc] * (double)yk[(size_t)cc * RS];
        }
        V* o = ring + (base + r);
        if constexpr (UPPER) {
          const double* d = sD + (size_t)rr * BB;
#pragma unroll
          for (int a = 0; a < B; ++a) {
            double v = 0.0;
#pragma unroll
            for (int cc = 0; cc < B; ++cc) v += d[a * B + cc] * sum[cc];
            o[(size_t)a * RS] = (V)v;
          }
        } else {
#pragma unroll
          for (int a = 0; a < B; ++a) o[(size_t)a * RS] = (V)sum[a];
        }
      }
      __syncthreads();
    }
    for (int l = L0; l < L1; ++l) {
      const int b0 = sLev[l - L0], b1 = sLev[l + 1 - L0];
      const int base = (l % t.ring) * t.width - b0;
      for (int r = b0 + tid; r < b1; r += nt) {
        const V* o = ring + (base + r);
#pragma unroll
        for (int a = 0; a < B; ++a) res[(size_t)sRow[r - r0] * B + a] = o[(size_t)a * RS];
      }
    }
    __syncthreads();
  }
}

// Both triangular solves of ILU0 in one workgroup, the levels in chunks
// staged through LDS (ilu_chunk_sweep): out = U^-1 L^-1 in, tmp = L^-1 in.
// K > 0: the register-pipelined level loop (rows of at most K blocks, levels
// of at most kIluThreads rows; K * B * B <= 32 keeps two rows in registers).
extern __shared__ char ilu_chunk_lds[];
template <int B, typename V, int K>
__global__ void __launch_bounds__(kIluThreads) ilu_apply_chunk_k(IluSweep t, const double* __restrict__ lu,
                                                                 const double* __restrict__ dinv,
                                                                 const V* __restrict__ in, V* __restrict__ tmp,
                                                                 V* __restrict__ out,
                                                                 const BicgScalars* __restrict__ sc) {
  if (sc && finished(sc)) return;
  ilu_chunk_sweep<B, V, false, K>(t, t.nlo_chunk, t.lo_chunk, t.lo_lev, t.lo_rows, t.lo_ent, t.lo_blk, t.lo_src,
                                  lu, nullptr, in, tmp, ilu_chunk_lds);
  ilu_chunk_sweep<B, V, true, K>(t, t.nup_chunk, t.up_chunk, t.up_lev, t.up_rows, t.up_ent, t.up_blk, t.up_src, lu,
                                 dinv, tmp, out, ilu_chunk_lds);
}

// ----------------------------------------------------------------- state I/O helpers
// packed (per compartment, node-major, b species) <-> padded (slot-major, B)
__global__ void unpack_k(long nslot, int B, const int* __restrict__ slot_b, const long* __restrict__ slot_dof,
                         const double* __restrict__ packed, double* __restrict__ padded) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nslot * B) return;
  long s = i / B;
  int c = (int)(i % B);
  padded[i] = c < slot_b[s] ? packed[slot_dof[s] + c] : 0.0;
}

__global__ void pack_k(long nslot, int B, const int* __restrict__ slot_b, const long* __restrict__ slot_dof,
                       const double* __restrict__ padded, double* __restrict__ packed) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nslot * B) return;
  long s = i / B;
  int c = (int)(i % B);
  if (c < slot_b[s]) packed[slot_dof[s] + c] = padded[i];
}

// precision conversion
template <typename A, typename Bt>
__global__ void convert_k(long n, const A* __restrict__ a, Bt* __restrict__ b) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (Bt)a[i];
}

// x -= z (Newton update), y = a*x + b*y
__global__ void axpby_k(long n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

// per-workgroup partial sums of w[s] * x[s*B+i] for every species i of a slot range
__global__ void __launch_bounds__(kThreads) weighted_sum_k(long first, long n, int B, const double* __restrict__ w,
                                                           const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double red[16 * 8];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long s = first + k;
    for (int i = 0; i < B; ++i) acc[i] += w[s] * x[s * B + i];
  }
  block_sum(acc, red);
  if (threadIdx.x == 0)
    for (int i = 0; i < 8; ++i) part[blockIdx.x * 8 + i] = acc[i];
}

__global__ void weighted_sum_final_k(int nparts, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[16 * 8];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = threadIdx.x; p < nparts; p += blockDim.x)
    for (int i = 0; i < 8; ++i) acc[i] += part[p * 8 + i];
  block_sum(acc, red);
  if (threadIdx.x == 0)
    for (int i = 0; i < 8; ++i) out[i] = acc[i];
}

// ----------------------------------------------------------------- host launchers
static inline unsigned grid_for(long n, int t = kThreads) { return (unsigned)((n + t - 1) / t); }

void fill(double* p, double v, long n, hipStream_t s) {
  if (n > 0) fill_k<double><<<grid_for(n), kThreads, 0, s>>>(p, v, n);
}
void fill(float* p, float v, long n, hipStream_t s) {
  if (n > 0) fill_k<float><<<grid_for(n), kThreads, 0, s>>>(p, v, n);
}
void fill(int* p, int v, long n, hipStream_t s) {
  if (n > 0) fill_k<int><<<grid_for(n), kThreads, 0, s>>>(p, v, n);
}

void init_jacobian(double* val, long nval, const int* diag_pos, const int* slot_b, int nslot, int B, hipStream_t s) {
  fill(val, 0.0, nval, s);
  pad_identity_k<<<grid_for((long)nslot * B), kThreads, 0, s>>>(val, diag_pos, slot_b, nslot, B);
}

void probe_copy(const void* a, void* b, long bytes, hipStream_t s) {
  const long n = bytes / (long)sizeof(double4);
  probe_copy_k<<<2048, kThreads, 0, s>>>(n, static_cast<const double4*>(a), static_cast<double4*>(b));
}

void zero_blocks(double* val, const int* pos, long n, int B, hipStream_t s) {
  if (n > 0) zero_blocks_k<<<grid_for(n * B * B), kThreads, 0, s>>>(val, pos, n, B * B);
}

void axpy(long n, double a, const double* x, double* y, hipStream_t s) {
  if (n > 0) axpy_k<<<grid_for(n), kThreads, 0, s>>>(n, a, x, y);
}
void axpby(long n, double a, const double* x, double b, double* y, hipStream_t s) {
  if (n > 0) axpby_k<<<grid_for(n), kThreads, 0, s>>>(n, a, x, b, y);
}

int reduce_parts(long n) {
  long g = grid_for(n);
  return (int)(g < 1024 ? (g < 1 ? 1 : g) : 1024);
}

void reduce(long n, const double* x, const double* y, int op, double* part, double* out, hipStream_t s) {
  int g = reduce_parts(n);
  reduce_k<<<g, kThreads, 0, s>>>(n, x, y, op, part);
  reduce_final_k<<<1, kThreads, 0, s>>>(g, part, op, out);
}

template <int B, int G, typename T, typename V>
static void spmv_rows_t(const SpmvPlan& pl, const T* val, const V* x, V* y, const V* w1, const V* w2, double* part,
                        BicgScalars* sc, hipStream_t s) {
  if constexpr (G % B == 0) {
    const unsigned grid = (unsigned)pl.row_wgs;  // a multiple of 8
    spmv_rows_k<B, G, T, V><<<grid, kThreads, 0, s>>>(pl.nrows, pl.rows_per_wg, pl.rp, pl.col, val, x, y, w1, w2, part,
                                                      sc);
  }
}

template <int B, typename T, typename V>
static void spmv_t(const SpmvPlan& pl, const T* val, const V* x, V* y, const V* w1, const V* w2, double* part,
                   BicgScalars* sc, int stage, hipStream_t s) {
  (void)stage;
  switch (pl.variant) {
    case 5: {
      const unsigned grid = (unsigned)(((pl.nchunk + 7) / 8) * 8);
      spmv_blk_k<B, T, V><<<grid, kThreads, 0, s>>>(pl.nchunk, pl.chunk_row, pl.rp, pl.col, val, x, y, w1, w2, part,
                                                     sc);
      return;
    }
    case 1: spmv_rows_t<B, 8, T, V>(pl, val, x, y, w1, w2, part, sc, s); return;
    case 2: spmv_rows_t<B, 16, T, V>(pl, val, x, y, w1, w2, part, sc, s); return;
    case 3: spmv_rows_t<B, 32, T, V>(pl, val, x, y, w1, w2, part, sc, s); return;
    case 4: spmv_rows_t<B, 64, T, V>(pl, val, x, y, w1, w2, part, sc, s); return;
    default: break;
  }
  // grid rounded up to a multiple of 8 so the XCD remap covers every chunk
  const unsigned grid = (unsigned)(((pl.nchunk + 7) / 8) * 8);
  spmv_k<B, T, V><<<grid, kThreads, 0, s>>>(pl.nchunk, pl.chunk_row, pl.rp, pl.col, val, x, y, w1, w2, part, sc, 0);
}

#define DCG_DISPATCH_B(B, CALL)                     \
  switch (B) {                                      \
    case 1: { constexpr int BB = 1; CALL; break; }  \
    case 2: { constexpr int BB = 2; CALL; break; }  \
    case 3: { constexpr int BB = 3; CALL; break; }  \
    case 4: { constexpr int BB = 4; CALL; break; }  \
    case 5: { constexpr int BB = 5; CALL; break; }  \
    case 6: { constexpr int BB = 6; CALL; break; }  \
    case 7: { constexpr int BB = 7; CALL; break; }  \
    case 8: { constexpr int BB = 8; CALL; break; }  \
    default: break;                                 \
  }

void spmv(int B, const SpmvPlan& pl, const double* val, const double* x, double* y, const double* w1,
          const double* w2, double* part, BicgScalars* sc, int stage, hipStream_t s) {
  DCG_DISPATCH_B(B, (spmv_t<BB, double, double>(pl, val, x, y, w1, w2, part, sc, stage, s)));
}
void spmv(int B, const SpmvPlan& pl, const float* val, const float* x, float* y, const float* w1, const float* w2,
          double* part, BicgScalars* sc, int stage, hipStream_t s) {
  DCG_DISPATCH_B(B, (spmv_t<BB, float, float>(pl, val, x, y, w1, w2, part, sc, stage, s)));
}

void dot_parts(long n, const double* y, const double* w1, const double* w2, double* part, int groups,
               const BicgScalars* sc, hipStream_t s) {
  dot_parts_k<double><<<groups, kThreads, 0, s>>>(n, y, w1, w2, part, sc);
}
void dot_parts(long n, const float* y, const float* w1, const float* w2, double* part, int groups,
               const BicgScalars* sc, hipStream_t s) {
  dot_parts_k<float><<<groups, kThreads, 0, s>>>(n, y, w1, w2, part, sc);
}

void block_inverse(int B, int nslot, const double* val, const int* diag_pos, double* dinv, int* status, bool point,
                   hipStream_t s) {
  DCG_DISPATCH_B(B, (block_inverse_k<BB><<<grid_for(nslot), kThreads, 0, s>>>(nslot, val, diag_pos, dinv, status,
                                                                              point ? 1 : 0)));
}

template <typename V>
void bicg_init(long n, const double* b, V* x, V* r, V* rh, V* p, V* v, double* part, int nparts, BicgScalars* sc,
               hipStream_t s) {
  bicg_init_k<V, double><<<nparts, kThreads, 0, s>>>(n, b, x, r, rh, p, v, part, sc);
  bicg_fin_k<<<1, 1024, 0, s>>>(nparts, part, 1, 0, sc);
}
template void bicg_init<double>(long, const double*, double*, double*, double*, double*, double*, double*, int,
                                BicgScalars*, hipStream_t);
template void bicg_init<float>(long, const double*, float*, float*, float*, float*, float*, double*, int,
                               BicgScalars*, hipStream_t);

template <typename V>
void bicg_p(int B, int nslot, const V* r, V* p, const V* v, V* ph, const double* dinv, bool diag, const BicgScalars* sc,
            hipStream_t s) {
  const int grid = (nslot + kThreads * kVecSlots - 1) / (kThreads * kVecSlots);
  if (grid > 0)
    DCG_DISPATCH_B(B, (bicg_p_k<BB, V><<<grid, kThreads, 0, s>>>(nslot, r, p, v, ph, dinv, diag ? 1 : 0, sc)));
}
template void bicg_p<double>(int, int, const double*, double*, const double*, double*, const double*, bool,
                             const BicgScalars*, hipStream_t);
template void bicg_p<float>(int, int, const float*, float*, const float*, float*, const double*, bool,
                            const BicgScalars*, hipStream_t);

template <typename V>
void bicg_s(int B, int nslot, const V* r, const V* v, V* sv, V* sh, const double* dinv, bool diag, const BicgScalars* sc,
            double* part, hipStream_t s) {
  const int grid = bicg_s_parts(nslot);
  if (grid > 0)
    DCG_DISPATCH_B(B, (bicg_s_k<BB, V><<<grid, kThreads, 0, s>>>(nslot, r, v, sv, sh, dinv, diag ? 1 : 0, sc, part)));
}
int bicg_s_parts(int nslot) { return (nslot + kThreads * kVecSlots - 1) / (kThreads * kVecSlots); }
template void bicg_s<double>(int, int, const double*, const double*, double*, double*, const double*, bool,
                             const BicgScalars*, double*, hipStream_t);
template void bicg_s<float>(int, int, const float*, const float*, float*, float*, const double*, bool,
                            const BicgScalars*, double*, hipStream_t);

template <typename V>
int bicg_xr(int B, int nslot, V* x, const V* ph, const V* sh, const V* sv, const V* t, V* r, const V* rh,
            const BicgScalars* sc, double* part, hipStream_t s) {
  const int grid = (nslot + kThreads * kXrSlots - 1) / (kThreads * kXrSlots);
  if (grid > 0)
    DCG_DISPATCH_B(B, (bicg_xr_k<BB, V><<<grid, kThreads, 0, s>>>(nslot, x, ph, sh, sv, t, r, rh, sc, part)));
  return grid;
}
template int bicg_xr<double>(int, int, double*, const double*, const double*, const double*, const double*, double*,
                             const double*, const BicgScalars*, double*, hipStream_t);
template int bicg_xr<float>(int, int, float*, const float*, const float*, const float*, const float*, float*,
                            const float*, const BicgScalars*, double*, hipStream_t);

void bicg_fin(int nparts, const double* part, int stride, int stage, BicgScalars* sc, hipStream_t s) {
  bicg_fin_k<<<1, 1024, 0, s>>>(nparts, part, stride, stage, sc);
}

void bicg_fin23(int na, const double* pa, int nb, const double* pb, BicgScalars* sc, hipStream_t s) {
  bicg_fin23_k<<<1, 1024, 0, s>>>(na, pa, nb, pb, sc);
}

void ilu_factor(int B, int nrows, const int* rows, const int* rp, const int* col, const int* diag_pos, double* lu,
                double* dinv, int* status, hipStream_t s) {
  if (nrows <= 0) return;
  DCG_DISPATCH_B(B, (ilu_factor_k<BB><<<grid_for(nrows, 64), 64, 0, s>>>(nrows, rows, rp, col, diag_pos, lu, dinv, status)));
}

void ilu_factor_sweep(int B, const IluSweep& t, const int* rp, const int* col, const int* diag_pos, double* lu,
                      double* dinv, int* status, hipStream_t s) {
  if (t.nlo <= 0) return;
  if (t.inbox_path && B <= 2) {
    const long n = (long)t.nrows * kIluRegRow * B * B;
    if (B == 1) {
      ilu_pack_k<1><<<grid_for(n, kThreads), kThreads, 0, s>>>(t.nrows, t, lu, t.packed);
      ilu_factor_inbox_k<1><<<1, t.factor_threads, (size_t)t.factor_threads * kIluRegRow * sizeof(double), s>>>(
          t, t.packed, t.inbox, lu, dinv, status);
    } else {
      ilu_pack_k<2><<<grid_for(n, kThreads), kThreads, 0, s>>>(t.nrows, t, lu, t.packed);
      ilu_factor_inbox_k<2><<<1, t.factor_threads, (size_t)t.factor_threads * kIluRegRow * 4 * sizeof(double), s>>>(
          t, t.packed, t.inbox, lu, dinv, status);
    }
    return;
  }
  DCG_DISPATCH_B(B, (ilu_factor_sweep_k<BB><<<1, 256, 0, s>>>(t.nlo, t.lo_lev, t.lo_rows, rp, col, diag_pos, lu,
                                                              dinv, status)));
}

// the chunked apply with `lds` bytes of dynamic LDS; above 64 KiB the
// kernel's limit is raised first (once per instantiation and size)
template <int BB, typename V, int K>
void launch_ilu_chunk_k(size_t lds, const IluSweep& t, const double* lu, const double* dinv, const V* in, V* tmp,
                        V* out, const BicgScalars* sc, hipStream_t s) {
  static std::atomic<size_t> raised{64 * 1024};
  if (lds > raised.load()) {
    DCG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&ilu_apply_chunk_k<BB, V, K>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    size_t cur = raised.load();
    while (lds > cur && !raised.compare_exchange_weak(cur, lds)) {
    }
  }
  ilu_apply_chunk_k<BB, V, K><<<1, kIluThreads, lds, s>>>(t, lu, dinv, in, tmp, out, sc);
}
template <int BB, typename V>
void launch_ilu_chunk(size_t lds, const IluSweep& t, const double* lu, const double* dinv, const V* in, V* tmp, V* out,
                      const BicgScalars* sc, hipStream_t s) {
  // the register path where the host planned one (K * B * B <= 32)
  if constexpr (BB <= 2) {
    if (t.reg_k == 4) return launch_ilu_chunk_k<BB, V, 4>(lds, t, lu, dinv, in, tmp, out, sc, s);
    if (t.reg_k == 8) return launch_ilu_chunk_k<BB, V, 8>(lds, t, lu, dinv, in, tmp, out, sc, s);
  }
  if constexpr (BB == 1)
    if (t.reg_k == 16) return launch_ilu_chunk_k<BB, V, 16>(lds, t, lu, dinv, in, tmp, out, sc, s);
  launch_ilu_chunk_k<BB, V, 0>(lds, t, lu, dinv, in, tmp, out, sc, s);
}

template <typename V>
void ilu_apply_sweep(int B, const IluSweep& t, const int* rp, const int* col, const int* diag_pos, const double* lu,
                     const double* dinv, const V* in, V* tmp, V* out, const BicgScalars* sc, hipStream_t s) {
  if (t.chunked) {
    const size_t lds = IluChunkLayout(t, B, static_cast<int>(sizeof(V))).total;
    DCG_DISPATCH_B(B, (launch_ilu_chunk<BB, V>(lds, t, lu, dinv, in, tmp, out, sc, s)));
    return;
  }
  DCG_DISPATCH_B(B, (ilu_apply_sweep_k<BB, V><<<1, 256, 0, s>>>(t.nlo, t.lo_lev, t.lo_rows, t.nup, t.up_lev,
                                                                t.up_rows, rp, col, diag_pos, lu, dinv, in, tmp, out,
                                                                sc)));
}
size_t ilu_chunk_lds_bytes(const IluSweep& t, int B, int vbytes) { return IluChunkLayout(t, B, vbytes).total; }

template void ilu_apply_sweep<double>(int, const IluSweep&, const int*, const int*, const int*, const double*,
                                      const double*, const double*, double*, double*, const BicgScalars*, hipStream_t);
template void ilu_apply_sweep<float>(int, const IluSweep&, const int*, const int*, const int*, const double*,
                                     const double*, const float*, float*, float*, const BicgScalars*, hipStream_t);

template <typename V>
void ilu_lower(int B, int nrows, const int* rows, const int* rp, const int* col, const int* diag_pos,
               const double* lu, const V* b, V* y, const BicgScalars* sc, hipStream_t s) {
  if (nrows <= 0) return;
  DCG_DISPATCH_B(B, (ilu_lower_k<BB, V><<<grid_for(nrows, 64), 64, 0, s>>>(nrows, rows, rp, col, diag_pos, lu, b, y, sc)));
}
template void ilu_lower<double>(int, int, const int*, const int*, const int*, const int*, const double*,
                                const double*, double*, const BicgScalars*, hipStream_t);
template void ilu_lower<float>(int, int, const int*, const int*, const int*, const int*, const double*, const float*,
                               float*, const BicgScalars*, hipStream_t);

template <typename V>
void ilu_upper(int B, int nrows, const int* rows, const int* rp, const int* col, const int* diag_pos,
               const double* lu, const double* dinv, const V* y, V* x, const BicgScalars* sc, hipStream_t s) {
  if (nrows <= 0) return;
  DCG_DISPATCH_B(B, (ilu_upper_k<BB, V><<<grid_for(nrows, 64), 64, 0, s>>>(nrows, rows, rp, col, diag_pos, lu, dinv, y, x, sc)));
}
template void ilu_upper<double>(int, int, const int*, const int*, const int*, const int*, const double*,
                                const double*, const double*, double*, const BicgScalars*, hipStream_t);
template void ilu_upper<float>(int, int, const int*, const int*, const int*, const int*, const double*,
                               const double*, const float*, float*, const BicgScalars*, hipStream_t);

void unpack_state(long nslot, int B, const int* slot_b, const long* slot_dof, const double* packed, double* padded,
                  hipStream_t s) {
  unpack_k<<<grid_for(nslot * B), kThreads, 0, s>>>(nslot, B, slot_b, slot_dof, packed, padded);
}
void pack_state(long nslot, int B, const int* slot_b, const long* slot_dof, const double* padded, double* packed,
                hipStream_t s) {
  pack_k<<<grid_for(nslot * B), kThreads, 0, s>>>(nslot, B, slot_b, slot_dof, padded, packed);
}

void to_float(long n, const double* a, float* b, hipStream_t s) {
  if (n > 0) convert_k<double, float><<<grid_for(n), kThreads, 0, s>>>(n, a, b);
}
void to_double(long n, const float* a, double* b, hipStream_t s) {
  if (n > 0) convert_k<float, double><<<grid_for(n), kThreads, 0, s>>>(n, a, b);
}

void weighted_species_sum(long first, long n, int B, const double* w, const double* x, double* part, double* out,
                          hipStream_t s) {
  int g = reduce_parts(n);
  weighted_sum_k<<<g, kThreads, 0, s>>>(first, n, B, w, x, part);
  weighted_sum_final_k<<<1, kThreads, 0, s>>>(g, part, out);
}

}  // namespace dcopasi::k
