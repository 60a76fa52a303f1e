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

#include "kernels.hh"
#include "common.hh"

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
  jacobi_inverse_dev<B>(val + (size_t)diag_pos[s] * B * B, dinv + (size_t)s * B * B, status, point != 0);
}

// ----------------------------------------------------------------- BiCGSTAB
// Scalars live on the device (BicgScalars); every kernel returns at once when
// the solve has finished, so the host enqueues iterations without waiting.
__device__ __forceinline__ bool finished(const BicgScalars* sc) { return sc->done != 0; }

// p = r + beta (p - omega v); ph = M^-1 p (Jacobi fused, else copy)
template <int B, typename V>
__global__ void bicg_p_k(int nslot, const V* __restrict__ r, V* __restrict__ p, const V* __restrict__ v,
                         V* __restrict__ ph, const double* __restrict__ dinv, const BicgScalars* __restrict__ sc) {
  if (finished(sc)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  const double beta = sc->beta, omega = sc->omega;
  V pv[B];
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const size_t k = (size_t)s * B + i;
    const double nv = (double)r[k] + beta * ((double)p[k] - omega * (double)v[k]);
    pv[i] = (V)nv;
    p[k] = pv[i];
  }
  if (dinv) dinv_apply<B>(dinv, s, pv, ph + (size_t)s * B);
  else
#pragma unroll
    for (int i = 0; i < B; ++i) ph[(size_t)s * B + i] = pv[i];
}

// s = r - alpha v; sh = M^-1 s; partial ||s||^2
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) bicg_s_k(int nslot, const V* __restrict__ r, const V* __restrict__ v,
                                                     V* __restrict__ sv, V* __restrict__ sh,
                                                     const double* __restrict__ dinv,
                                                     const BicgScalars* __restrict__ sc, double* __restrict__ part) {
  if (finished(sc)) return;
  __shared__ double red[16];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (s < nslot) {
    const double alpha = sc->alpha;
    V t[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const size_t k = (size_t)s * B + i;
      t[i] = (V)((double)r[k] - alpha * (double)v[k]);
      sv[k] = t[i];
      acc[0] += (double)t[i] * (double)t[i];
    }
    if (dinv) dinv_apply<B>(dinv, s, t, sh + (size_t)s * B);
    else
#pragma unroll
      for (int i = 0; i < B; ++i) sh[(size_t)s * B + i] = t[i];
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

// x += alpha ph + omega sh ; r = s - omega t ; partial (rh.r), (r.r)
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) bicg_xr_k(int nslot, V* __restrict__ x, const V* __restrict__ ph,
                                                      const V* __restrict__ sh, const V* __restrict__ sv,
                                                      const V* __restrict__ t, V* __restrict__ r,
                                                      const V* __restrict__ rh, const BicgScalars* __restrict__ sc,
                                                      double* __restrict__ part) {
  if (finished(sc)) return;
  __shared__ double red[32];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  const bool half = sc->half != 0;
  if (s < nslot) {
    const double alpha = sc->alpha, omega = sc->omega;
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const size_t k = (size_t)s * B + i;
      if (half) {
        x[k] = (V)((double)x[k] + alpha * (double)ph[k]);
      } else {
        x[k] = (V)((double)x[k] + alpha * (double)ph[k] + omega * (double)sh[k]);
        const V rv = (V)((double)sv[k] - omega * (double)t[k]);
        r[k] = rv;
        acc[0] += (double)rh[k] * (double)rv;
        acc[1] += (double)rv * (double)rv;
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
  // invert the (updated) diagonal block
  double a[B][B], inv[B][B];
  const double* d = lu + (size_t)diag_pos[i] * B * B;
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
  double* o = dinv + (size_t)i * B * B;
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int c = 0; c < B; ++c) o[r * B + c] = inv[r][c];
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
void bicg_p(int B, int nslot, const V* r, V* p, const V* v, V* ph, const double* dinv, const BicgScalars* sc,
            hipStream_t s) {
  DCG_DISPATCH_B(B, (bicg_p_k<BB, V><<<grid_for(nslot), kThreads, 0, s>>>(nslot, r, p, v, ph, dinv, sc)));
}
template void bicg_p<double>(int, int, const double*, double*, const double*, double*, const double*,
                             const BicgScalars*, hipStream_t);
template void bicg_p<float>(int, int, const float*, float*, const float*, float*, const double*, const BicgScalars*,
                            hipStream_t);

template <typename V>
void bicg_s(int B, int nslot, const V* r, const V* v, V* sv, V* sh, const double* dinv, const BicgScalars* sc,
            double* part, hipStream_t s) {
  DCG_DISPATCH_B(B, (bicg_s_k<BB, V><<<grid_for(nslot), kThreads, 0, s>>>(nslot, r, v, sv, sh, dinv, sc, part)));
}
template void bicg_s<double>(int, int, const double*, const double*, double*, double*, const double*,
                             const BicgScalars*, double*, hipStream_t);
template void bicg_s<float>(int, int, const float*, const float*, float*, float*, const double*, const BicgScalars*,
                            double*, hipStream_t);

template <typename V>
void bicg_xr(int B, int nslot, V* x, const V* ph, const V* sh, const V* sv, const V* t, V* r, const V* rh,
             const BicgScalars* sc, double* part, hipStream_t s) {
  DCG_DISPATCH_B(B, (bicg_xr_k<BB, V><<<grid_for(nslot), kThreads, 0, s>>>(nslot, x, ph, sh, sv, t, r, rh, sc, part)));
}
template void bicg_xr<double>(int, int, double*, const double*, const double*, const double*, const double*,
                              double*, const double*, const BicgScalars*, double*, hipStream_t);
template void bicg_xr<float>(int, int, float*, const float*, const float*, const float*, const float*, float*,
                             const float*, const BicgScalars*, double*, hipStream_t);

void bicg_fin(int nparts, const double* part, int stride, int stage, BicgScalars* sc, hipStream_t s) {
  bicg_fin_k<<<1, 1024, 0, s>>>(nparts, part, stride, stage, sc);
}

void ilu_factor(int B, int nrows, const int* rows, const int* rp, const int* col, const int* diag_pos, double* lu,
                double* dinv, int* status, hipStream_t s) {
  if (nrows <= 0) return;
  DCG_DISPATCH_B(B, (ilu_factor_k<BB><<<grid_for(nrows, 64), 64, 0, s>>>(nrows, rows, rp, col, diag_pos, lu, dinv, status)));
}

void ilu_factor_sweep(int B, int nlev, const int* lev, const int* rows, const int* rp, const int* col,
                      const int* diag_pos, double* lu, double* dinv, int* status, hipStream_t s) {
  if (nlev <= 0) return;
  DCG_DISPATCH_B(B, (ilu_factor_sweep_k<BB><<<1, 256, 0, s>>>(nlev, lev, rows, rp, col, diag_pos, lu, dinv, status)));
}

template <typename V>
void ilu_apply_sweep(int B, int nlo, const int* lo_lev, const int* lo_rows, int nup, const int* up_lev,
                     const int* up_rows, const int* rp, const int* col, const int* diag_pos, const double* lu,
                     const double* dinv, const V* in, V* tmp, V* out, const BicgScalars* sc, hipStream_t s) {
  DCG_DISPATCH_B(B, (ilu_apply_sweep_k<BB, V><<<1, 256, 0, s>>>(nlo, lo_lev, lo_rows, nup, up_lev, up_rows, rp, col,
                                                                diag_pos, lu, dinv, in, tmp, out, sc)));
}
template void ilu_apply_sweep<double>(int, int, const int*, const int*, int, const int*, const int*, const int*,
                                      const int*, const int*, const double*, const double*, const double*, double*,
                                      double*, const BicgScalars*, hipStream_t);
template void ilu_apply_sweep<float>(int, int, const int*, const int*, int, const int*, const int*, const int*,
                                     const int*, const int*, const double*, const double*, const float*, float*,
                                     float*, const BicgScalars*, hipStream_t);

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
