// Device helpers shared by the solver and batch kernels: wavefront and
// workgroup reductions, the XCD-aware chunk map, the BiCGSTAB scalar update,
// and B x B block inversion / application.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hh"

namespace dcopasi::k {

constexpr int kThreads = 256;
constexpr int kWave = 64;

// ----------------------------------------------------------------- helpers
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// sum over the workgroup, result valid in thread 0
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* lds) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) lds[wid * N + i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += lds[w * N + i];
      v[i] = s;
    }
  }
}

// XCD-aware chunk mapping: hardware dispatches consecutive workgroups round
// robin over the 8 XCDs; give each XCD one contiguous run of chunks so its
// L2 sees neighbouring rows (and their shared x entries).
__device__ __forceinline__ int xcd_chunk(int bid, int nblocks) {
  const int per = (nblocks + 7) / 8;
  const int c = (bid % 8) * per + bid / 8;
  return c;
}

// ----------------------------------------------------------------- BiCGSTAB scalars
// Updates the device-resident scalars after one reduction stage; v0/v1 are the
// two reduced sums of that stage.
__device__ void bicg_update(BicgScalars* sc, int stage, double v0, double v1) {
  switch (stage) {
    case 0: {  // initial: v0 = r.r ; rho = r.r since rh = r
      sc->r0 = sqrt(v0);
      sc->rnorm = sc->r0;
      sc->rho = v0;
      sc->alpha = sc->omega = 1.0;
      sc->beta = 0.0;
      sc->iter = 0;
      sc->half = 0;
      if (!(sc->r0 > 0.0)) sc->done = isfinite(sc->r0) ? 1 : 3;
      if (sc->r0 <= sc->abstol) sc->done = 1;
      break;
    }
    case 1: {  // after v = A ph: alpha = rho / (rh.v)
      if (v0 == 0.0 || !isfinite(v0)) {
        sc->done = isfinite(v0) ? 4 : 3;
        break;
      }
      sc->alpha = sc->rho / v0;
      break;
    }
    case 2: {  // after s: half-step convergence test
      const double ns = sqrt(v0);
      sc->iter += 1;
      if (!isfinite(ns)) {
        sc->done = 3;
        break;
      }
      if (ns <= sc->tol * sc->r0 || ns <= sc->abstol) {
        sc->half = 1;
        sc->rnorm = ns;
      }
      break;
    }
    case 3: {  // after t = A sh: omega = (t.s)/(t.t)
      if (sc->half) break;
      sc->omega = v1 == 0.0 ? 0.0 : v0 / v1;
      break;
    }
    case 4: {  // after x, r update: convergence, beta for the next step
      if (sc->half) {
        sc->done = 1;
        break;
      }
      const double nr = sqrt(v1);
      sc->rnorm = nr;
      if (!isfinite(nr)) {
        sc->done = 3;
        break;
      }
      if (nr <= sc->tol * sc->r0 || nr <= sc->abstol) {
        sc->done = 1;
        break;
      }
      if (sc->omega == 0.0 || v0 == 0.0) {
        sc->done = 4;
        break;
      }
      sc->beta = (v0 / sc->rho) * (sc->alpha / sc->omega);
      sc->rho = v0;
      if (sc->iter >= sc->maxit) sc->done = 2;
      break;
    }
  }
}

// out = M^-1 in at slot s: a diagonal inverse (diag: B values per slot, the
// reference's point Jacobi) or a B x B block per slot
template <int B, typename V>
__device__ __forceinline__ void dinv_apply(const double* __restrict__ dinv, int s, const V* in, V* out, bool diag) {
  if (diag) {
#pragma unroll
    for (int i = 0; i < B; ++i) out[i] = (V)(dinv[(size_t)s * B + i] * (double)in[i]);
    return;
  }
  const double* d = dinv + (size_t)s * B * B;
  double t[B];
#pragma unroll
  for (int i = 0; i < B; ++i) t[i] = (double)in[i];
#pragma unroll
  for (int i = 0; i < B; ++i) {
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < B; ++j) v += d[i * B + j] * t[j];
    out[i] = (V)v;
  }
}

// inverse of one B x B block d into o (Gauss-Jordan with partial pivoting);
// a singular block sets status[0] = 4
template <int B>
__device__ __forceinline__ void block_inverse_dev(const double* __restrict__ d, double* __restrict__ o,
                                                  int* __restrict__ status) {
  double a[B][B], inv[B][B];
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) {
      a[i][j] = d[i * B + j];
      inv[i][j] = i == j ? 1.0 : 0.0;
    }
  // Gauss-Jordan with partial pivoting
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
      best = 1.0;
      a[piv][c] = 1.0;
    }
    if (piv != c)
      for (int j = 0; j < B; ++j) {
        double t = a[c][j];
        a[c][j] = a[piv][j];
        a[piv][j] = t;
        t = inv[c][j];
        inv[c][j] = inv[piv][j];
        inv[piv][j] = t;
      }
    const double p = 1.0 / a[c][c];
    for (int j = 0; j < B; ++j) {
      a[c][j] *= p;
      inv[c][j] *= p;
    }
    for (int r = 0; r < B; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      if (f == 0.0) continue;
      for (int j = 0; j < B; ++j) {
        a[r][j] -= f * a[c][j];
        inv[r][j] -= f * inv[c][j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) o[i * B + j] = inv[i][j];
}

// point Jacobi in the same B x B layout: 1/a_ii on the diagonal, zeros
// elsewhere, so every block-Jacobi application applies it unchanged; a zero
// diagonal entry sets status[0] = 4
template <int B>
__device__ __forceinline__ void diag_inverse_dev(const double* __restrict__ d, double* __restrict__ o,
                                                 int* __restrict__ status) {
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const double a = d[i * B + i];
    if (a == 0.0) status[0] = 4;
    o[i] = a == 0.0 ? 1.0 : 1.0 / a;
  }
}

// the Jacobi inverse of slot s's diagonal block d into dinv: B values per
// slot (point: the inverse of the scalar diagonal) or a B x B block
template <int B>
__device__ __forceinline__ void jacobi_inverse_dev(const double* __restrict__ d, double* __restrict__ dinv, int s,
                                                   int* __restrict__ status, bool point) {
  if (point) diag_inverse_dev<B>(d, dinv + (size_t)s * B, status);
  else block_inverse_dev<B>(d, dinv + (size_t)s * B * B, status);
}


// ----------------------------------------------------------------- SpMV chunk bodies
// LDS stage element: the FP32 matrix stages its per-block products in FP32
// (half the LDS, twice the resident workgroups), rows are summed in FP64.
template <typename T>
using StageT = typename std::conditional<std::is_same<T, float>::value, float, double>::type;

// One workgroup computes rows [r0, r1) of y = A x, whose blocks are
// [k0, k1) = [rp[r0], rp[r1]); the chunk's block-row entries (blocks x B)
// fit the LDS stage.  Phase 1 streams the chunk's blocks
// with consecutive threads on consecutive (block, row-of-block) pairs, so
// value loads are fully coalesced; products go to LDS.  Phase 2 sums each row
// from LDS and adds y.w1, y.w2 to acc.  IP, MP, XP, YP: the pointer types of
// the pattern, the values, the vectors read and y (plain pointers, or
// global-memory ones where they come from a batch member's descriptor).
template <int B, typename T, typename V, typename S, typename IP, typename MP, typename XP, typename YP>
__device__ __forceinline__ void spmv_entries_chunk(int r0, int r1, int k0, int k1, IP __restrict__ rp,
                                                   IP __restrict__ col, MP __restrict__ val, XP __restrict__ x,
                                                   YP __restrict__ y, XP __restrict__ w1, XP __restrict__ w2,
                                                   double (&acc)[2], S* __restrict__ staged) {
  const int ne = (k1 - k0) * B;
  constexpr int U = 4;  // independent entries per thread in flight
  for (int e0 = threadIdx.x; e0 < ne; e0 += U * kThreads) {
    T av[U][B];
    int cv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * kThreads;
      const int kb = k0 + (e < ne ? e : 0) / B, i = (e < ne ? e : 0) % B;
      const auto a = val + (size_t)kb * (B * B) + i * B;
#pragma unroll
      for (int j = 0; j < B; ++j) av[u][j] = __builtin_nontemporal_load(a + j);
      cv[u] = col[kb];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * kThreads;
      if (e >= ne) break;
      const auto xv = x + (size_t)cv[u] * B;
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < B; ++j) s += (double)av[u][j] * (double)xv[j];
      staged[e] = (S)s;
    }
  }
  __syncthreads();
  const int nout = (r1 - r0) * B;
  for (int o = threadIdx.x; o < nout; o += blockDim.x) {
    const int row = r0 + o / B, i = o % B;
    double s = 0.0;
    for (int kb = rp[row]; kb < rp[row + 1]; ++kb) s += (double)staged[(kb - k0) * B + i];
    const size_t yi = (size_t)row * B + i;
    y[yi] = (V)s;
    const double ys = (double)(V)s;
    if (w1) acc[0] += ys * (double)w1[yi];
    if (w2) acc[1] += ys * (double)w2[yi];
  }
}

// Block form: one thread per whole B x B block (B*B contiguous values, one
// column index, one B-vector of x), writing B products to the stage; halves
// the per-entry index and gather work for B = 2 and makes FP32 loads 16 bytes.
template <int B, typename T, typename V, typename S, typename IP, typename MP, typename XP, typename YP>
__device__ __forceinline__ void spmv_blocks_chunk(int r0, int r1, int k0, int k1, IP __restrict__ rp,
                                                  IP __restrict__ col, MP __restrict__ val, XP __restrict__ x,
                                                  YP __restrict__ y, XP __restrict__ w1, XP __restrict__ w2,
                                                  double (&acc)[2], S* __restrict__ staged) {
  const int nb = k1 - k0;
  constexpr int U = 2;
  for (int b0 = threadIdx.x; b0 < nb; b0 += U * kThreads) {
    T av[U][B * B];
    int cv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * kThreads;
      const int kb = k0 + (b < nb ? b : 0);
      const auto a = val + (size_t)kb * (B * B);
#pragma unroll
      for (int j = 0; j < B * B; ++j) av[u][j] = __builtin_nontemporal_load(a + j);
      cv[u] = col[kb];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * kThreads;
      if (b >= nb) break;
      const auto xv = x + (size_t)cv[u] * B;
      double xl[B];
#pragma unroll
      for (int j = 0; j < B; ++j) xl[j] = (double)xv[j];
#pragma unroll
      for (int i = 0; i < B; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < B; ++j) s += (double)av[u][i * B + j] * xl[j];
        staged[b * B + i] = (S)s;
      }
    }
  }
  __syncthreads();
  const int nout = (r1 - r0) * B;
  for (int o = threadIdx.x; o < nout; o += blockDim.x) {
    const int row = r0 + o / B, i = o % B;
    double s = 0.0;
    for (int kb = rp[row]; kb < rp[row + 1]; ++kb) s += (double)staged[(kb - k0) * B + i];
    const size_t yi = (size_t)row * B + i;
    y[yi] = (V)s;
    const double ys = (double)(V)s;
    if (w1) acc[0] += ys * (double)w1[yi];
    if (w2) acc[1] += ys * (double)w2[yi];
  }
}

}  // namespace dcopasi::k
