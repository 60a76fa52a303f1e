// Batched solver kernels on gfx950: one launch advances the same operation
// for every member of a batch of independent models.  A workgroup takes one
// item of a work list -- (member, 256-slot block) for vector kernels,
// (member, SpMV chunk) for the matrix product -- and finds that member's
// buffers in the MemberView array.  Every member keeps its own BiCGSTAB
// scalars, so members converge, stop and are skipped independently, and the
// reductions are segmented: the partial sums of a member occupy a contiguous
// range of the work list and one finaliser workgroup per member reduces them
// in a fixed order.  These kernels restate, per member, the BiCGSTAB and
// Jacobi iteration of solver.hip (dune-istl BiCGSTABSolver / SeqJac as
// configured by model/make_step_operator.hh:102-146).
#include <hip/hip_runtime.h>

#include "common.hh"
#include "kernels.hh"

namespace dcopasi::k {

namespace {

__device__ __forceinline__ bool skipped(const Skip& sk, int m) {
  return sk.p && ((sk.p[(size_t)m * sk.stride] != 0) == (sk.when != 0));
}

// A member's pointers come from its descriptor in memory, so the compiler
// cannot tell their address space and would access them with flat
// instructions; they all point into device global memory, and passing them
// through that address space lets it use global loads and stores (and pair
// adjacent ones into wide accesses).
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ global_ptr<T> in_global(T* p) {
  return (global_ptr<T>)p;
}

// vector selectors: FP64 vectors by BatchVec index, FP32 Krylov vectors at
// the same index as their FP64 twins (kVKry + i), the FP32 x at kVKry + 9
template <typename V>
__device__ __forceinline__ V* vsel(const MemberView& m, int code);
template <>
__device__ __forceinline__ double* vsel<double>(const MemberView& m, int code) {
  return code >= 0 ? m.v[code] : nullptr;
}
template <>
__device__ __forceinline__ float* vsel<float>(const MemberView& m, int code) {
  return code >= 0 ? m.f[code - kVKry] : nullptr;
}
template <typename T>
__device__ __forceinline__ const T* matrix(const MemberView& m);
template <>
__device__ __forceinline__ const double* matrix<double>(const MemberView& m) {
  return m.val;
}
template <>
__device__ __forceinline__ const float* matrix<float>(const MemberView& m) {
  return m.val32;
}

// ----------------------------------------------------------------- vector kernels
template <int B>
__global__ void __launch_bounds__(kThreads) bcopy_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                    Skip sk, int dst, int src) {
  const int2 w = work[blockIdx.x];
  if (skipped(sk, w.x)) return;
  const MemberView& m = mv[w.x];
  const int s = w.y * kThreads + threadIdx.x;
  if (s >= m.nslot) return;
  double* d = m.v[dst] + (size_t)s * B;
  const double* a = src >= 0 ? m.v[src] + (size_t)s * B : nullptr;
#pragma unroll
  for (int i = 0; i < B; ++i) d[i] = a ? a[i] : 0.0;
}

template <int B>
__global__ void __launch_bounds__(kThreads) bzerof_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                     Skip sk, int sel) {
  const int2 w = work[blockIdx.x];
  if (skipped(sk, w.x)) return;
  const MemberView& m = mv[w.x];
  const int s = w.y * kThreads + threadIdx.x;
  if (s >= m.nslot) return;
  float* d = vsel<float>(m, sel) + (size_t)s * B;
#pragma unroll
  for (int i = 0; i < B; ++i) d[i] = 0.0f;
}

template <int B>
__global__ void __launch_bounds__(kThreads) baxpy_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                    Skip sk, double a, int x, int y) {
  const int2 w = work[blockIdx.x];
  if (skipped(sk, w.x)) return;
  const MemberView& m = mv[w.x];
  const int s = w.y * kThreads + threadIdx.x;
  if (s >= m.nslot) return;
  const double* xv = m.v[x] + (size_t)s * B;
  double* yv = m.v[y] + (size_t)s * B;
#pragma unroll
  for (int i = 0; i < B; ++i) yv[i] += a * xv[i];
}

// per work item partial of x.x, sum |x| or max |x| (0 for skipped members)
template <int B>
__global__ void __launch_bounds__(kThreads) bnorm_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                    Skip sk, int x, int op, double* __restrict__ part) {
  __shared__ double red[16];
  const int2 w = work[blockIdx.x];
  double v[1] = {0.0};
  if (!skipped(sk, w.x)) {
    const MemberView& m = mv[w.x];
    const int s = w.y * kThreads + threadIdx.x;
    if (s < m.nslot) {
      const double* a = m.v[x] + (size_t)s * B;
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const double t = a[i];
        if (op == 0) v[0] += t * t;
        else if (op == 1) v[0] += fabs(t);
        else v[0] = fmax(v[0], fabs(t));
      }
    }
  }
  if (op == 2) {
    double mx = v[0];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, kWave));
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = 0.0;
      for (int i = 0; i < (int)(blockDim.x / kWave); ++i) r = fmax(r, red[i]);
      part[blockIdx.x] = r;
    }
    return;
  }
  block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// one workgroup per member: reduce its range of partials (sum or max)
__global__ void __launch_bounds__(kThreads) bnorm_fin_k(const int2* __restrict__ range,
                                                        const double* __restrict__ part, int op,
                                                        double* __restrict__ out) {
  __shared__ double red[16];
  const int2 r = range[blockIdx.x];
  double v[1] = {0.0};
  for (int i = r.x + threadIdx.x; i < r.y; i += blockDim.x) v[0] = op == 2 ? fmax(v[0], part[i]) : v[0] + part[i];
  if (op == 2) {
    double mx = v[0];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, kWave));
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      double q = 0.0;
      for (int i = 0; i < (int)(blockDim.x / kWave); ++i) q = fmax(q, red[i]);
      out[blockIdx.x] = q;
    }
    return;
  }
  block_sum(v, red);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

template <int B>
__global__ void __launch_bounds__(kThreads) binv_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                   Skip sk, int* __restrict__ status, int point) {
  const int2 w = work[blockIdx.x];
  if (skipped(sk, w.x)) return;
  const MemberView& m = mv[w.x];
  const int s = w.y * kThreads + threadIdx.x;
  if (s >= m.nslot) return;
  jacobi_inverse_dev<B>(m.val + (size_t)m.diag_pos[s] * B * B, m.dinv, s, status, point != 0);
}

// FP32 copy of the matrix rows of one 256-slot block
template <int B>
__global__ void __launch_bounds__(kThreads) btof_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                   Skip sk) {
  const int2 w = work[blockIdx.x];
  if (skipped(sk, w.x)) return;
  const MemberView& m = mv[w.x];
  if (m.npc) {
    // point storage: the member's coefficients in equal shares per block
    const long nblk = (m.nslot + kThreads - 1) / kThreads, share = (m.npc + nblk - 1) / nblk;
    const long v0 = (long)w.y * share, v1 = min(m.npc, v0 + share);
    for (long i = v0 + threadIdx.x; i < v1; i += kThreads) m.pc32[i] = (float)m.pc[i];
    return;
  }
  const int s0 = w.y * kThreads, s1 = min(m.nslot, s0 + kThreads);
  const long v0 = (long)m.rp[s0] * B * B, v1 = (long)m.rp[s1] * B * B;
  for (long i = v0 + threadIdx.x; i < v1; i += kThreads) m.val32[i] = (float)m.val[i];
}

// ----------------------------------------------------------------- BiCGSTAB
__global__ void bsetup_k(int members, Skip sk, const double* __restrict__ tol, const IrState* __restrict__ ir,
                         double inner, int maxit, BicgScalars* __restrict__ sc) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= members) return;
  BicgScalars s{};
  s.maxit = maxit;
  s.abstol = 0.0;
  bool off = skipped(sk, m);
  if (ir) {
    off = off || ir[m].done;
    s.tol = off ? 0.5 : fmin(0.5, fmax(inner, 0.5 * ir[m].target / ir[m].rn));
  } else {
    s.tol = tol[m];
  }
  s.done = off ? 1 : 0;
  sc[m] = s;
}

// x = 0, r = rh = b, p = v = 0, partial b.b
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) binit_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                    int b, int x, double* __restrict__ part,
                                                    const BicgScalars* __restrict__ sc) {
  __shared__ double red[16];
  const int2 w = work[blockIdx.x];
  double acc[1] = {0.0};
  if (!sc[w.x].done) {
    const MemberView& m = mv[w.x];
    const int s = w.y * kThreads + threadIdx.x;
    if (s < m.nslot) {
      const double* bv = m.v[b];
      V *xv = vsel<V>(m, x), *r = vsel<V>(m, kVKry + 0), *rh = vsel<V>(m, kVKry + 1), *p = vsel<V>(m, kVKry + 2),
        *v = vsel<V>(m, kVKry + 3);
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const size_t k = (size_t)s * B + i;
        const V bi = (V)bv[k];
        xv[k] = (V)0;
        r[k] = bi;
        rh[k] = bi;
        p[k] = (V)0;
        v[k] = (V)0;
        acc[0] += (double)bi * (double)bi;
      }
    }
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// The BiCGSTAB updates go over vwork4 items (member, block of kBVecSlots *
// kThreads slots): each thread takes kBVecSlots slots kThreads apart, so
// every pass of the workgroup is contiguous and more bytes are in flight.
// (Batch::rebuild sizes both work lists in blocks of 256 slots.)
static_assert(kThreads == 256, "batch work lists are built for 256-thread workgroups");

// p = r + beta (p - omega v); ph = M^-1 p
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) bp_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                 int jacobi, const BicgScalars* __restrict__ sc) {
  const int2 w = work[blockIdx.x];
  const BicgScalars& S = sc[w.x];
  if (S.done) return;
  const MemberView& m = mv[w.x];
  const auto r = in_global(vsel<V>(m, kVKry + 0));
  const auto p = in_global(vsel<V>(m, kVKry + 2));
  const auto v = in_global(vsel<V>(m, kVKry + 3));
  const auto ph = in_global(vsel<V>(m, kVKry + 4));
  const double beta = S.beta, omega = S.omega;
#pragma unroll
  for (int u = 0; u < kBVecSlots; ++u) {
    const int s = (w.y * kBVecSlots + u) * kThreads + threadIdx.x;
    if (s >= m.nslot) break;
    // each vector's B values read together (adjacent accesses pair up)
    const size_t k0 = (size_t)s * B;
    V rv[B], pold[B], vv[B], pv[B], o[B];
#pragma unroll
    for (int i = 0; i < B; ++i) rv[i] = r[k0 + i];
#pragma unroll
    for (int i = 0; i < B; ++i) pold[i] = p[k0 + i];
#pragma unroll
    for (int i = 0; i < B; ++i) vv[i] = v[k0 + i];
#pragma unroll
    for (int i = 0; i < B; ++i) pv[i] = (V)((double)rv[i] + beta * ((double)pold[i] - omega * (double)vv[i]));
#pragma unroll
    for (int i = 0; i < B; ++i) p[k0 + i] = pv[i];
    if (jacobi) dinv_apply<B>((const double*)in_global(m.dinv), s, pv, o, jacobi == 2);
    else
#pragma unroll
      for (int i = 0; i < B; ++i) o[i] = pv[i];
#pragma unroll
    for (int i = 0; i < B; ++i) ph[k0 + i] = o[i];
  }
}

// s = r - alpha v; sh = M^-1 s; partial ||s||^2
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) bs_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                 int jacobi, const BicgScalars* __restrict__ sc,
                                                 double* __restrict__ part) {
  __shared__ double red[16];
  const int2 w = work[blockIdx.x];
  const BicgScalars& S = sc[w.x];
  if (S.done) return;  // uniform over the workgroup
  const MemberView& m = mv[w.x];
  double acc[1] = {0.0};
  const auto r = in_global(vsel<V>(m, kVKry + 0));
  const auto v = in_global(vsel<V>(m, kVKry + 3));
  const auto sv = in_global(vsel<V>(m, kVKry + 5));
  const auto sh = in_global(vsel<V>(m, kVKry + 6));
  const double alpha = S.alpha;
#pragma unroll
  for (int u = 0; u < kBVecSlots; ++u) {
    const int s = (w.y * kBVecSlots + u) * kThreads + threadIdx.x;
    if (s >= m.nslot) break;
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
    if (jacobi) dinv_apply<B>((const double*)in_global(m.dinv), s, t, o, jacobi == 2);
    else
#pragma unroll
      for (int i = 0; i < B; ++i) o[i] = t[i];
#pragma unroll
    for (int i = 0; i < B; ++i) sh[k0 + i] = o[i];
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// x += alpha ph + omega sh; r = s - omega t; partial (rh.r), (r.r)
template <int B, typename V>
__global__ void __launch_bounds__(kThreads) bxr_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                  int x, const BicgScalars* __restrict__ sc,
                                                  double* __restrict__ part) {
  __shared__ double red[32];
  const int2 w = work[blockIdx.x];
  const BicgScalars& S = sc[w.x];
  if (S.done) return;
  const MemberView& m = mv[w.x];
  double acc[2] = {0.0, 0.0};
  const bool half = S.half != 0;
  const auto xv = in_global(vsel<V>(m, x));
  const auto r = in_global(vsel<V>(m, kVKry + 0));
  const auto rh = in_global(vsel<V>(m, kVKry + 1));
  const auto ph = in_global(vsel<V>(m, kVKry + 4));
  const auto sv = in_global(vsel<V>(m, kVKry + 5));
  const auto sh = in_global(vsel<V>(m, kVKry + 6));
  const auto t = in_global(vsel<V>(m, kVKry + 7));
  const double alpha = S.alpha, omega = S.omega;
#pragma unroll
  for (int u = 0; u < kBVecSlots; ++u) {
    const int s = (w.y * kBVecSlots + u) * kThreads + threadIdx.x;
    if (s >= m.nslot) break;
    const size_t k0 = (size_t)s * B;
    V x0[B], pv[B];
#pragma unroll
    for (int i = 0; i < B; ++i) x0[i] = xv[k0 + i];
#pragma unroll
    for (int i = 0; i < B; ++i) pv[i] = ph[k0 + i];
    if (half) {
#pragma unroll
      for (int i = 0; i < B; ++i) xv[k0 + i] = (V)((double)x0[i] + alpha * (double)pv[i]);
    } else {
      V hv[B], sw[B], tv[B], r0[B], xn[B], rn[B];
#pragma unroll
      for (int i = 0; i < B; ++i) hv[i] = sh[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) sw[i] = sv[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) tv[i] = t[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) r0[i] = rh[k0 + i];
#pragma unroll
      for (int i = 0; i < B; ++i) {
        xn[i] = (V)((double)x0[i] + alpha * (double)pv[i] + omega * (double)hv[i]);
        rn[i] = (V)((double)sw[i] - omega * (double)tv[i]);
        acc[0] += (double)r0[i] * (double)rn[i];
        acc[1] += (double)rn[i] * (double)rn[i];
      }
#pragma unroll
      for (int i = 0; i < B; ++i) xv[k0 + i] = xn[i];
#pragma unroll
      for (int i = 0; i < B; ++i) r[k0 + i] = rn[i];
    }
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// y = A x per (member, chunk); partial dots y.w1, y.w2 per work item.  The
// work index goes through the XCD map like the single-model SpMV.
template <int B, typename T, typename V, bool BLOCKS>
__global__ void __launch_bounds__(kThreads) bspmv_k(int nwork, const int4* __restrict__ cwork,
                                                    const MemberView* __restrict__ mv, Skip sk,
                                                    const BicgScalars* __restrict__ sc, int x, int y, int w1,
                                                    int w2, double* __restrict__ part) {
  __shared__ StageT<T> staged[kSpmvStage];
  __shared__ double red[32];
  const int c = xcd_chunk(blockIdx.x, gridDim.x);
  if (c >= nwork) return;
  const int4 w = cwork[c];
  if (skipped(sk, w.x) || (sc && sc[w.x].done)) return;
  const MemberView& m = mv[w.x];
  double acc[2] = {0.0, 0.0};
  const int r0 = w.y, r1 = w.y + (w.w >> 16), k0 = w.z, k1 = w.z + (w.w & 0xffff);
  const auto rp = in_global(m.rp), col = in_global(m.col);
  const auto val = in_global(matrix<T>(m));
  const auto xv = in_global((const V*)vsel<V>(m, x));
  const auto yv = in_global(vsel<V>(m, y));
  const auto w1v = in_global((const V*)vsel<V>(m, w1));
  const auto w2v = in_global((const V*)vsel<V>(m, w2));
  if (BLOCKS)
    spmv_blocks_chunk<B, T, V, StageT<T>>(r0, r1, k0, k1, rp, col, val, xv, yv, w1v, w2v, acc, staged);
  else
    spmv_entries_chunk<B, T, V, StageT<T>>(r0, r1, k0, k1, rp, col, val, xv, yv, w1v, w2v, acc, staged);
  if (part) {
    block_sum(acc, red);
    if (threadIdx.x == 0) {
      part[2 * c] = acc[0];
      part[2 * c + 1] = acc[1];
    }
  }
}

// one workgroup per member: reduce its partials and update its scalars
__global__ void __launch_bounds__(kThreads) bfin_k(const int2* __restrict__ range, const double* __restrict__ part,
                                                   int stride, int stage, BicgScalars* __restrict__ sc) {
  const int m = blockIdx.x;
  if (sc[m].done) return;
  __shared__ double red[32];
  const int2 r = range[m];
  double v[2] = {0.0, 0.0};
  for (int i = r.x + threadIdx.x; i < r.y; i += blockDim.x) {
    v[0] += part[(size_t)i * stride];
    if (stride == 2) v[1] += part[(size_t)i * stride + 1];
  }
  block_sum(v, red);
  if (threadIdx.x == 0) bicg_update(sc + m, stage, v[0], v[1]);
}

// ----------------------------------------------------------------- refinement
__global__ void birbegin_k(int members, Skip sk, const double* __restrict__ rtol, const double* __restrict__ sums,
                           IrState* __restrict__ ir) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= members) return;
  const double bn = sqrt(sums[m]);
  IrState q{};
  q.target = rtol[m] * bn;
  q.rn = bn;
  q.status = isfinite(bn) ? 0 : 3;
  q.done = (skipped(sk, m) || !(bn > 0.0) || q.status) ? 1 : 0;
  ir[m] = q;
}

// x += d (the FP32 correction)
template <int B>
__global__ void __launch_bounds__(kThreads) birupd_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                     const IrState* __restrict__ ir, int x) {
  const int2 w = work[blockIdx.x];
  if (ir[w.x].done) return;
  const MemberView& m = mv[w.x];
  const int s = w.y * kThreads + threadIdx.x;
  if (s >= m.nslot) return;
  double* xv = m.v[x] + (size_t)s * B;
  const float* d = m.f[9] + (size_t)s * B;
#pragma unroll
  for (int i = 0; i < B; ++i) xv[i] += (double)d[i];
}

// res = b - ax, partial ||res||^2
template <int B>
__global__ void __launch_bounds__(kThreads) birres_k(const int2* __restrict__ work, const MemberView* __restrict__ mv,
                                                     const IrState* __restrict__ ir, int b, int ax, int res,
                                                     double* __restrict__ part) {
  __shared__ double red[16];
  const int2 w = work[blockIdx.x];
  double acc[1] = {0.0};
  if (!ir[w.x].done) {
    const MemberView& m = mv[w.x];
    const int s = w.y * kThreads + threadIdx.x;
    if (s < m.nslot) {
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const size_t k = (size_t)s * B + i;
        const double v = m.v[b][k] - m.v[ax][k];
        m.v[res][k] = v;
        acc[0] += v * v;
      }
    }
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ void bircheck_k(int members, const double* __restrict__ sums, const BicgScalars* __restrict__ sc,
                           IrState* __restrict__ ir) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= members) return;
  IrState q = ir[m];
  if (q.done) return;
  q.its += sc[m].iter;
  q.refs += 1;
  const double rn = sqrt(sums[m]);
  if (sc[m].done == 3 || !isfinite(rn)) {
    q.status = 3;
    q.done = 1;
  } else {
    q.rn = rn;
    if (rn <= q.target) q.done = 1;
  }
  ir[m] = q;
}

inline unsigned grid8(int n) { return (unsigned)(((n + 7) / 8) * 8); }
inline unsigned grid_for(int n, int t = kThreads) { return (unsigned)((n + t - 1) / t); }

template <int B, typename T, typename V>
void bspmv_launch(int variant, const BatchWork& w, const MemberView* mv, Skip sk, const BicgScalars* sc, int x, int y,
                  int w1, int w2, double* part, hipStream_t st) {
  if (w.ncwork <= 0) return;
  if (variant == 5)
    bspmv_k<B, T, V, true><<<grid8(w.ncwork), kThreads, 0, st>>>(w.ncwork, w.cwork, mv, sk, sc, x, y, w1, w2, part);
  else
    bspmv_k<B, T, V, false><<<grid8(w.ncwork), kThreads, 0, st>>>(w.ncwork, w.cwork, mv, sk, sc, x, y, w1, w2, part);
}

#define DCG_BATCH_B(B, ...)                                \
  switch (B) {                                             \
    case 1: { constexpr int BB = 1; __VA_ARGS__; break; }  \
    case 2: { constexpr int BB = 2; __VA_ARGS__; break; }  \
    case 3: { constexpr int BB = 3; __VA_ARGS__; break; }  \
    case 4: { constexpr int BB = 4; __VA_ARGS__; break; }  \
    case 5: { constexpr int BB = 5; __VA_ARGS__; break; }  \
    case 6: { constexpr int BB = 6; __VA_ARGS__; break; }  \
    case 7: { constexpr int BB = 7; __VA_ARGS__; break; }  \
    case 8: { constexpr int BB = 8; __VA_ARGS__; break; }  \
    default: break;                                        \
  }

}  // namespace

// ----------------------------------------------------------------- launchers
void b_copy(int B, const BatchWork& w, const MemberView* mv, Skip sk, int dst, int src, hipStream_t s) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (bcopy_k<BB><<<w.nvwork, kThreads, 0, s>>>(w.vwork, mv, sk, dst, src)));
}

void b_axpy(int B, const BatchWork& w, const MemberView* mv, Skip sk, double a, int x, int y, hipStream_t s) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (baxpy_k<BB><<<w.nvwork, kThreads, 0, s>>>(w.vwork, mv, sk, a, x, y)));
}

void b_norm(int B, const BatchWork& w, const MemberView* mv, Skip sk, int x, int op, double* part, double* out,
            hipStream_t s) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (bnorm_k<BB><<<w.nvwork, kThreads, 0, s>>>(w.vwork, mv, sk, x, op, part)));
  bnorm_fin_k<<<w.members, kThreads, 0, s>>>(w.vrange, part, op, out);
}

void b_block_inverse(int B, const BatchWork& w, const MemberView* mv, Skip sk, int* status, bool point,
                     hipStream_t s) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (binv_k<BB><<<w.nvwork, kThreads, 0, s>>>(w.vwork, mv, sk, status, point ? 1 : 0)));
}

void b_to_float(int B, const BatchWork& w, const MemberView* mv, Skip sk, hipStream_t s) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (btof_k<BB><<<w.nvwork, kThreads, 0, s>>>(w.vwork, mv, sk)));
}

void b_bicg_setup(const BatchWork& w, Skip sk, const double* tol, const IrState* ir, double inner, int maxit,
                  BicgScalars* sc, hipStream_t s) {
  bsetup_k<<<grid_for(w.members, 64), 64, 0, s>>>(w.members, sk, tol, ir, inner, maxit, sc);
}

template <typename V>
void b_bicg_init(int B, const BatchWork& w, const MemberView* mv, int b, int x, double* part, BicgScalars* sc,
                 hipStream_t st) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (binit_k<BB, V><<<w.nvwork, kThreads, 0, st>>>(w.vwork, mv, b, x, part, sc)));
  bfin_k<<<w.members, kThreads, 0, st>>>(w.vrange, part, 1, 0, sc);
}
template void b_bicg_init<double>(int, const BatchWork&, const MemberView*, int, int, double*, BicgScalars*,
                                  hipStream_t);
template void b_bicg_init<float>(int, const BatchWork&, const MemberView*, int, int, double*, BicgScalars*,
                                 hipStream_t);

template <typename V>
void b_bicg_iteration(int B, int variant, const BatchWork& w, const MemberView* mv, int x, int jacobi,
                      double* part, BicgScalars* sc, hipStream_t st, hipEvent_t* ev, const BatchOperator* op) {
  if (w.nvwork <= 0) return;
  using T = V;  // the matrix copy has the precision of the Krylov vectors
  const int jac = jacobi;  // 0 none, 1 vertex blocks, 2 the scalar diagonal
  const Skip none{};
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  // the operator: the members' BCSR matrices, or the given one (point storage)
  const int2* orange = op ? op->range : w.crange;
  auto matvec = [&](int xs, int ys, int w1, int w2) {
    if (op) {
      op->apply(xs, ys, w1, w2, part, 0);
      return;
    }
    DCG_BATCH_B(B, (bspmv_launch<BB, T, V>(variant, w, mv, none, sc, xs, ys, w1, w2, part, st)));
  };
  DCG_BATCH_B(B, {
    bp_k<BB, V><<<w.nvwork4, kThreads, 0, st>>>(w.vwork4, mv, jac, sc);
    mark(0);
    matvec(kVKry + 4, kVKry + 3, kVKry + 1, -1);
    mark(1);
    bfin_k<<<w.members, kThreads, 0, st>>>(orange, part, 2, 1, sc);
    bs_k<BB, V><<<w.nvwork4, kThreads, 0, st>>>(w.vwork4, mv, jac, sc, part);
    bfin_k<<<w.members, kThreads, 0, st>>>(w.vrange4, part, 1, 2, sc);
    mark(2);
    matvec(kVKry + 6, kVKry + 7, kVKry + 5, kVKry + 7);
    mark(3);
    bfin_k<<<w.members, kThreads, 0, st>>>(orange, part, 2, 3, sc);
    bxr_k<BB, V><<<w.nvwork4, kThreads, 0, st>>>(w.vwork4, mv, x, sc, part);
    bfin_k<<<w.members, kThreads, 0, st>>>(w.vrange4, part, 2, 4, sc);
  });
}
template void b_bicg_iteration<double>(int, int, const BatchWork&, const MemberView*, int, int, double*,
                                       BicgScalars*, hipStream_t, hipEvent_t*, const BatchOperator*);
template void b_bicg_iteration<float>(int, int, const BatchWork&, const MemberView*, int, int, double*,
                                      BicgScalars*, hipStream_t, hipEvent_t*, const BatchOperator*);

void b_spmv64(int B, int variant, const BatchWork& w, const MemberView* mv, Skip sk, int x, int y, hipStream_t st) {
  DCG_BATCH_B(B, (bspmv_launch<BB, double, double>(variant, w, mv, sk, nullptr, x, y, -1, -1, nullptr, st)));
}

void b_spmv32(int B, int variant, const BatchWork& w, const MemberView* mv, Skip sk, int x, int y, hipStream_t st) {
  DCG_BATCH_B(B, (bspmv_launch<BB, float, float>(variant, w, mv, sk, nullptr, x, y, -1, -1, nullptr, st)));
}

void b_copy_f(int B, const BatchWork& w, const MemberView* mv, Skip sk, int sel, hipStream_t st) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (bzerof_k<BB><<<w.nvwork, kThreads, 0, st>>>(w.vwork, mv, sk, sel)));
}

void b_ir_begin(const BatchWork& w, Skip sk, const double* rtol, const double* sums, IrState* ir, hipStream_t st) {
  birbegin_k<<<grid_for(w.members, 64), 64, 0, st>>>(w.members, sk, rtol, sums, ir);
}

void b_ir_update(int B, const BatchWork& w, const MemberView* mv, const IrState* ir, int x, hipStream_t st) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (birupd_k<BB><<<w.nvwork, kThreads, 0, st>>>(w.vwork, mv, ir, x)));
}

void b_ir_residual(int B, const BatchWork& w, const MemberView* mv, const IrState* ir, int b, int ax, int res,
                   double* part, double* sums, hipStream_t st) {
  if (w.nvwork <= 0) return;
  DCG_BATCH_B(B, (birres_k<BB><<<w.nvwork, kThreads, 0, st>>>(w.vwork, mv, ir, b, ax, res, part)));
  bnorm_fin_k<<<w.members, kThreads, 0, st>>>(w.vrange, part, 0, sums);
}

void b_ir_check(const BatchWork& w, const double* sums, const BicgScalars* sc, IrState* ir, hipStream_t st) {
  bircheck_k<<<grid_for(w.members, 64), 64, 0, st>>>(w.members, sums, sc, ir);
}

}  // namespace dcopasi::k
