// Host-side launchers of the ahead-of-time solver kernels (solver.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace dcopasi::k {

// LDS stage of the SpMV: block-row entries (blocks x B) per workgroup chunk.
constexpr int kSpmvStage = 4096;

// Chunking of BCSR rows so each chunk's blocks x B fit the LDS stage.
struct SpmvPlan {
  // variant 0: LDS-staged chunks, one thread per block row entry; 5: LDS-staged
  // chunks, one thread per whole block; 1..4: row-split with 8/16/32/64 lanes
  // per row (only when the block size divides the group); chosen by the autotuner
  int variant = 0;
  int nchunk = 0;
  const int* chunk_row = nullptr;  // nchunk + 1 row boundaries (device)
  const int* rp = nullptr;         // BCSR row pointer (device)
  const int* col = nullptr;        // BCSR columns (device)
  int nrows = 0;                   // row-split: rows, workgroups (multiple of 8), rows per workgroup
  int row_wgs = 0;
  int rows_per_wg = 0;
  // number of partial-sum pairs a launch writes
  int parts() const { return (variant == 0 || variant == 5) ? nchunk : row_wgs; }
};

// Device-resident BiCGSTAB state shared by its kernels.
struct BicgScalars {
  double rho, alpha, omega, beta, r0, rnorm, tol, abstol;
  int iter, maxit, done, half;
};

void fill(double* p, double v, long n, hipStream_t s);
void fill(float* p, float v, long n, hipStream_t s);
void fill(int* p, int v, long n, hipStream_t s);
void init_jacobian(double* val, long nval, const int* diag_pos, const int* slot_b, int nslot, int B, hipStream_t s);
// copy `bytes` (multiple of 32) from a to b with a plain streaming kernel
void probe_copy(const void* a, void* b, long bytes, hipStream_t s);
// zero the B x B blocks at the given block positions
void zero_blocks(double* val, const int* pos, long n, int B, hipStream_t s);
void axpy(long n, double a, const double* x, double* y, hipStream_t s);
void axpby(long n, double a, const double* x, double b, double* y, hipStream_t s);
int reduce_parts(long n);
// op 0: sum x*y (y null: x*x), 1: sum |x|, 2: max |x|; `part` holds reduce_parts(n) doubles
void reduce(long n, const double* x, const double* y, int op, double* part, double* out, hipStream_t s);

// y = A x; with `part` the per-chunk partial dots y.w1, y.w2 are written
// (part[2c], part[2c+1]); with `sc` the launch is skipped once the solve is
// done.  `stage` is unused by the kernel.
void spmv(int B, const SpmvPlan& pl, const double* val, const double* x, double* y, const double* w1,
          const double* w2, double* part, BicgScalars* sc, int stage, hipStream_t s);
void spmv(int B, const SpmvPlan& pl, const float* val, const float* x, float* y, const float* w1, const float* w2,
          double* part, BicgScalars* sc, int stage, hipStream_t s);

void block_inverse(int B, int nslot, const double* val, const int* diag_pos, double* dinv, int* status, hipStream_t s);

template <typename V>
void bicg_init(long n, const double* b, V* x, V* r, V* rh, V* p, V* v, double* part, int nparts, BicgScalars* sc,
               hipStream_t s);
template <typename V>
void bicg_p(int B, int nslot, const V* r, V* p, const V* v, V* ph, const double* dinv, const BicgScalars* sc,
            hipStream_t s);
template <typename V>
void bicg_s(int B, int nslot, const V* r, const V* v, V* sv, V* sh, const double* dinv, const BicgScalars* sc,
            double* part, hipStream_t s);
template <typename V>
void bicg_xr(int B, int nslot, V* x, const V* ph, const V* sh, const V* sv, const V* t, V* r, const V* rh,
             const BicgScalars* sc, double* part, hipStream_t s);

void ilu_factor(int B, int nrows, const int* rows, const int* rp, const int* col, const int* diag_pos, double* lu,
                double* dinv, int* status, hipStream_t s);
template <typename V>
void ilu_lower(int B, int nrows, const int* rows, const int* rp, const int* col, const int* diag_pos,
               const double* lu, const V* b, V* y, const BicgScalars* sc, hipStream_t s);
template <typename V>
void ilu_upper(int B, int nrows, const int* rows, const int* rp, const int* col, const int* diag_pos,
               const double* lu, const double* dinv, const V* y, V* x, const BicgScalars* sc, hipStream_t s);

void unpack_state(long nslot, int B, const int* slot_b, const long* slot_dof, const double* packed, double* padded,
                  hipStream_t s);
void pack_state(long nslot, int B, const int* slot_b, const long* slot_dof, const double* padded, double* packed,
                hipStream_t s);
// out[i] = sum over slots [first, first+n) of w[s] * x[s*B+i], i < B (<= 8); part: 8*reduce_parts(n)
void weighted_species_sum(long first, long n, int B, const double* w, const double* x, double* part, double* out,
                          hipStream_t s);
// reduce per-workgroup partials (stride 1 or 2) and update the scalars of `stage`
void bicg_fin(int nparts, const double* part, int stride, int stage, BicgScalars* sc, hipStream_t s);
void to_float(long n, const double* a, float* b, hipStream_t s);
void to_double(long n, const float* a, double* b, hipStream_t s);

}  // namespace dcopasi::k
