// Host-side launchers of the ahead-of-time solver kernels (solver.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>

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

// ----------------------------------------------------------------- batches
// A batch advances many independent models with one launch per operation:
// every batched kernel walks a work list of (member, block) pairs and finds
// the member's buffers in a device-resident MemberView array, so members keep
// their own allocations and join or leave without moving data.

// FP64 vectors of a member, by selector
enum BatchVec : int {
  kVU = 0,      // state u_n
  kVHist = 1,   // history of earlier stages
  kVR = 2,      // Newton residual
  kVZ = 3,      // Newton correction
  kVXs = 4,     // stage iterate
  kVStage = 5,  // stage values 0..2 (5, 6, 7)
  kVKry = 8,    // FP64 Krylov vectors r rh p v ph s sh t tmp (8..16)
  kBatchVec = 17
};
constexpr int kBatchVecF = 10;  // FP32 Krylov vectors r rh p v ph s sh t tmp x

struct MemberView {
  int nslot = 0, nchunk = 0;
  const int* rp = nullptr;
  const int* col = nullptr;
  const int* diag_pos = nullptr;
  const int* chunk_row = nullptr;
  double* val = nullptr;  // BCSR values, or the rows' diagonal blocks (point storage, diag_pos the identity)
  float* val32 = nullptr;
  double* dinv = nullptr;
  double* v[kBatchVec] = {};
  float* f[kBatchVecF] = {};
  BicgScalars* sc = nullptr;
  // point storage: the coefficients, their FP32 copy and their count (0:
  // the member holds an assembled matrix)
  const double* pc = nullptr;
  float* pc32 = nullptr;
  long npc = 0;
};

// One batch member as the batched point-coefficient apply sees it (host twin
// of the generated dcg_pc_member, same field order and types): its
// compartment's tables and coefficients, the linearisation's time and
// stiffness weight, its vectors by selector and the flags that end its work.
struct PcMember {
  const int* g2l = nullptr;  // null: the slots are the grid's vertices in order
  const unsigned char* cmask = nullptr;
  const double* pc = nullptr;
  const float* pc32 = nullptr;
  const double* tca = nullptr;
  const double* prm = nullptr;
  int n[4] = {1, 1, 1, 0};
  double* v[kBatchVec] = {};
  float* f[kBatchVecF] = {};
  const int* done = nullptr;    // BicgScalars::done of the member's Krylov solve
  const int* irdone = nullptr;  // IrState::done of its refinement
};

// The operator of a batched Krylov solve when it is not the members' BCSR
// matrices: apply(x, y, w1, w2, part, skip) computes y = A x for vector
// selectors (FP32 solves: kVKry + i for the FP32 vector i) with the partial
// dots (y.w1, y.w2) per work item into part (w1 < 0: none), skipping members
// whose solve is done (skip 0), whose refinement is done (1) or none (2);
// range[m] = [first, last) partial pairs of member m.
struct BatchOperator {
  std::function<void(int x, int y, int w1, int w2, double* part, int skip)> apply;
  const int2* range = nullptr;
};

// per-member state of FP64 iterative refinement around the FP32 Krylov solve
struct IrState {
  double target, rn;
  int done, status, its, refs;
};

// skip member m when (p[m * stride] != 0) == when; p == nullptr: never
struct Skip {
  const int* p = nullptr;
  int stride = 0;
  int when = 0;
};

// work lists: (member, 256-slot block) for vectors, (member, block of
// kBVecSlots * 256 slots) for the BiCGSTAB updates (vwork4, vrange4); for
// SpMV one item per chunk of rows, {member, first row, first block, rows <<
// 16 | blocks}, so a workgroup starts streaming after a single dependent
// load; range[m] = [first, last) work items of member m
constexpr int kBVecSlots = 4;
struct BatchWork {
  const int2* vwork = nullptr;
  int nvwork = 0;
  const int2* vrange = nullptr;
  const int2* vwork4 = nullptr;
  int nvwork4 = 0;
  const int2* vrange4 = nullptr;
  const int4* cwork = nullptr;
  int ncwork = 0;
  const int2* crange = nullptr;
  int members = 0;
};

// Host twin of the generated dcg_asm_member (same field order and types):
// one batch member as the batched tile kernels see it.
constexpr int kAsmComps = 4;
struct AsmMember {
  double* vec[8] = {};  // BatchVec selectors kVU .. kVStage + 2
  double* val = nullptr;
  const int* own_start = nullptr;
  const unsigned* own_mask = nullptr;
  const double* prm = nullptr;
  const int* g2l[kAsmComps] = {};
  const unsigned char* cmask[kAsmComps] = {};
  int n[4] = {1, 1, 1, 0};
  double origin[3] = {0, 0, 0};
  double h[3] = {0, 0, 0};
  // point storage: coefficients per compartment, the rows' diagonal blocks,
  // and where the set-up leaves the linearisation's (time, stiffness weight)
  double* pc[kAsmComps] = {};
  double* pdiag = nullptr;
  double* tca = nullptr;
};

void b_copy(int B, const BatchWork& w, const MemberView* mv, Skip sk, int dst, int src, hipStream_t s);
void b_axpy(int B, const BatchWork& w, const MemberView* mv, Skip sk, double a, int x, int y, hipStream_t s);
// per-member reduction of vector x (op 0: x.x, 1: sum |x|, 2: max |x|) into out[m]
void b_norm(int B, const BatchWork& w, const MemberView* mv, Skip sk, int x, int op, double* part, double* out,
            hipStream_t s);
void b_block_inverse(int B, const BatchWork& w, const MemberView* mv, Skip sk, int* status, bool point, hipStream_t s);
void b_to_float(int B, const BatchWork& w, const MemberView* mv, Skip sk, hipStream_t s);
// Krylov set-up: tolerances from tol[m] (or, with ir, the refinement's inner
// tolerance max(inner, target/(2 rn)) capped at 0.5); members skipped by sk or
// finished refinements start done
void b_bicg_setup(const BatchWork& w, Skip sk, const double* tol, const IrState* ir, double inner, int maxit,
                  BicgScalars* sc, hipStream_t s);
// b: FP64 selector; x: FP64 selector (FP64 Krylov) or the FP32 x vector
template <typename V>
void b_bicg_init(int B, const BatchWork& w, const MemberView* mv, int b, int x, double* part, BicgScalars* sc,
                 hipStream_t st);
// one BiCGSTAB iteration on every unfinished member (jacobi: apply dinv)
// ev (optional, 4 events): recorded before and after each of the two SpMVs
template <typename V>
void b_bicg_iteration(int B, int variant, const BatchWork& w, const MemberView* mv, int x, int jacobi,
                      double* part, BicgScalars* sc, hipStream_t st, hipEvent_t* ev = nullptr,
                      const BatchOperator* op = nullptr);
// y = A x on the FP64 matrix (selectors), members skipped by sk
void b_spmv64(int B, int variant, const BatchWork& w, const MemberView* mv, Skip sk, int x, int y, hipStream_t st);
// y = A x on the FP32 matrix copy, FP32 Krylov vectors (selectors kVKry + i)
void b_spmv32(int B, int variant, const BatchWork& w, const MemberView* mv, Skip sk, int x, int y, hipStream_t st);
// zero the FP32 Krylov vector at selector sel (kVKry + i)
void b_copy_f(int B, const BatchWork& w, const MemberView* mv, Skip sk, int sel, hipStream_t st);
// refinement: ir from ||b||^2 in sums; x += d; res = b - A x; residual check
void b_ir_begin(const BatchWork& w, Skip sk, const double* rtol, const double* sums, IrState* ir, hipStream_t st);
void b_ir_update(int B, const BatchWork& w, const MemberView* mv, const IrState* ir, int x, hipStream_t st);
void b_ir_residual(int B, const BatchWork& w, const MemberView* mv, const IrState* ir, int b, int ax, int res,
                   double* part, double* sums, hipStream_t st);
void b_ir_check(const BatchWork& w, const double* sums, const BicgScalars* sc, IrState* ir, hipStream_t st);

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

// partial dots (y.w1, y.w2) per workgroup into part[2 g], part[2 g + 1] over
// `groups` workgroups (w2 may be null); nothing once sc->done
void dot_parts(long n, const double* y, const double* w1, const double* w2, double* part, int groups,
               const BicgScalars* sc, hipStream_t s);
void dot_parts(long n, const float* y, const float* w1, const float* w2, double* part, int groups,
               const BicgScalars* sc, hipStream_t s);

// Jacobi preconditioner blocks: the inverse of each diagonal B x B block, or
// with point set only the inverse of its diagonal (in the same layout)
void block_inverse(int B, int nslot, const double* val, const int* diag_pos, double* dinv, int* status, bool point,
                   hipStream_t s);

template <typename V>
void bicg_init(long n, const double* b, V* x, V* r, V* rh, V* p, V* v, double* part, int nparts, BicgScalars* sc,
               hipStream_t s);
template <typename V>
void bicg_p(int B, int nslot, const V* r, V* p, const V* v, V* ph, const double* dinv, bool diag, const BicgScalars* sc,
            hipStream_t s);
template <typename V>
void bicg_s(int B, int nslot, const V* r, const V* v, V* sv, V* sh, const double* dinv, bool diag, const BicgScalars* sc,
            double* part, hipStream_t s);
// workgroup partials bicg_s writes for nslot slots
int bicg_s_parts(int nslot);
// returns the number of partial pairs it writes (the finaliser's count)
template <typename V>
int bicg_xr(int B, int nslot, V* x, const V* ph, const V* sh, const V* sv, const V* t, V* r, const V* rh,
             const BicgScalars* sc, double* part, hipStream_t s);

// The fixed device work space of the allocation-free entry points: its size,
// and its device address on `device` (looked up once per device, with that
// device current; every call on a device shares the one array).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
double* noalloc_scratch(int device);

void ilu_factor(int B, int nrows, const int* rows, const int* rp, const int* col, const int* diag_pos, double* lu,
                double* dinv, int* status, hipStream_t s);
// Tables of the single-workgroup ILU0 sweeps (every level in one workgroup,
// barriers between levels), built at set-up from the level schedules:
//   lo_lev / up_lev: level boundaries into lo_rows / up_rows (row positions);
//   lo_ent[r] .. lo_ent[r+1]: the strictly lower blocks of the row at
//   position r: block lo_blk[e] of the matrix and the LDS ring slot lo_src[e]
//   of its column's result; up_*: the same for the strictly upper blocks;
//   lo_chunk / up_chunk: the levels in chunks whose blocks and rows fit the
//   stage (at most stage_ent blocks, stage_rows rows, stage_levels levels);
//   ring, width: ring levels and the widest level (chunked: the apply runs
//   ilu_apply_chunk_k, else the plain one-workgroup sweep);
//   inbox (blocks of at most 2 species and rows within the kIluReg* bounds):
//   the factorisation reads each row's blocks from `packed` and its
//   operands from its `inbox` slots, both at fixed positions in level order,
//   with the tables frow (row, first block, blocks, lower blocks), fub, fm
//   and fpush (see ilu_factor_inbox_k); factor_threads threads.
constexpr int kIluRegLower = 4, kIluRegPairs = 16, kIluRegRow = 9, kIluRegPush = 24;
constexpr int kIluThreads = 256;  // workgroup of the chunked apply
constexpr int kIluFubStride = 8;  // ints per row of fub (kIluRegLower + 1, padded to 16-byte loads)
struct IluSweep {
  int nlo = 0, nup = 0;
  const int *lo_lev = nullptr, *lo_rows = nullptr, *up_lev = nullptr, *up_rows = nullptr;
  bool chunked = false;
  int nlo_chunk = 0, nup_chunk = 0, stage_ent = 0, stage_rows = 0, stage_levels = 0, ring = 0, width = 0;
  int reg_k = 0;  // rows of at most reg_k blocks, levels of at most kIluThreads rows: the register path
  const int *lo_chunk = nullptr, *lo_ent = nullptr, *lo_blk = nullptr, *lo_src = nullptr;
  const int *up_chunk = nullptr, *up_ent = nullptr, *up_blk = nullptr, *up_src = nullptr;
  int factor_threads = 0;
  bool inbox_path = false;
  int nrows = 0;
  const int4* frow = nullptr;
  const int *fub = nullptr, *fm = nullptr;
  const int2* fpush = nullptr;
  double *packed = nullptr, *inbox = nullptr;
};
// every level of the factorisation / of both triangular solves in one workgroup
void ilu_factor_sweep(int B, const IluSweep& t, const int* rp, const int* col, const int* diag_pos, double* lu,
                      double* dinv, int* status, hipStream_t s);
template <typename V>
void ilu_apply_sweep(int B, const IluSweep& t, const int* rp, const int* col, const int* diag_pos, const double* lu,
                     const double* dinv, const V* in, V* tmp, V* out, const BicgScalars* sc, hipStream_t s);
// LDS bytes of the chunked apply's stage and ring for values of vbytes bytes
size_t ilu_chunk_lds_bytes(const IluSweep& t, int B, int vbytes);
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
// stages 2 and 3 of one iteration in one launch: na partial norms of the
// s-update (stride 1), then nb partial dot pairs of the second operator
void bicg_fin23(int na, const double* pa, int nb, const double* pb, BicgScalars* sc, hipStream_t s);
void to_float(long n, const double* a, float* b, hipStream_t s);
void to_double(long n, const float* a, double* b, hipStream_t s);

}  // namespace dcopasi::k
