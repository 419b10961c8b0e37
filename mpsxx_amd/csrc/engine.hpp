// Device tensors, the block-sparse contraction planner and composite-op plans.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "kernels.hpp"
#include "layout.hpp"

namespace mpx {

void hip_check(hipError_t e, const char* what);

// Device allocations since the library was loaded: hipMalloc/hipFree calls and their host time, buffers handed out
// again from the pool (and the device synchronizations that took), bytes held by DevBufs and by the pool.
struct DeviceMemoryStats {
  int64_t mallocs = 0, frees = 0, reuses = 0, pool_syncs = 0, bytes_live = 0, cached_bytes = 0;
  double malloc_ms = 0.0, free_ms = 0.0;
};
DeviceMemoryStats device_memory_stats();
// frac of the current device's memory (fallback_gb GiB when it cannot be queried): default cache bounds.
size_t device_memory_fraction(double frac, double fallback_gb);

// Owning device buffer (from the library's caching pool; see engine.cpp).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  size_t cls = 0;  // size of the pooled allocation behind p
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  ~DevBuf();
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), cls(o.cls) {
    o.p = nullptr;
    o.bytes = 0;
    o.cls = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept;
  void alloc(size_t n);
  void release();
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Block placement in one launch: y[dst + r] = x[src + r] for every segment (CopySeg with stride 1); the segment list
// is uploaded with the call (replaces one hipMemcpyAsync per block).
void copy_segments(const std::vector<CopySeg>& segs, const double* x, double* y, hipStream_t s);

// Host vector -> new device buffer (synchronous copy).
template <typename T>
DevBuf upload_vec(const std::vector<T>& v) {
  DevBuf b(v.size() * sizeof(T));
  if (!v.empty()) hip_check(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
  return b;
}

// Device-resident block-sparse FP64 array.  Every Tensor object registers its address while it exists, so that
// code holding a pointer recorded earlier can check that the object is still alive (Tensor::alive).
struct Tensor {
  Layout L;
  DevBuf data;      // nelem doubles
  bool empty = true;  // size() == 0 in the reference sense (no structure yet)

  Tensor();
  ~Tensor();
  Tensor(Tensor&& o) noexcept;
  Tensor& operator=(Tensor&& o) noexcept = default;
  Tensor(const Tensor&) = delete;
  Tensor& operator=(const Tensor&) = delete;
  static bool alive(const Tensor* t);

  double* ptr() const { return data.as<double>(); }
  void set_layout(const Layout& l, hipStream_t s);       // allocate + zero
  void relayout(const Layout& l, hipStream_t s);         // keep data of blocks present in both
};

// View of a tensor operand, optionally conjugated (QSTArray::conjugate: QNs negated, data shared).
struct View {
  const Layout* L = nullptr;
  bool conj = false;
  QN q() const { return conj ? -L->q : L->q; }
  QN qs(int i, int k) const { return conj ? -L->qs[i][k] : L->qs[i][k]; }
};

// Result of the symbolic Contract: output structure + one pair list per output block.
struct ContractSym {
  Layout C;                               // finalized output structure
  std::vector<GemmTask> tasks;            // one per output block receiving contributions; c_off in C's layout
  std::vector<GemmPair> pairs;
  int64_t flops = 0;                      // 2 * sum m*n*k over pairs
};

// Symbolic BTAS Contract(a[ca], b[cb]) (include/btas/QSPARSE/QSTCONTRACT.h:16-57 + STBLAS.h:298-359).  If `existing`
// is non-null and non-empty, its q/qshape are checked and its blocks kept (output = existing U new).  If `mask` is
// non-null, quantum-number symmetric and has the output's sector counts, only output blocks that `mask`'s quantum
// numbers allow receive pairs (for an output that is afterwards up-cast onto `mask`, which keeps exactly those);
// otherwise it is ignored.
ContractSym contract_symbolic(const View& A, const std::vector<int>& ca, const View& B, const std::vector<int>& cb,
                              const Layout* existing, const Layout* mask = nullptr);

// Planner threading: threads > 0 fixes the host thread count (0: MPSXX_HOST_THREADS or the hardware concurrency, at
// most 16); a contraction uses one thread per min_pairs candidate pairs at most.  Plans do not depend on either.
void set_planner_threads(int threads, int64_t min_pairs);
// Return every pooled (released, unused) device buffer to the device, after a device-wide synchronization.
void device_pool_trim();
int host_threads();
// Run f(0..n-1) on n host threads (f(0) on the caller); the first exception is rethrown after all have finished.
void parallel_for(int n, const std::function<void(int)>& f);

// Index lists of btas::Gemm(transa, transb) for ranks (NA, NB, NC) (QSTBLAS.h:183-226).
void gemm_indices(int transa, int transb, int NA, int NB, int NC, std::vector<int>& ca, std::vector<int>& cb);

// Structure of btas::Tie (STREINDEX.h:95-147) and its gather segments.
Layout tie_layout(const Layout& X, const std::vector<int>& tie, std::vector<CopySeg>& segs);

// Host copies of a step's task lists (batched launches concatenate them across instances).
struct HostLists {
  std::vector<GemmTask> gtasks;
  std::vector<GemmPair> gpairs;
  std::vector<AxpyTask> atasks;
  std::vector<AxpyPair> apairs;
  std::vector<Chunk> chunks;
};

// How the blocks of a GEMM step are mapped to kernels: blocks with both sides <= small_limit run as 32x32 tiles, split
// over the 4 waves of a work-group when they reduce over >= 2 pairs and >= split_units (pair, 16-wide k panel) units,
// one wave per tile otherwise; larger blocks run as 64x64 LDS tiles, whose pair lists are split over several
// work-groups (partial tiles + an ordered sum) while the step has fewer than split_big such tiles (0: never).  The
// split 32x32 tiles are cut the same way into pair ranges of at least split_seg units, each on its own work-group,
// until the list holds about split_small work-groups (0: never).  max_groups > 0 caps the grid of every launch.
struct GemmPolicy {
  int small_limit = 64;
  int split_units = 2;
  int max_groups = 0;
  // in 64x64-tile equivalents (a 128x128 tile counts four); measured on the norb=24 M=1024 merged sigma, whose
  // second step holds a few hundred long-K tiles: 2048 -> 16384 takes FP64 from 32.9 to 29.1 ms, FP32 from 19.7 to
  // 17.3 ms and split-BF16x3 from 14.5 to 13.7 ms, with the L=32 M=512 Hubbard sigma unchanged; with the later
  // 128x128 tiles 16384 -> 32768 takes FP32 from 11.50 to 11.35 ms and FP64 from 24.16 to 24.06 ms
  int split_big = 32768;
  // 1024 -> 4096: the merged sigma on the converged M=1024 QC state 4.14 -> 3.99 ms FP32, 3.24 -> 3.10 ms split-BF16x3,
  // 7.33 -> 7.28 ms FP64; the L=32 M=512 Hubbard sigma (39.6 us), its 64-chain batch and sweeps unchanged
  int split_small = 4096;
  int split_seg = 32;
  // blocks with both sides >= this run as 128x128 tiles (gemm128_kernel); 0: never.  The FP64 and FP32 default;
  // split-BF16x3 uses kBigTileMinX3 (the measurements are at default_policy(int))
  int big_tile_min = 192;
};
constexpr int kBigTileMinX3 = 64;  // split-BF16x3 threshold of the 128x128 tiles
// the defaults above, overridable by MPSXX_SMALL_GEMM_LIMIT, MPSXX_SPLIT_UNITS, MPSXX_SPLIT_BIG, MPSXX_SPLIT_SMALL,
// MPSXX_SPLIT_SEG, MPSXX_BIG_TILE_MIN
GemmPolicy default_policy();
// the defaults for the block GEMMs of one precision (MPSXX_FP64 / MPSXX_FP32 / MPSXX_BF16X3): 128x128 tiles for
// blocks with both sides >= big_tile_min in every precision (MPSXX_BIG_TILE_MIN_F64 sets the FP64 threshold, 0: none)
GemmPolicy default_policy(int precision);
// `p` with the tile settings `precision` has kernels for
GemmPolicy policy_for(GemmPolicy p, int precision);

// Tile lists of one GEMM step (heaviest blocks first in each list).
struct GemmTiling {
  std::vector<BigTile> big;
  std::vector<RedTile> red;
  int nslots = 0;  // partial 64x64 tiles of the split big reductions
  std::vector<GemmTile> small;
  std::vector<BigTile> split;      // 32x32 tiles over a pair range; slot >= 0: partial 32x32 tile
  std::vector<RedTile> split_red;  // 32x32 tiles whose partial sums are added in slot order
  int split_slots = 0;
  std::vector<BigTile> big2;       // 128x128 tiles of the largest blocks (split reductions as for the 64x64 tiles)
  std::vector<RedTile> red2;
  int nslots2 = 0;
};
void gemm_tiles(const std::vector<GemmTask>& tasks, const std::vector<GemmPair>& pairs, const GemmPolicy& pol,
                GemmTiling& out);

// One device step of a plan.
struct Step {
  enum Kind { GEMM, AXPY, COPY } kind = GEMM;
  int a = -1, b = -1, c = -1;  // slots: GEMM (A,B,C); AXPY (V,S,C); COPY (x,-,y)
  int b2 = -1;                 // fused AXPY: slot of the second coefficient tensor
  int accumulate = 0;
  double alpha = 1.0, beta = 0.0;
  int nwork = 0;               // GEMM: 64x64 tiles (BigTile); AXPY: chunks
  DevBuf work, tasks, pairs;   // device copies of the task lists
  int nwork_red = 0;           // GEMM: tiles of split reductions (RedTile) and their partial buffer
  DevBuf work_red, partial;
  int nwork_small = 0;         // GEMM: 32x32 tiles of the small blocks (gemm_small_kernel)
  DevBuf work_small;
  int nwork_split = 0;         // GEMM: 32x32 tiles with long reductions (gemm_split_kernel, BigTile)
  DevBuf work_split;
  int nwork_split_red = 0;     // GEMM: split 32x32 tiles cut over work-groups (RedTile) and their partial buffer
  DevBuf work_split_red, split_partial;
  int nwork2 = 0;              // GEMM: 128x128 tiles of the largest blocks (gemm128_kernel), their split reductions
  int nwork2_layout[4] = {0, 0, 0, 0};  // of those, tiles of tasks with layout flags 3, 1, 2, 0, listed in that order
  int nwork_layout[4] = {0, 0, 0, 0};   // the same for the 64x64 tiles (work, nwork)
  DevBuf work2;
  int nwork_red2 = 0;
  DevBuf work_red2, partial2;
  int max_groups = 0;          // grid cap of the step's launches (0: one work-group per work item)
  GemmPolicy policy;           // GEMM: how the blocks were mapped to the tile lists
  int64_t flops = 0;
  HostLists host;
};

// A copy of the step with its own device lists (copied on the device) and host lists.
Step clone_step(const Step& st);

// One kernel launch of a step, with its share of the step's algorithmic cost: flops 2*rows*cols*k summed over the
// tiles of the launch (GEMM) or 2*len per pair (AXPY); bytes = operand elements read once per tile + outputs, at their
// storage size.
struct KernelPart {
  enum Kind { GEMM64 = 0, GEMM_SMALL = 1, GEMM_SPLIT = 2, AXPY = 3, GEMM_X3 = 4, GEMM128 = 5, GEMM_X3_128 = 6 };
  int kind;
  unsigned mask;  // `parts` bit of launch_step
  int items;
  int64_t flops, bytes;
};
// ea, ec: bytes per operand element and per output element of the launch (8 for FP64, 4 for FP32 storage)
std::vector<KernelPart> step_parts(const Step& st, int ea = 8, int ec = 8);
// (Re)build the device tile lists of a GEMM step from its host task lists.
void set_gemm_policy(Step& st, const GemmPolicy& pol);
const char* part_kernel_name(int kind);

// Slot table: inputs are bound per call; intermediates are owned by the plan.
// A plan's launch sequence captured once in a hipGraph, its kernels reading their operand bases from a device table
// of slot pointers (the batched kernels' pointer tables with one instance) that is refreshed, outside the graph, only
// when a slot's address changed (another trial vector, a restored buffer).
struct PlanGraph {
  DevBuf table;
  void* pinned = nullptr;
  std::vector<void*> shadow;
  hipEvent_t uploaded = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  PlanGraph() = default;
  PlanGraph(const PlanGraph&) = delete;
  PlanGraph& operator=(const PlanGraph&) = delete;
  ~PlanGraph();
};

struct Plan {
  std::vector<Step> steps;
  std::vector<Layout> slot_layout;     // per slot
  std::vector<int> slot_input;         // >= 0: index of the bound input tensor; -1 intermediate; -2 output
  std::vector<DevBuf> slot_buf;        // owned storage of intermediates (FP64 or FP32)
  std::vector<int> slot_prec;          // 0 FP64, 1 FP32
  Layout out;                          // output structure
  int64_t flops = 0;                   // algorithmic flops (reference order)
  int64_t bytes = 0;                   // algorithmic bytes (inputs + output, FP64)
  int precision = 0;
  bool stripped = false;               // slot buffers released by strip(); run_plan re-allocates them first
  // per slot: the FP64 input pointer the slot's FP32 copy was last converted from (nullptr: the copy holds nothing).
  // A copy is converted again when its bit is set in convert_mask, when it was never converted, or when the slot is
  // bound to a different input than the one it was converted from.
  std::vector<const double*> conv_src;
  std::shared_ptr<PlanGraph> graph;    // run_plan_graph's capture (created on first use)
  // Release the slot buffers (intermediates, FP32 operand copies), keeping the task lists; restore() allocates them
  // again (FP32 operand copies must then be converted again).  Returns whether anything was allocated.
  void strip();
  bool restore();
  // device memory held by the plan (intermediates, FP32 operand copies, task lists)
  size_t device_bytes() const {
    size_t n = 0;
    for (const DevBuf& b : slot_buf) n += b.bytes;
    for (const Step& st : steps)
      n += st.work.bytes + st.tasks.bytes + st.pairs.bytes + st.work_red.bytes + st.partial.bytes +
           st.work_small.bytes + st.work_split.bytes + st.work_split_red.bytes + st.split_partial.bytes +
           st.work2.bytes + st.work_red2.bytes + st.partial2.bytes;
    return n;
  }
};

// Builders of device steps from symbolic results.
Step make_contract_step(const ContractSym& sym, int a, int b, int c, int accumulate, double alpha,
                        const GemmPolicy& pol = default_policy());
Step make_copy_step(const std::vector<CopySeg>& segs, int x, int y, double alpha, double beta);
// Two consecutive coefficient contractions X = Contract(Y, W1), Z = Contract(X, W2) (both with a scalar B side)
// executed as one pass Z[r] += sum W1*W2*Y[..] without materialising X; returns false (no step) when either
// contraction is not of that form.
bool make_fused_axpy_step(const ContractSym& first, const ContractSym& second, int y, int w1, int w2, int z,
                          int accumulate, Step& out);

// Launch the kernels of one GEMM or AXPY step (FP64, FP32, or FP32 operands into an FP64 output); `parts` selects
// the tile lists (bit 0: 64x64 tiles or the AXPY chunks, bit 1: small tiles, bit 2: split tiles, bit 3: 128x128
// tiles).  x3: FP32 operands of the 64x64 tiles multiplied as split-BF16 pairs (precision MPSXX_BF16X3).
constexpr unsigned kAllParts = 15u;
void launch_step(const Step& st, const Operands& ops, bool f32, bool cf32, hipStream_t s, unsigned parts = kAllParts,
                 bool x3 = false);

// Execute a plan: inputs[i] are the FP64 device pointers of the bound inputs, out the FP64 output pointer.
void run_plan(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s);
// Same, recording an event before each step and after the last one (events.size() >= steps + 1).
void run_plan_timed(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s,
                    const std::vector<hipEvent_t>& events);
// Full form: FP32 plans convert only the inputs whose bit is set in convert_mask (the others reuse the FP32 copy
// made by an earlier call with the same input contents).
// run_plan_ex with the plan's launches replayed from its hipGraph (FP32 operand conversions are launched before the
// graph, the pointer table uploaded when it changed).  Plans with copy steps run as run_plan_ex does.
void run_plan_graph(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s,
                    unsigned convert_mask);
// Whether MPSXX_SIGMA_GRAPH (default on; "0" switches it off) lets the Davidson solvers replay their sigma plans.
bool sigma_graphs_enabled();
int set_sigma_graphs(int on);  // returns the previous setting
void run_plan_ex(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s,
                 const std::vector<hipEvent_t>& events, unsigned convert_mask);

// Device pointers of every slot of a plan for one call (FP32 plans convert the inputs whose bit is set).
std::vector<void*> slot_pointers(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s,
                                 unsigned convert_mask);

// Mean device time (HIP events on s) of `launch` over reps calls, after one untimed call.
double time_launches(const std::function<void()>& launch, int reps, hipStream_t s);

// Per kernel launch of a plan: device time and algorithmic cost (the slots must hold data from an earlier run).
struct PartTiming {
  int step;
  KernelPart part;
  double ms;
};
std::vector<PartTiming> profile_plan_parts(Plan& P, const std::vector<const double*>& inputs, double* out, int reps,
                                           hipStream_t s);

const char* kind_name(int k);

}  // namespace mpx

// Opaque handle behind the C ABI's mpsxx_tensor; the tag lets entry points reject freed or foreign handles.
struct mpsxx_tensor_s {
  static constexpr uint64_t kMagic = 0x4d50535854454e53ull;
  uint64_t magic = kMagic;
  mpx::Tensor t;
  ~mpsxx_tensor_s() { magic = 0; }
};
