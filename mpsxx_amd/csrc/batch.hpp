// Batched two-site sigma vectors: many independent problems (different chains, bond dimensions, quantum-number
// structures) run through ONE grouped launch per contraction step.  Each instance keeps its own plan (its
// intermediates are that instance's scratch); instances whose plans have the same step sequence form a group, the
// group's task lists are concatenated with the instance's index in every task, and the kernels read their operand
// bases from per-instance pointer tables.  The launch sequence can be captured once in a hipGraph and replayed.
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "engine.hpp"

namespace mpx {

class BatchSigma {
 public:
  // One instance = the six operands of compute_sigma_vector (src/compute_sigma_vector.hpp:54-70); sgv accumulates.
  struct Operand6 {
    Tensor *lmpo, *rmpo, *lopr, *ropr, *wfn, *sgv;
  };

  BatchSigma(const std::vector<Operand6>& inst, int prec, hipStream_t s);
  ~BatchSigma();
  BatchSigma(const BatchSigma&) = delete;
  BatchSigma& operator=(const BatchSigma&) = delete;

  // sgv_i += H_i * wfn_i for every instance.  FP32 batches refresh the FP32 copies of the inputs whose bit is set in
  // convert_mask (bit 0 lmpo, 1 rmpo, 2 lopr, 3 ropr, 4 wfn); use_graph replays a captured graph of the launches.
  // verify: check first that no input's block structure changed since creation (O(blocks) per tensor; a caller
  // that holds the inputs fixed, like the batched Davidson, checks once).
  void run(unsigned convert_mask, bool use_graph, hipStream_t s, bool verify = true);

  // Redirect instance i's wfn input to wfn[i] and its sgv output to sgv[i]: device arrays laid out as that instance's
  // wfn and sgv tensors, which then only describe the structure (a batched Davidson moves its trial and sigma vectors
  // from call to call).  The pointer tables follow at the next run; empty lists switch the redirection off.
  void set_vectors(const std::vector<const double*>& wfn, const std::vector<double*>& sgv);

  // Average device time of one run over `reps` runs (HIP events on s).
  double time(int reps, unsigned convert_mask, bool use_graph, hipStream_t s);
  // Device time and algorithmic cost of every kernel launch (outputs redirected to scratch).
  std::vector<PartTiming> profile(int reps, hipStream_t s);

  // Autotuner: for every step, time each candidate tile policy (block-size cut between the 32x32 and 64x64 kernels,
  // K-split threshold, grid cap) on the real task lists and keep the fastest.  Outputs are redirected to scratch while
  // tuning, so the instances' sgv are untouched.  Choices are remembered per structure for later batches.
  std::vector<GemmPolicy> tune(int reps, hipStream_t s);
  void set_policies(const std::vector<GemmPolicy>& pol);
  const std::vector<GemmPolicy>& policies() const { return pol_; }

  int size() const { return n_; }
  int nsteps() const { return (int)launches_.size(); }
  int precision() const { return prec_; }
  int ngroups() const { return (int)groups_.size(); }
  int64_t flops() const { return flops_; }
  int64_t bytes() const { return bytes_; }
  int64_t step_flops(int k) const { return launches_[k].st.flops; }
  int64_t work_items() const;

 private:
  // instances with identical plan shapes (slot roles and step kinds)
  struct Group {
    std::vector<int> members;  // instance indices; a task's `inst` is the position in this list
    int nslots = 0;
    size_t table_off = 0;      // first pointer of the group's tables: [slot][member]
  };
  // one contraction step of one group over all its members
  struct Launch {
    int group = 0;
    Step st;
    bool f32 = false, cf32 = false;
  };

  void check_structures(hipStream_t s);
  void upload_tables(hipStream_t s, const std::vector<void*>* out_override);
  void enqueue(unsigned convert_mask, hipStream_t s, const std::vector<hipEvent_t>* events);
  void launch_one(int k, hipStream_t s, unsigned parts = kAllParts);
  void build_step_lists(int k);
  void drop_graphs();
  std::vector<uint64_t> structure_key() const;

  int n_ = 0, prec_ = 0;
  std::vector<Operand6> inst_;
  std::vector<uint64_t> in_hash_;  // 5 input structure hashes per instance, fixed at creation (tuned-policy key)
  std::vector<Layout> in_layout_;  // the 5 input layouts per instance the plans were built for
  std::vector<Plan> plans_;
  std::vector<Group> groups_;
  std::vector<Launch> launches_;   // group-major, each group's steps in plan order
  std::vector<GemmPolicy> pol_;    // per launch
  // FP32 input copies: one segment per (instance, input), chunk lists per input index
  std::vector<int> seg_inst_, seg_slot_;
  std::vector<DevBuf> conv_chunks_;
  std::vector<int> nconv_chunks_;
  // device pointer tables of every group, then the conversion segments
  DevBuf table_;
  size_t table_ptrs_ = 0, table_bytes_ = 0;
  void* pinned_ = nullptr;
  std::vector<char> shadow_;
  hipEvent_t upload_done_ = nullptr;
  hipStream_t cap_stream_ = nullptr;
  std::map<unsigned, hipGraphExec_t> graphs_;
  std::map<unsigned, hipGraph_t> graph_src_;
  int64_t flops_ = 0, bytes_ = 0;
  std::vector<const double*> vwfn_;  // set_vectors redirections (empty: the tensors' own buffers)
  std::vector<double*> vsgv_;
  const double* input_ptr(int i, int idx) const;
  double* output_ptr(int i) const;
};

// One instance of a batched Davidson solve: the two-site problem of compute_sigma_vector (one operator group), the
// diagonal from compute_diagonal_elements (may be empty), and wfnc, the start vector, overwritten by the normalized
// eigenvector on the instance's Davidson layout.
struct BatchDavInstance {
  Tensor *lmpo, *rmpo, *lopr, *ropr, *diag, *wfnc;
};
struct BatchDavidsonStats {
  std::vector<double> energy;    // lowest Ritz value of each instance at its convergence (or last step)
  std::vector<double> rnorm;     // its last ||r||^2
  std::vector<int> nsigma;       // sigma vectors applied to the instance while it was iterating
  std::vector<int> nsigma_low;   // of them in FP32 / split-BF16x3
  std::vector<int> converged;    // ||r||^2 < tol reached in the final (FP64) phase
  int batch_sigmas = 0;          // batched applications (every instance at once)
  int batch_sigmas_low = 0;
  int macro_iterations = 0;
  double sigma_ms = 0.0;         // device time of the batched applications
  double seconds = 0.0;          // wall time of the whole solve
};
// Davidson (src/diagonalize.hpp:283-449) on every instance at once, in lockstep: one batched sigma application and
// one set of segmented vector operations per step over all instances; precisions as davidson_twosite (FP32 or
// split-BF16x3 sigma vectors until their noise floor, then FP64 refinement).  use_graph replays the sigma launches
// from a hipGraph.
void batch_davidson_twosite(std::vector<BatchDavInstance>& inst, double tol, int prec, bool use_graph,
                            BatchDavidsonStats& st, hipStream_t s);

}  // namespace mpx

// Opaque handle behind the C ABI's mpsxx_batch.
struct mpsxx_batch_s {
  static constexpr uint64_t kMagic = 0x4d50535842415443ull;
  uint64_t magic = kMagic;
  std::unique_ptr<mpx::BatchSigma> b;
  ~mpsxx_batch_s() { magic = 0; }
};
