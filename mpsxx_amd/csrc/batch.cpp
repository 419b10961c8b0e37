// Batched two-site sigma vectors (one grouped launch per contraction step for many instances, hipGraph replay,
// autotuned tile policies).
#include "batch.hpp"

#include <algorithm>
#include <climits>
#include <cstring>

#include "mpsxx_amd.h"
#include "ops.hpp"

namespace mpx {

namespace {

std::map<std::vector<uint64_t>, std::vector<GemmPolicy>>& tuned_policies() {
  static std::map<std::vector<uint64_t>, std::vector<GemmPolicy>> m;
  return m;
}

bool same_shape(const Layout& a, const Layout& b) { return a.rank == b.rank && a.q == b.q && a.qs == b.qs && a.ds == b.ds; }

}  // namespace

BatchSigma::BatchSigma(const std::vector<Operand6>& inst, int prec, hipStream_t s) : inst_(inst), prec_(prec) {
  n_ = (int)inst_.size();
  if (n_ <= 0) throw Error(MPSXX_ERR_ARG, "batch: no instances");
  if (prec != MPSXX_FP64 && prec != MPSXX_FP32 && prec != MPSXX_BF16X3)
    throw Error(MPSXX_ERR_ARG, "batch: unknown precision");
  plans_.reserve(n_);
  for (int i = 0; i < n_; ++i) {
    const Operand6& o = inst_[i];
    for (Tensor* t : {o.lmpo, o.rmpo, o.lopr, o.ropr, o.wfn})
      if (!t || t->empty) throw Error(MPSXX_ERR_ARG, "batch compute_sigma_vector: an input array is empty");
    if (!o.sgv) throw Error(MPSXX_ERR_ARG, "batch compute_sigma_vector: missing output");
    plans_.push_back(build_plan_sigma_twosite(o.lmpo->L, o.rmpo->L, o.lopr->L, o.ropr->L, o.wfn->L,
                                              o.sgv->empty ? nullptr : &o.sgv->L, prec));
    for (Tensor* t : {o.lmpo, o.rmpo, o.lopr, o.ropr, o.wfn}) {
      in_hash_.push_back(t->L.hash);
      in_layout_.push_back(t->L);
    }
    flops_ += plans_.back().flops;
    bytes_ += plans_.back().bytes;
  }
  // group the instances by plan shape (slot roles, step kinds): boundary sites, for example, turn a GEMM step into
  // a coefficient AXPY
  std::map<std::vector<int64_t>, int> gid;
  for (int i = 0; i < n_; ++i) {
    const Plan& P = plans_[i];
    std::vector<int64_t> sig{(int64_t)P.slot_layout.size()};
    sig.insert(sig.end(), P.slot_input.begin(), P.slot_input.end());
    sig.insert(sig.end(), P.slot_prec.begin(), P.slot_prec.end());
    for (const Step& x : P.steps) {
      if (x.kind == Step::COPY) throw Error(MPSXX_ERR_UNSUPPORTED, "batch: copy steps are not batched");
      int64_t alpha_bits;
      std::memcpy(&alpha_bits, &x.alpha, 8);
      sig.insert(sig.end(), {(int64_t)x.kind, x.a, x.b, x.b2, x.c, x.accumulate, alpha_bits});
    }
    auto it = gid.find(sig);
    if (it == gid.end()) {
      it = gid.emplace(sig, (int)groups_.size()).first;
      groups_.emplace_back();
      groups_.back().nslots = (int)P.slot_layout.size();
    }
    groups_[it->second].members.push_back(i);
  }
  for (Group& g : groups_) {
    g.table_off = table_ptrs_;
    table_ptrs_ += (size_t)g.nslots * g.members.size();
  }
  // concatenated task lists, one launch per (group, step)
  for (int gi = 0; gi < (int)groups_.size(); ++gi) {
    const Group& g = groups_[gi];
    const Plan& P0 = plans_[g.members[0]];
    for (size_t k = 0; k < P0.steps.size(); ++k) {
      launches_.emplace_back();
      Launch& L = launches_.back();
      L.group = gi;
      Step& st = L.st;
      const Step& s0 = P0.steps[k];
      st.kind = s0.kind;
      st.a = s0.a;
      st.b = s0.b;
      st.b2 = s0.b2;
      st.c = s0.c;
      st.accumulate = s0.accumulate;
      st.alpha = s0.alpha;
      L.f32 = P0.slot_prec[st.a] == 1;
      L.cf32 = P0.slot_prec[st.c] == 1;
      for (int m = 0; m < (int)g.members.size(); ++m) {
        Step& si = plans_[g.members[m]].steps[k];
        st.flops += si.flops;
        if (st.kind == Step::GEMM) {
          const int32_t pbase = (int32_t)st.host.gpairs.size();
          for (GemmTask t : si.host.gtasks) {
            t.pair_begin += pbase;
            t.inst = m;
            st.host.gtasks.push_back(t);
          }
          st.host.gpairs.insert(st.host.gpairs.end(), si.host.gpairs.begin(), si.host.gpairs.end());
        } else {
          const int32_t tbase = (int32_t)st.host.atasks.size(), pbase = (int32_t)st.host.apairs.size();
          for (AxpyTask t : si.host.atasks) {
            t.pair_begin += pbase;
            t.inst = m;
            st.host.atasks.push_back(t);
          }
          st.host.apairs.insert(st.host.apairs.end(), si.host.apairs.begin(), si.host.apairs.end());
          for (Chunk c : si.host.chunks) {
            c.task += tbase;
            st.host.chunks.push_back(c);
          }
        }
        si = Step();  // the per-instance lists are never launched
      }
      if (st.kind == Step::GEMM) {
        st.tasks = upload_vec(st.host.gtasks);
        st.pairs = upload_vec(st.host.gpairs);
      } else {
        st.tasks = upload_vec(st.host.atasks);
        st.pairs = upload_vec(st.host.apairs);
        st.work = upload_vec(st.host.chunks);
        st.nwork = (int)st.host.chunks.size();
      }
    }
  }
  // tile policies: remembered ones for this structure, else the defaults
  auto it = tuned_policies().find(structure_key());
  if (it != tuned_policies().end() && it->second.size() == launches_.size()) pol_ = it->second;
  else pol_.assign(launches_.size(), default_policy(prec_));
  for (GemmPolicy& p : pol_) p = policy_for(p, prec_);
  for (int k = 0; k < (int)launches_.size(); ++k) build_step_lists(k);
  // FP32 input copies
  conv_chunks_.resize(5);
  nconv_chunks_.assign(5, 0);
  if (prec_ != MPSXX_FP64) {
    for (int idx = 0; idx < 5; ++idx) {
      std::vector<Chunk> ch;
      for (int i = 0; i < n_; ++i)
        for (int j = 0; j < (int)plans_[i].slot_layout.size(); ++j) {
          if (plans_[i].slot_input[j] != idx || plans_[i].slot_prec[j] != 1) continue;
          const int seg = (int)seg_inst_.size();
          seg_inst_.push_back(i);
          seg_slot_.push_back(j);
          for (int64_t r = 0; r < plans_[i].slot_layout[j].nelem; r += kAxpyChunk) ch.push_back({seg, (int32_t)r});
        }
      nconv_chunks_[idx] = (int)ch.size();
      conv_chunks_[idx] = upload_vec(ch);
    }
  }
  table_bytes_ = table_ptrs_ * sizeof(void*) + seg_inst_.size() * sizeof(ConvSeg);
  table_.alloc(table_bytes_);
  hip_check(hipHostMalloc(&pinned_, table_bytes_, hipHostMallocDefault), "hipHostMalloc");
  hip_check(hipEventCreateWithFlags(&upload_done_, hipEventDisableTiming), "event");
  hip_check(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking), "stream");
  check_structures(s);
}

BatchSigma::~BatchSigma() {
  drop_graphs();
  if (upload_done_) {
    (void)hipEventSynchronize(upload_done_);
    (void)hipEventDestroy(upload_done_);
  }
  if (pinned_) (void)hipHostFree(pinned_);
  if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
}

std::vector<uint64_t> BatchSigma::structure_key() const {
  std::vector<uint64_t> k{(uint64_t)prec_, (uint64_t)n_};
  k.insert(k.end(), in_hash_.begin(), in_hash_.end());
  for (const Plan& P : plans_) k.push_back(P.out.hash);
  return k;
}

int64_t BatchSigma::work_items() const {
  int64_t w = 0;
  for (const Launch& L : launches_) w += L.st.nwork + L.st.nwork_small + L.st.nwork_split;
  return w;
}

void BatchSigma::drop_graphs() {
  for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
  for (auto& g : graph_src_) (void)hipGraphDestroy(g.second);
  graphs_.clear();
  graph_src_.clear();
}

void BatchSigma::build_step_lists(int k) {
  Step& st = launches_[k].st;
  if (st.kind == Step::GEMM) {
    set_gemm_policy(st, pol_[k]);
  } else {
    st.max_groups = pol_[k].max_groups;
    st.policy = pol_[k];
  }
}

void BatchSigma::set_policies(const std::vector<GemmPolicy>& pol) {
  if (pol.size() != launches_.size()) throw Error(MPSXX_ERR_ARG, "batch: one policy per launch expected");
  hip_check(hipDeviceSynchronize(), "batch sync");  // launches in flight may still read the old tile lists
  pol_ = pol;
  for (GemmPolicy& p : pol_) p = policy_for(p, prec_);
  for (int k = 0; k < (int)launches_.size(); ++k) build_step_lists(k);
  drop_graphs();
}

void BatchSigma::check_structures(hipStream_t s) {
  for (int i = 0; i < n_; ++i) {
    const Operand6& o = inst_[i];
    const Tensor* in[5] = {o.lmpo, o.rmpo, o.lopr, o.ropr, o.wfn};
    for (int j = 0; j < 5; ++j)
      if (in[j]->empty || !in[j]->L.same_structure(in_layout_[5 * i + j]))
        throw Error(MPSXX_ERR_SHAPE, "batch: an input's block structure changed since the batch was built");
    const Layout& out = plans_[i].out;
    Tensor& y = *o.sgv;
    if (y.empty) {
      y.set_layout(out, s);
    } else if (!y.L.same_structure(out)) {
      if (!y.ptr() && y.L.nelem > 0)  // a structure-only output (set_vectors) cannot be laid out again
        throw Error(MPSXX_ERR_SHAPE, "batch: an output without storage must have the batch's output structure");
      // the output may only have lost blocks (e.g. a caller-side clear and refill); every block must still exist
      if (!same_shape(y.L, out)) throw Error(MPSXX_ERR_QN, "batch compute_sigma_vector: output quantum shape mismatch");
      for (int64_t b = 0; b < y.L.nblocks(); ++b)
        if (out.find(y.L.tag[b]) < 0)
          throw Error(MPSXX_ERR_SHAPE, "batch: the output holds blocks outside the batch's output structure");
      y.relayout(out, s);
    }
  }
}

void BatchSigma::set_vectors(const std::vector<const double*>& wfn, const std::vector<double*>& sgv) {
  if ((!wfn.empty() && (int)wfn.size() != n_) || (!sgv.empty() && (int)sgv.size() != n_))
    throw Error(MPSXX_ERR_ARG, "batch: one wavefunction and one sigma array per instance expected");
  vwfn_ = wfn;
  vsgv_ = sgv;
}

const double* BatchSigma::input_ptr(int i, int idx) const {
  if (idx == 4 && !vwfn_.empty()) return vwfn_[i];
  const Operand6& o = inst_[i];
  const Tensor* in[5] = {o.lmpo, o.rmpo, o.lopr, o.ropr, o.wfn};
  return in[idx]->ptr();
}

double* BatchSigma::output_ptr(int i) const { return vsgv_.empty() ? inst_[i].sgv->ptr() : vsgv_[i]; }

void BatchSigma::upload_tables(hipStream_t s, const std::vector<void*>* out_override) {
  std::vector<char> buf(table_bytes_);
  void** T = reinterpret_cast<void**>(buf.data());
  for (const Group& g : groups_) {
    const size_t nm = g.members.size();
    for (size_t m = 0; m < nm; ++m) {
      const int i = g.members[m];
      const Plan& P = plans_[i];
      for (int j = 0; j < g.nslots; ++j) {
        void* p;
        const int si = P.slot_input[j];
        if (si >= 0) p = P.slot_prec[j] == 1 ? P.slot_buf[j].p : (void*)input_ptr(i, si);
        else if (si == -2) p = out_override ? (*out_override)[i] : (void*)output_ptr(i);
        else p = P.slot_buf[j].p;
        T[g.table_off + (size_t)j * nm + m] = p;
      }
    }
  }
  ConvSeg* segs = reinterpret_cast<ConvSeg*>(buf.data() + table_ptrs_ * sizeof(void*));
  for (size_t g = 0; g < seg_inst_.size(); ++g) {
    const int i = seg_inst_[g], j = seg_slot_[g];
    const Plan& P = plans_[i];
    ConvSeg sg;
    sg.src = input_ptr(i, P.slot_input[j]);
    sg.dst = P.slot_buf[j].as<float>();
    sg.n = P.slot_layout[j].nelem;
    std::memcpy(&segs[g], &sg, sizeof(sg));
  }
  if (buf == shadow_) return;
  // the previous upload may still be reading the pinned staging buffer
  hip_check(hipEventSynchronize(upload_done_), "batch table sync");
  std::memcpy(pinned_, buf.data(), table_bytes_);
  hip_check(hipMemcpyAsync(table_.p, pinned_, table_bytes_, hipMemcpyHostToDevice, s), "batch table upload");
  hip_check(hipEventRecord(upload_done_, s), "event");
  shadow_.swap(buf);
}

void BatchSigma::launch_one(int k, hipStream_t s, unsigned parts) {
  const Launch& L = launches_[k];
  const Step& st = L.st;
  const Group& g = groups_[L.group];
  const size_t nm = g.members.size();
  void** T = table_.as<void*>() + g.table_off;
  const int b2 = st.b2 >= 0 ? st.b2 : (st.b >= 0 ? st.b : st.a);
  Operands ops{T + (size_t)st.a * nm, st.b >= 0 ? T + (size_t)st.b * nm : nullptr, T + (size_t)b2 * nm,
               T + (size_t)st.c * nm, 1};
  launch_step(st, ops, L.f32, L.cf32, s, parts, prec_ == MPSXX_BF16X3);
}

void BatchSigma::enqueue(unsigned convert_mask, hipStream_t s, const std::vector<hipEvent_t>* events) {
  int e = 0;
  if (events) hip_check(hipEventRecord((*events)[e++], s), "event");
  if (prec_ != MPSXX_FP64) {
    const ConvSeg* segs =
        reinterpret_cast<const ConvSeg*>(table_.as<char>() + table_ptrs_ * sizeof(void*));
    for (int idx = 0; idx < 5; ++idx)
      if (convert_mask & (1u << idx))
        launch_convert_segments(conv_chunks_[idx].as<Chunk>(), nconv_chunks_[idx], segs, s);
  }
  for (int k = 0; k < (int)launches_.size(); ++k) {
    if (events) hip_check(hipEventRecord((*events)[e++], s), "event");
    launch_one(k, s);
  }
  if (events) hip_check(hipEventRecord((*events)[e++], s), "event");
}

void BatchSigma::run(unsigned convert_mask, bool use_graph, hipStream_t s, bool verify) {
  if (prec_ == MPSXX_FP64) convert_mask = 0;
  if (verify) check_structures(s);
  upload_tables(s, nullptr);
  if (!use_graph) {
    enqueue(convert_mask, s, nullptr);
  } else {
    auto it = graphs_.find(convert_mask);
    if (it == graphs_.end()) {
      // kernels only (no memset nodes); the pointer tables they read are refreshed outside the graph
      hipGraph_t g = nullptr;
      hip_check(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal), "begin capture");
      try {
        enqueue(convert_mask, cap_stream_, nullptr);
      } catch (...) {
        (void)hipStreamEndCapture(cap_stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hip_check(hipStreamEndCapture(cap_stream_, &g), "end capture");
      hipGraphExec_t ex = nullptr;
      hip_check(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "graph instantiate");
      graph_src_[convert_mask] = g;
      it = graphs_.emplace(convert_mask, ex).first;
    }
    hip_check(hipGraphLaunch(it->second, s), "graph launch");
  }
  hip_check(hipGetLastError(), "batch launch");
}

double BatchSigma::time(int reps, unsigned convert_mask, bool use_graph, hipStream_t s) {
  if (reps <= 0) throw Error(MPSXX_ERR_ARG, "batch time: reps must be positive");
  // the untimed first call also captures the graph
  return time_launches([&] { run(convert_mask, use_graph, s); }, reps, s);
}

std::vector<PartTiming> BatchSigma::profile(int reps, hipStream_t s) {
  if (reps <= 0) throw Error(MPSXX_ERR_ARG, "batch profile: reps must be positive");
  check_structures(s);
  std::vector<DevBuf> scratch(n_);
  std::vector<void*> outp(n_);
  for (int i = 0; i < n_; ++i) {
    scratch[i].alloc(std::max<int64_t>(plans_[i].out.nelem, 1) * sizeof(double));
    outp[i] = scratch[i].p;
  }
  upload_tables(s, &outp);
  enqueue(~0u, s, nullptr);
  std::vector<PartTiming> res;
  for (int k = 0; k < (int)launches_.size(); ++k)
    for (KernelPart kp : step_parts(launches_[k].st, launches_[k].f32 ? 4 : 8, launches_[k].cf32 ? 4 : 8)) {
      const double ms = time_launches([&] { launch_one(k, s, kp.mask); }, reps, s);
      if (prec_ == MPSXX_BF16X3 && kp.kind == KernelPart::GEMM64) kp.kind = KernelPart::GEMM_X3;
      if (prec_ == MPSXX_BF16X3 && kp.kind == KernelPart::GEMM128) kp.kind = KernelPart::GEMM_X3_128;
      res.push_back({k, kp, ms});
    }
  hip_check(hipStreamSynchronize(s), "profile sync");
  upload_tables(s, nullptr);
  hip_check(hipStreamSynchronize(s), "profile sync");
  return res;
}

std::vector<GemmPolicy> BatchSigma::tune(int reps, hipStream_t s) {
  if (reps <= 0) throw Error(MPSXX_ERR_ARG, "batch tune: reps must be positive");
  check_structures(s);
  // redirect the outputs to scratch so that repeated accumulations leave the callers' sgv untouched
  std::vector<DevBuf> scratch(n_);
  std::vector<void*> outp(n_);
  for (int i = 0; i < n_; ++i) {
    scratch[i].alloc(std::max<int64_t>(plans_[i].out.nelem, 1) * sizeof(double));
    outp[i] = scratch[i].p;
  }
  upload_tables(s, &outp);
  enqueue(~0u, s, nullptr);  // every intermediate now holds real data
  auto time_step = [&](int k) { return time_launches([&] { launch_one(k, s); }, reps, s); };
  const int grid_caps[] = {0, 1024, 2048, 4096, 8192};
  for (int k = 0; k < (int)launches_.size(); ++k) {
    std::vector<GemmPolicy> cand;
    if (launches_[k].st.kind == Step::GEMM) {
      for (int lim : {0, 32, 64, 96, 128, 192})
        for (int su : {2, 4, 8, INT_MAX}) {
          if (lim == 0 && su != 2) continue;  // no small blocks: the split threshold is irrelevant
          for (int cap : grid_caps) {
            GemmPolicy c = pol_[k];  // keeps the 32x32 split settings
            c.small_limit = lim;
            c.split_units = su;
            c.max_groups = cap;
            cand.push_back(c);
          }
        }
    } else {
      for (int cap : grid_caps) {
        GemmPolicy c = pol_[k];
        c.small_limit = 0;
        c.split_units = 0;
        c.max_groups = cap;
        c.split_big = 0;
        cand.push_back(c);
      }
    }
    // the current policy first: a candidate must beat it by 1 % to replace it
    cand.insert(cand.begin(), pol_[k]);
    double best = 1e300;
    GemmPolicy bp = pol_[k];
    for (size_t c = 0; c < cand.size(); ++c) {
      pol_[k] = cand[c];
      build_step_lists(k);
      const double t = time_step(k);
      if (t < (c == 0 ? best : 0.99 * best)) {
        best = t;
        bp = cand[c];
      }
    }
    // second pass: the split threshold of the 64x64 tiles, with the other choices fixed
    if (launches_[k].st.kind == Step::GEMM) {
      for (int sb : {0, 512, 2048, 8192}) {
        if (sb == bp.split_big) continue;
        pol_[k] = bp;
        pol_[k].split_big = sb;
        build_step_lists(k);
        const double t = time_step(k);
        if (t < 0.99 * best) {
          best = t;
          bp = pol_[k];
        }
      }
    }
    pol_[k] = bp;
    build_step_lists(k);
  }
  hip_check(hipStreamSynchronize(s), "tune sync");
  upload_tables(s, nullptr);
  hip_check(hipStreamSynchronize(s), "tune sync");
  drop_graphs();
  tuned_policies()[structure_key()] = pol_;
  return pol_;
}

}  // namespace mpx
