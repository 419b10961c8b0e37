// The 64x64 grouped-GEMM tiles shared by kernels.hip (FP32 operands) and kernels_f64.hip (FP64 operands, compiled
// with MFMA accumulators in VGPRs, which keeps them out of per-panel register copies); device code, included only by
// .hip files.
//
// grouped GEMM: one 256-thread work-group per 64x64 C tile; 4 waves, each owning a 32x32 quarter computed with
// 2x2 v_mfma_{f64,f32}_16x16x4 tiles.  A and B tiles (64x16, 16x64) are staged in LDS k-major so that the MFMA
// operand of lane l is the LDS word [k + l/16][i + l%16].  The global->LDS mapping picks the unit-stride direction of
// each operand so loads are coalesced whether the block is used as stored or transposed.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mpx {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Operand layout of both 16x16x4 instructions: lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Result layout (measured on gfx950): FP32 lane l, register r holds D[4*(l/16) + r][l%16];
//                                     FP64 lane l, register r holds D[(l/16) + 4*r][l%16].
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef dbl4 acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct Mfma<float> {
  typedef flt4 acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

constexpr int BM = 64, BN = 64, BK = 16, LDP = 64 + 1;

// Work-group -> work item of a launch with one work-group per item.  Work-groups are dispatched round-robin over the
// 8 XCDs (work-group b runs on XCD b % 8), each XCD with an L2 of its own.  Items are taken in chunks of `chunk`
// consecutive items, which share operand panels: the i-th work-group of XCD x takes item i % chunk of chunk
// 8 (i / chunk) + x, so one chunk runs on one XCD and shares its L2, while the chunks, and with them the heavy items at
// the head of the list, are dealt over the XCDs in turn.  A bijection: the tail that does not fill a round of 8 chunks
// keeps the identity, as do persistent launches (fewer work-groups than items) and chunk <= 1.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item(int b, int nwg, int nitems, int chunk) {
  if (nwg != nitems || chunk <= 1) return b;
  const int full = nitems / (kXcds * chunk) * (kXcds * chunk);
  if (b >= full) return b;
  const int x = b % kXcds, i = b / kXcds;
  return ((i / chunk) * kXcds + x) * chunk + i % chunk;
}

// operand base of the task's instance: Tab = the launch passes per-instance pointer tables (batched launches, which
// are distinct kernel symbols in profiles)
template <typename T, bool Tab>
__device__ __forceinline__ T* opnd(const void* p, int inst) {
  if constexpr (Tab) return reinterpret_cast<T* const*>(p)[inst];
  else return reinterpret_cast<T*>(const_cast<void*>(p));
}

// Panel staging of the 64x64 and 128x128 tiles through buffer loads.  A 16-wide k panel of LINES lines (rows of A or
// columns of B) is moved by 256 threads, LINES/16 elements each, into LDS k-major ([k][line]).  Every load goes
// against a buffer resource covering the operand block: an element outside the block gets an offset beyond the
// extent and reads as zero, so every panel (inside the block or on its edge, and the empty panel past the end of a
// tile) goes through the same branch-free loads and the compiler counts the loads in flight exactly (guarded loads in
// several code paths made it wait for every outstanding load, the prefetch included, before each LDS store).  Element
// -> thread maps, chosen per operand layout:
//   unit stride along k: a thread takes 4 consecutive k (4*(t%4)..+3) of lines t/4 + 64*run, run < LINES/64, one
//     16-byte load per line for FP32 and two for FP64; k past the pair's end is zeroed when the panel is stored;
//   otherwise: a thread takes line t%LINES at k t/LINES + (256/LINES)*i, one element per load, LINES consecutive
//     lines per load instruction over the work-group.
constexpr int32_t kOutOfBlock = 0x7fffffff;  // a byte offset beyond every block extent (the host keeps blocks < 2^30 B)

// One operand of the current pair: the buffer resource of its block and the thread's offsets.
struct PanelView {
  __amdgpu_buffer_rsrc_t rsrc;
  int32_t boff0;  // byte offset of the thread's first element for k0 = 0
  int32_t bstep;  // byte offset between the thread's lines (unit-stride-k layout) or k values (otherwise)
  int32_t bsk;    // bytes per unit of k
  int32_t lim;    // lines of the block from the tile's first line on
};

template <bool Contig, int LINES>
struct PanelMap {
  static constexpr int E = LINES * BK / 256;      // elements per thread
  static constexpr int KSTEP = 256 / LINES;        // k step between the thread's elements (general layout)
  static __device__ __forceinline__ int l0(int tid) { return Contig ? tid >> 2 : tid % LINES; }
  static __device__ __forceinline__ int kk0(int tid) { return Contig ? 4 * (tid & 3) : tid / LINES; }
};

template <typename T, bool Contig, int LINES>
__device__ __forceinline__ void panel_view(PanelView& v, int tid, const T* __restrict__ block, int line0, int nline,
                                           int klen, int64_t s_line, int64_t s_k) {
  typedef PanelMap<Contig, LINES> M;
  // elements spanned by the block (0 for an empty one: then every element reads as zero)
  const int64_t extent = nline > 0 && klen > 0 ? (int64_t)(nline - 1) * s_line + (int64_t)(klen - 1) * s_k + 1 : 0;
  v.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(block), (short)0, (int)(extent * (int64_t)sizeof(T)),
                                             0x00020000);
  v.boff0 = (int32_t)(sizeof(T) * ((int64_t)(line0 + M::l0(tid)) * s_line + (int64_t)M::kk0(tid) * s_k));
  v.bstep = (int32_t)(sizeof(T) * (Contig ? 64 * s_line : M::KSTEP * s_k));
  v.bsk = (int32_t)(sizeof(T) * s_k);
  v.lim = nline - line0;
}

template <typename T>
__device__ __forceinline__ T buffer_load1(__amdgpu_buffer_rsrc_t r, int32_t off) {
  if constexpr (sizeof(T) == 8) return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// The thread's elements of the panel at k0 (valid = false past the tile's last panel: zeros).  kl = k of the
// thread's first element left in the pair (unit-stride-k layout: elements i >= kl of a line are zeroed at the store).
template <typename T, bool Contig, int LINES>
__device__ __forceinline__ void panel_load(T (&r)[(PanelMap<Contig, LINES>::E)], int& kl, const PanelView& v, int tid,
                                           int k0, int klen, bool valid) {
  typedef PanelMap<Contig, LINES> M;
  const int32_t base = v.boff0 + k0 * v.bsk;
  kl = klen - k0 - M::kk0(tid);
  if constexpr (Contig) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef double d2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int run = 0; run < LINES / 64; ++run) {
      const bool ok = valid && M::l0(tid) + 64 * run < v.lim && kl > 0;
      const int32_t off = ok ? base + run * v.bstep : kOutOfBlock;
      if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const d2 x =
              __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, ok ? off + 16 * h : off, 0, 0));
          r[4 * run + 2 * h] = x[0];
          r[4 * run + 2 * h + 1] = x[1];
        }
      } else {
        const f4 x = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, off, 0, 0));
#pragma unroll
        for (int i = 0; i < 4; ++i) r[4 * run + i] = x[i];
      }
    }
  } else {
    const bool lok = valid && M::l0(tid) < v.lim;
#pragma unroll
    for (int i = 0; i < M::E; ++i) {
      const int32_t off = (lok && M::KSTEP * i < kl) ? base + i * v.bstep : kOutOfBlock;
      r[i] = buffer_load1<T>(v.rsrc, off);
    }
  }
}

template <typename T, bool Contig, int LINES, int LD>
__device__ __forceinline__ void panel_store(T (*S)[LD], const T (&r)[(PanelMap<Contig, LINES>::E)], int kl, int tid) {
  typedef PanelMap<Contig, LINES> M;
  const int line = M::l0(tid), k = M::kk0(tid);
  if constexpr (Contig) {
#pragma unroll
    for (int run = 0; run < LINES / 64; ++run)
#pragma unroll
      for (int i = 0; i < 4; ++i) S[k + i][line + 64 * run] = i < kl ? r[4 * run + i] : T(0);
  } else {
#pragma unroll
    for (int i = 0; i < M::E; ++i) S[k + M::KSTEP * i][line] = r[i];
  }
}

// 64x64 tiles of large blocks.  The (pair, 16-wide k panel) sequence of a tile (or of its pair range, when the
// reduction is split over several work-groups) is software-pipelined: the next panel is loaded into registers while
// the MFMAs run on the current one from one of two LDS buffers, one barrier per panel.  A tile with slot >= 0 writes
// its unscaled partial sum to partial[slot] (64x64, row-major); gemm_reduce_kernel adds the slots of a tile in order.
// One kernel per operand layout (CA, CB: unit stride along k for every pair of the task's A, B).
template <typename T, typename TC, bool Tab, bool CA, bool CB>
__global__ __launch_bounds__(256) void gemm_grouped_kernel(const BigTile* __restrict__ tiles, int ntiles,
                                                           const GemmTask* __restrict__ tasks,
                                                           const GemmPair* __restrict__ pairs, Operands ops,
                                                           double alpha, int accumulate, T* __restrict__ partial,
                                                           int xchunk) {
  typedef typename Mfma<T>::acc_t acc_t;
  typedef PanelMap<CA, BM> MA;
  typedef PanelMap<CB, BN> MB;
  __shared__ T As[2][BK][LDP];
  __shared__ T Bs[2][BK][LDP];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const BigTile tl = tiles[xcd_item(w, gridDim.x, ntiles, xchunk)];
    const GemmTask tk = tasks[tl.task];
    const T* __restrict__ A = opnd<const T, Tab>(ops.a, tk.inst);
    const T* __restrict__ B = opnd<const T, Tab>(ops.b, tk.inst);
    // 16x16 sub-tiles of this wave that lie wholly in the padding of an edge tile skip their MFMAs (uniform per wave)
    const bool live_r0 = tl.row0 + wr * 32 < tk.m, live_r1 = tl.row0 + wr * 32 + 16 < tk.m;
    const bool live_c0 = tl.col0 + wc * 32 < tk.n, live_c1 = tl.col0 + wc * 32 + 16 < tk.n;

    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // the load cursor (cp, ck0) walks the tile's (pair, k panel) sequence; va, vb view the cursor's pair
    int cp = tl.pair_lo, ck0 = 0;
    GemmPair cpr = pairs[cp];
    PanelView va, vb;
    panel_view<T, CA, BM>(va, tid, A + cpr.a_off, tl.row0, tk.m, cpr.k, cpr.sa_m, cpr.sa_k);
    panel_view<T, CB, BN>(vb, tid, B + cpr.b_off, tl.col0, tk.n, cpr.k, cpr.sb_n, cpr.sb_k);
    bool cur = true;  // the cursor is on a panel of the tile (uniform)
    T ra[MA::E], rb[MB::E];
    int kla, klb;
    auto load = [&]() {
      panel_load<T, CA, BM>(ra, kla, va, tid, ck0, cpr.k, cur);
      panel_load<T, CB, BN>(rb, klb, vb, tid, ck0, cpr.k, cur);
      if (!cur) return;
      ck0 += BK;
      if (ck0 < cpr.k) return;
      ck0 = 0;
      if (++cp >= tl.pair_hi) {
        cur = false;
        return;
      }
      cpr = pairs[cp];
      panel_view<T, CA, BM>(va, tid, A + cpr.a_off, tl.row0, tk.m, cpr.k, cpr.sa_m, cpr.sa_k);
      panel_view<T, CB, BN>(vb, tid, B + cpr.b_off, tl.col0, tk.n, cpr.k, cpr.sb_n, cpr.sb_k);
    };
    load();
    panel_store<T, CA, BM, LDP>(As[0], ra, kla, tid);
    panel_store<T, CB, BN, LDP>(Bs[0], rb, klb, tid);
    __syncthreads();
    int buf = 0;
    while (true) {
      const bool more = cur;  // a next panel (uniform)
      load();                 // zeros past the tile's last panel
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const T a0 = As[buf][kk + lk][wr * 32 + lr];
        const T a1 = As[buf][kk + lk][wr * 32 + 16 + lr];
        const T b0 = Bs[buf][kk + lk][wc * 32 + lr];
        const T b1 = Bs[buf][kk + lk][wc * 32 + 16 + lr];
        if (live_r0 && live_c0) acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
        if (live_r0 && live_c1) acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
        if (live_r1 && live_c0) acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
        if (live_r1 && live_c1) acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
      }
      if (!more) break;
      panel_store<T, CA, BM, LDP>(As[buf ^ 1], ra, kla, tid);
      panel_store<T, CB, BN, LDP>(Bs[buf ^ 1], rb, klb, tid);
      __syncthreads();
      buf ^= 1;
    }

    if (tl.slot < 0) {
      TC* __restrict__ C = opnd<TC, Tab>(ops.c, tk.inst);
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = tl.row0 + wr * 32 + mi * 16 + Mfma<T>::row(lk, r);
            const int col = tl.col0 + wc * 32 + ni * 16 + lr;
            if (row < tk.m && col < tk.n) {
              const int64_t idx = tk.c_off + (int64_t)row * tk.n + col;
              const double v = alpha * (double)acc[mi][ni][r];
              C[idx] = accumulate ? (TC)((double)C[idx] + v) : (TC)v;
            }
          }
    } else {
      T* __restrict__ P = partial + (int64_t)tl.slot * (BM * BN);
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = wr * 32 + mi * 16 + Mfma<T>::row(lk, r);
            const int col = wc * 32 + ni * 16 + lr;
            P[row * BN + col] = acc[mi][ni][r];
          }
    }
    __syncthreads();  // every wave is done with the LDS buffers before the next tile's prologue
  }
}

// Sum of the partial tiles of split reductions, in slot order (deterministic), into C: one work-group per 16-row
// strip of a tile (kRedStrips per tile), four consecutive elements per thread read as one vector, the slot loop
// unrolled so that several loads are in flight.
constexpr int kRedStrips = BM / 16;
template <typename T, typename TC, bool Tab>
__global__ __launch_bounds__(256) void gemm_reduce_kernel(const RedTile* __restrict__ red, int nred,
                                                          const GemmTask* __restrict__ tasks,
                                                          const T* __restrict__ partial, Operands ops, double alpha,
                                                          int accumulate) {
  typedef typename Mfma<T>::acc_t vec_t;
  for (int w = blockIdx.x; w < nred * kRedStrips; w += gridDim.x) {
    const RedTile rt = red[w / kRedStrips];
    const GemmTask tk = tasks[rt.task];
    const int lr = (w % kRedStrips) * 16 + (threadIdx.x >> 4), lc = (threadIdx.x & 15) * 4;
    const int row = rt.row0 + lr, col = rt.col0 + lc;
    if (row >= tk.m || col >= tk.n) continue;
    const vec_t* __restrict__ P =
        reinterpret_cast<const vec_t*>(partial + (int64_t)rt.slot0 * (BM * BN) + lr * BN + lc);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll 4
    for (int q = 0; q < rt.nslots; ++q) {
      const vec_t v = P[(int64_t)q * (BM * BN / 4)];
      s0 += (double)v[0];
      s1 += (double)v[1];
      s2 += (double)v[2];
      s3 += (double)v[3];
    }
    TC* __restrict__ C = opnd<TC, Tab>(ops.c, tk.inst);
    const int64_t idx = tk.c_off + (int64_t)row * tk.n + col;
    const double sum[4] = {s0, s1, s2, s3};
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (col + j < tk.n) {
        const double v = alpha * sum[j];
        C[idx + j] = accumulate ? (TC)((double)C[idx + j] + v) : (TC)v;
      }
  }
}

static inline int groups(int items, int max_groups) { return max_groups > 0 && items > max_groups ? max_groups : items; }

template <typename T, typename TC, bool Tab, bool CA, bool CB>
void launch_gemm_group(const BigTile* tiles, int ntiles, const GemmTask* tasks, const GemmPair* pairs, Operands ops,
                       double alpha, int accumulate, T* part, int max_groups, hipStream_t s) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL((gemm_grouped_kernel<T, TC, Tab, CA, CB>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s,
                     tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part, xcd_chunk());
}

template <typename T, typename TC, bool Tab>
void launch_gemm_groups(const BigTile* tiles, const int (&n)[4], const GemmTask* tasks, const GemmPair* pairs,
                        Operands ops, double alpha, int accumulate, T* part, int max_groups, hipStream_t s) {
  launch_gemm_group<T, TC, Tab, true, true>(tiles, n[0], tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  tiles += n[0];
  launch_gemm_group<T, TC, Tab, true, false>(tiles, n[1], tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  tiles += n[1];
  launch_gemm_group<T, TC, Tab, false, true>(tiles, n[2], tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  tiles += n[2];
  launch_gemm_group<T, TC, Tab, false, false>(tiles, n[3], tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
}

template <typename T, typename TC>
void launch_gemm(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                 const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                 hipStream_t s, int max_groups) {
  if (ntiles[0] + ntiles[1] + ntiles[2] + ntiles[3] <= 0) return;
  T* part = static_cast<T*>(partial);
  if (ops.table) launch_gemm_groups<T, TC, true>(tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  else launch_gemm_groups<T, TC, false>(tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  if (nred <= 0) return;
  if (ops.table)
    hipLaunchKernelGGL((gemm_reduce_kernel<T, TC, true>), dim3(groups(nred * kRedStrips, max_groups)), dim3(256), 0,
                       s, red, nred, tasks, part, ops, alpha, accumulate);
  else
    hipLaunchKernelGGL((gemm_reduce_kernel<T, TC, false>), dim3(groups(nred * kRedStrips, max_groups)), dim3(256), 0,
                       s, red, nred, tasks, part, ops, alpha, accumulate);
}

}  // namespace mpx
