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

// Panel staging of the 64x64 kernel: each of the 256 threads moves 4 elements of the 64x16 A panel and 4 of the 16x64
// B panel.  The element -> thread map follows the operand's unit-stride direction (k for A stored as used, rows for A
// used transposed; columns for B as stored, k for B transposed) so that global loads are coalesced either way.
template <typename T>
__device__ __forceinline__ void load_a(T (&r)[4], const T* __restrict__ Ap, const GemmPair& pr, int row0, int m, int k0,
                                       int tid) {
  if (pr.sa_k == 1) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j, kk = e & 15, i = e >> 4;
      const int gi = row0 + i, gk = k0 + kk;
      r[j] = (gi < m && gk < pr.k) ? Ap[(int64_t)gi * pr.sa_m + gk] : T(0);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j, i = e & 63, kk = e >> 6;
      const int gi = row0 + i, gk = k0 + kk;
      r[j] = (gi < m && gk < pr.k) ? Ap[(int64_t)gi * pr.sa_m + (int64_t)gk * pr.sa_k] : T(0);
    }
  }
}
template <typename T>
__device__ __forceinline__ void store_a(T (*As)[LDP], const T (&r)[4], bool contig_k, int tid) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int e = tid + 256 * j;
    if (contig_k) As[e & 15][e >> 4] = r[j];
    else As[e >> 6][e & 63] = r[j];
  }
}
template <typename T>
__device__ __forceinline__ void load_b(T (&r)[4], const T* __restrict__ Bp, const GemmPair& pr, int col0, int n, int k0,
                                       int tid) {
  if (pr.sb_n == 1) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j, jj = e & 63, kk = e >> 6;
      const int gj = col0 + jj, gk = k0 + kk;
      r[j] = (gj < n && gk < pr.k) ? Bp[(int64_t)gk * pr.sb_k + gj] : T(0);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j, kk = e & 15, jj = e >> 4;
      const int gj = col0 + jj, gk = k0 + kk;
      r[j] = (gj < n && gk < pr.k) ? Bp[(int64_t)gk * pr.sb_k + (int64_t)gj * pr.sb_n] : T(0);
    }
  }
}
template <typename T>
__device__ __forceinline__ void store_b(T (*Bs)[LDP], const T (&r)[4], bool contig_n, int tid) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int e = tid + 256 * j;
    if (contig_n) Bs[e >> 6][e & 63] = r[j];
    else Bs[e & 15][e >> 4] = r[j];
  }
}

// 64x64 tiles of large blocks.  The (pair, 16-wide k panel) sequence of a tile (or of its pair range, when the
// reduction is split over several work-groups) is software-pipelined: the next panel is loaded into registers while
// the MFMAs run on the current one from one of two LDS buffers, one barrier per panel.  A tile with slot >= 0 writes
// its unscaled partial sum to partial[slot] (64x64, row-major); gemm_reduce_kernel adds the slots of a tile in order.
template <typename T, typename TC, bool Tab>
__global__ __launch_bounds__(256) void gemm_grouped_kernel(const BigTile* __restrict__ tiles, int ntiles,
                                                           const GemmTask* __restrict__ tasks,
                                                           const GemmPair* __restrict__ pairs, Operands ops,
                                                           double alpha, int accumulate, T* __restrict__ partial,
                                                           int xchunk) {
  typedef typename Mfma<T>::acc_t acc_t;
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

    // the tile's k panels over its pair range, counted first so that the panel loop is a plain counted loop (the
    // accumulators then stay in the MFMA registers instead of being copied around a mid-loop exit)
    int npanel = 0;  // every pair takes at least one panel (k = 0 pairs contribute a zero panel)
    for (int q = tl.pair_lo; q < tl.pair_hi; ++q) npanel += max(1, (pairs[q].k + BK - 1) / BK);
    // prologue: first panel into buffer 0
    int p = tl.pair_lo, k0 = 0;
    GemmPair pr = pairs[p];
    T ra[4], rb[4];
    load_a(ra, A + pr.a_off, pr, tl.row0, tk.m, 0, tid);
    load_b(rb, B + pr.b_off, pr, tl.col0, tk.n, 0, tid);
    store_a(As[0], ra, pr.sa_k == 1, tid);
    store_b(Bs[0], rb, pr.sb_n == 1, tid);
    __syncthreads();
    int buf = 0;
    for (int it = 0; it < npanel; ++it) {
      int np = p, nk0 = k0 + BK;
      GemmPair npr = pr;
      if (nk0 >= pr.k) {
        np = p + 1;
        nk0 = 0;
        if (np < tl.pair_hi) npr = pairs[np];
      }
      const bool more = it + 1 < npanel;  // uniform over the work-group
      if (more) {
        load_a(ra, A + npr.a_off, npr, tl.row0, tk.m, nk0, tid);
        load_b(rb, B + npr.b_off, npr, tl.col0, tk.n, nk0, tid);
      }
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
      if (more) {
        store_a(As[buf ^ 1], ra, npr.sa_k == 1, tid);
        store_b(Bs[buf ^ 1], rb, npr.sb_n == 1, tid);
      }
      __syncthreads();
      buf ^= 1;
      p = np;
      k0 = nk0;
      pr = npr;
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

template <typename T, typename TC>
void launch_gemm(const BigTile* tiles, int ntiles, const RedTile* red, int nred, void* partial, const GemmTask* tasks,
                 const GemmPair* pairs, Operands ops, double alpha, int accumulate, hipStream_t s, int max_groups) {
  if (ntiles <= 0) return;
  T* part = static_cast<T*>(partial);
  const int xc = xcd_chunk();
  if (ops.table)
    hipLaunchKernelGGL((gemm_grouped_kernel<T, TC, true>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s, tiles,
                       ntiles, tasks, pairs, ops, alpha, accumulate, part, xc);
  else
    hipLaunchKernelGGL((gemm_grouped_kernel<T, TC, false>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s, tiles,
                       ntiles, tasks, pairs, ops, alpha, accumulate, part, xc);
  if (nred <= 0) return;
  if (ops.table)
    hipLaunchKernelGGL((gemm_reduce_kernel<T, TC, true>), dim3(groups(nred * kRedStrips, max_groups)), dim3(256), 0,
                       s, red, nred, tasks, part, ops, alpha, accumulate);
  else
    hipLaunchKernelGGL((gemm_reduce_kernel<T, TC, false>), dim3(groups(nred * kRedStrips, max_groups)), dim3(256), 0,
                       s, red, nred, tasks, part, ops, alpha, accumulate);
}

}  // namespace mpx
