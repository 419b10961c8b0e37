// 128x128 grouped-GEMM tiles for large blocks (both sides >= GemmPolicy::big_tile_min), FP64 and FP32 operands; device
// code, included only by .hip files (FP64 instantiated in kernels_f64.hip, FP32 in kernels.hip).
//
// One 256-thread work-group per 128x128 C tile, 4 waves each owning a 64x64 quarter computed as 4x4
// v_mfma_{f64,f32}_16x16x4 tiles: per 4-wide k step a wave reads 4 A and 4 B operands from LDS and issues 16 MFMAs,
// four times the MFMAs per operand byte of the 64x64 tiles.  A and B panels (128x16, 16x128) are staged in LDS k-major
// and double-buffered; each thread moves 8 elements of each per panel.  The element -> thread map follows the
// operand's unit-stride direction, and each thread's element offsets inside the block and its row (column) validity
// are set up once per pair, so a panel load costs a few integer operations per element.  16x16 sub-tiles that lie
// wholly in the padding of an edge tile skip their MFMAs.
#pragma once

#include "gemm_tiles.hpp"

namespace mpx {

constexpr int B2M = 128, B2N = 128, LDP2 = 128 + 4;

// Per-thread view of one operand of one pair.  Element j of the thread's 8 sits at line l0 + j*lstep (row of A or
// column of B) and k kk0 + j*kstep of the panel, at offset off0 + j*dstep inside the block for panel k0 = 0 (an
// arithmetic sequence in j, so the view stays in a few registers); kinc = offset increment per unit of k0, lim =
// lines left in the block from the tile's first line.
struct PanelView {
  int32_t off0, dstep;
  int32_t kk0, kstep;
  int32_t l0, lstep;
  int32_t kinc, lim;
};

// contig_k: the operand's unit stride runs along k (A as stored with sa_k == 1, B transposed)
__device__ __forceinline__ void panel_view(PanelView& v, bool contig_k, int tid, int line0, int nline, int64_t s_line,
                                           int64_t s_k) {
  if (contig_k) {  // 16 threads read 16 consecutive k of one line
    v.kk0 = tid & 15;
    v.kstep = 0;
    v.l0 = tid >> 4;
    v.lstep = 16;
  } else {         // 128 threads read 128 consecutive lines at one k
    v.kk0 = tid >> 7;
    v.kstep = 2;
    v.l0 = tid & 127;
    v.lstep = 0;
  }
  v.off0 = (int32_t)((int64_t)(line0 + v.l0) * s_line + (int64_t)v.kk0 * s_k);
  v.dstep = (int32_t)(v.lstep * s_line + v.kstep * s_k);
  v.kinc = (int32_t)s_k;
  v.lim = nline - line0;
}

template <typename T>
__device__ __forceinline__ void panel_load(T (&r)[8], const T* __restrict__ P, const PanelView& v, int k0, int klen) {
  const T* __restrict__ q = P + (v.off0 + k0 * v.kinc);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const bool ok = v.l0 + j * v.lstep < v.lim && k0 + v.kk0 + j * v.kstep < klen;
    r[j] = ok ? q[j * v.dstep] : T(0);
  }
}

template <typename T>
__device__ __forceinline__ void panel_store(T (*S)[LDP2], const T (&r)[8], const PanelView& v) {
#pragma unroll
  for (int j = 0; j < 8; ++j) S[v.kk0 + j * v.kstep][v.l0 + j * v.lstep] = r[j];
}

// FP64: one work-group per CU (each wave has the whole 512-entry register file for its 128 accumulator registers and
// the panel prefetch; the 64-cycle FP64 MFMAs of one panel cover the prefetch latency); FP32: two
template <typename T, typename TC, bool Tab>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 1 : 2) void gemm128_kernel(const BigTile* __restrict__ tiles, int ntiles,
                                                          const GemmTask* __restrict__ tasks,
                                                          const GemmPair* __restrict__ pairs, Operands ops,
                                                          double alpha, int accumulate, T* __restrict__ partial,
                                                          int xchunk) {
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[2][BK][LDP2];
  __shared__ T Bs[2][BK][LDP2];
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
    bool live_r[4], live_c[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      live_r[i] = tl.row0 + wr * 64 + i * 16 < tk.m;
      live_c[i] = tl.col0 + wc * 64 + i * 16 < tk.n;
    }
    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    int npanel = 0;
    for (int q = tl.pair_lo; q < tl.pair_hi; ++q) npanel += max(1, (pairs[q].k + BK - 1) / BK);
    int p = tl.pair_lo, k0 = 0;
    GemmPair pr = pairs[p];
    PanelView va, vb;
    panel_view(va, pr.sa_k == 1, tid, tl.row0, tk.m, pr.sa_m, pr.sa_k);
    panel_view(vb, pr.sb_n != 1, tid, tl.col0, tk.n, pr.sb_n, pr.sb_k);
    T ra[8], rb[8];
    panel_load(ra, A + pr.a_off, va, 0, pr.k);
    panel_load(rb, B + pr.b_off, vb, 0, pr.k);
    panel_store(As[0], ra, va);
    panel_store(Bs[0], rb, vb);
    __syncthreads();
    int buf = 0;
    // va, vb describe the panel held in ra, rb (the next one once it is loaded)
    for (int it = 0; it < npanel; ++it) {
      int np = p, nk0 = k0 + BK;
      if (nk0 >= pr.k) {
        np = p + 1;
        nk0 = 0;
      }
      const bool more = it + 1 < npanel;  // uniform over the work-group
      if (more) {
        if (np != p) {
          pr = pairs[np];
          panel_view(va, pr.sa_k == 1, tid, tl.row0, tk.m, pr.sa_m, pr.sa_k);
          panel_view(vb, pr.sb_n != 1, tid, tl.col0, tk.n, pr.sb_n, pr.sb_k);
        }
        panel_load(ra, A + pr.a_off, va, nk0, pr.k);
        panel_load(rb, B + pr.b_off, vb, nk0, pr.k);
      }
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        T a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = As[buf][kk + lk][wr * 64 + i * 16 + lr];
          b[i] = Bs[buf][kk + lk][wc * 64 + i * 16 + lr];
        }
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < 4; ++ni)
            if (live_r[mi] && live_c[ni]) acc[mi][ni] = Mfma<T>::op(a[mi], b[ni], acc[mi][ni]);
      }
      if (more) {
        panel_store(As[buf ^ 1], ra, va);
        panel_store(Bs[buf ^ 1], rb, vb);
      }
      __syncthreads();
      buf ^= 1;
      p = np;
      k0 = nk0;
    }

    if (tl.slot < 0) {
      TC* __restrict__ C = opnd<TC, Tab>(ops.c, tk.inst);
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = tl.row0 + wr * 64 + mi * 16 + Mfma<T>::row(lk, r);
            const int col = tl.col0 + wc * 64 + ni * 16 + lr;
            if (row < tk.m && col < tk.n) {
              const int64_t idx = tk.c_off + (int64_t)row * tk.n + col;
              const double v = alpha * (double)acc[mi][ni][r];
              C[idx] = accumulate ? (TC)((double)C[idx] + v) : (TC)v;
            }
          }
    } else {
      T* __restrict__ P = partial + (int64_t)tl.slot * (B2M * B2N);
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = wr * 64 + mi * 16 + Mfma<T>::row(lk, r);
            const int col = wc * 64 + ni * 16 + lr;
            P[row * B2N + col] = acc[mi][ni][r];
          }
    }
    __syncthreads();  // every wave is done with the LDS buffers before the next tile's prologue
  }
}

// Sum of the 128x128 partial tiles of split reductions in slot order (deterministic): one work-group per 8-row strip,
// four consecutive elements per thread.
constexpr int kRedStrips2 = B2M / 8;
template <typename T, typename TC, bool Tab>
__global__ __launch_bounds__(256) void gemm_reduce128_kernel(const RedTile* __restrict__ red, int nred,
                                                             const GemmTask* __restrict__ tasks,
                                                             const T* __restrict__ partial, Operands ops,
                                                             double alpha, int accumulate) {
  typedef typename Mfma<T>::acc_t vec_t;
  for (int w = blockIdx.x; w < nred * kRedStrips2; w += gridDim.x) {
    const RedTile rt = red[w / kRedStrips2];
    const GemmTask tk = tasks[rt.task];
    const int lr = (w % kRedStrips2) * 8 + (threadIdx.x >> 5), lc = (threadIdx.x & 31) * 4;
    const int row = rt.row0 + lr, col = rt.col0 + lc;
    if (row >= tk.m || col >= tk.n) continue;
    const vec_t* __restrict__ P =
        reinterpret_cast<const vec_t*>(partial + (int64_t)rt.slot0 * (B2M * B2N) + lr * B2N + lc);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll 4
    for (int q = 0; q < rt.nslots; ++q) {
      const vec_t v = P[(int64_t)q * (B2M * B2N / 4)];
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

template <typename T, typename TC>
void launch_gemm128(const BigTile* tiles, int ntiles, const RedTile* red, int nred, void* partial,
                    const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                    hipStream_t s, int max_groups) {
  if (ntiles <= 0) return;
  T* part = static_cast<T*>(partial);
  const int xc = xcd_chunk();
  if (ops.table)
    hipLaunchKernelGGL((gemm128_kernel<T, TC, true>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s, tiles,
                       ntiles, tasks, pairs, ops, alpha, accumulate, part, xc);
  else
    hipLaunchKernelGGL((gemm128_kernel<T, TC, false>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s, tiles,
                       ntiles, tasks, pairs, ops, alpha, accumulate, part, xc);
  if (nred <= 0) return;
  if (ops.table)
    hipLaunchKernelGGL((gemm_reduce128_kernel<T, TC, true>), dim3(groups(nred * kRedStrips2, max_groups)), dim3(256),
                       0, s, red, nred, tasks, part, ops, alpha, accumulate);
  else
    hipLaunchKernelGGL((gemm_reduce128_kernel<T, TC, false>), dim3(groups(nred * kRedStrips2, max_groups)), dim3(256),
                       0, s, red, nred, tasks, part, ops, alpha, accumulate);
}

}  // namespace mpx
