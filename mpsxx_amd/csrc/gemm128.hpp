// 128x128 grouped-GEMM tiles for large blocks (both sides >= GemmPolicy::big_tile_min), FP64 and FP32 operands; device
// code, included only by .hip files (FP32 instantiated in kernels.hip, FP64 in kernels_f64_128.hip).
//
// One 256-thread work-group per 128x128 C tile, 4 waves each owning a 64x64 quarter computed as 4x4
// v_mfma_{f64,f32}_16x16x4 tiles: per 4-wide k step a wave reads 4 A and 4 B operands from LDS and issues 16 MFMAs,
// four times the MFMAs per operand byte of the 64x64 tiles.  A and B panels (128x16, 16x128) are staged in LDS k-major
// and double-buffered; the next panel is loaded into registers while the MFMAs run on the current one.  Panels move
// through the buffer loads of gemm_tiles.hpp (8 elements of each operand per thread), one kernel per operand layout
// (the host groups the tiles by their task's layout flags, one launch per group).
// 16x16 sub-tiles that lie wholly in the padding of an edge tile skip their MFMAs.
#pragma once

#include <type_traits>

#include "gemm_tiles.hpp"

namespace mpx {

constexpr int B2M = 128, B2N = 128, LDP2 = 128 + 4;
// occupancy: FP64 two work-groups per CU (256 registers per lane: 128 for the accumulators, the panel prefetch, the
// operands), FP32 three (168 registers); more waves per SIMD hide the panel loads better (FP32: 13.0 ms with two, 11.7
// with three, 11.9 with four, which spills in the loop's set-up)
template <typename T, typename TC, bool Tab, bool CA, bool CB>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 2 : 3) void gemm128_kernel(const BigTile* __restrict__ tiles,
                                                                             int ntiles,
                                                                             const GemmTask* __restrict__ tasks,
                                                                             const GemmPair* __restrict__ pairs,
                                                                             Operands ops, double alpha,
                                                                             int accumulate, T* __restrict__ partial,
                                                                             int xchunk) {
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[2][BK][LDP2];
  __shared__ T Bs[2][BK][LDP2];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // the wave index is uniform across a wave: read it from the first lane, so that every condition derived from it is a
  // scalar branch, not an exec mask around each MFMA
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
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
    // every 16x16 sub-tile of the wave's quarter holds real elements (all but edge tiles): MFMAs without guards
    const bool full = live_r[3] && live_c[3];
    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // the load cursor (cp, ck0) walks the tile's (pair, k panel) sequence; va, vb view the cursor's pair
    int cp = tl.pair_lo, ck0 = 0;
    GemmPair cpr = pairs[cp];
    PanelView va, vb;
    panel_view<T, CA, B2M>(va, tid, A + cpr.a_off, tl.row0, tk.m, cpr.k, cpr.sa_m, cpr.sa_k);
    panel_view<T, CB, B2M>(vb, tid, B + cpr.b_off, tl.col0, tk.n, cpr.k, cpr.sb_n, cpr.sb_k);
    bool cur = true;  // the cursor is on a panel of the tile (uniform)
    T ra[8], rb[8];
    int kla, klb;
    auto load = [&]() {
      panel_load<T, CA, B2M>(ra, kla, va, tid, ck0, cpr.k, cur);
      panel_load<T, CB, B2M>(rb, klb, vb, tid, ck0, cpr.k, cur);
      if (!cur) return;
      ck0 += BK;
      if (ck0 < cpr.k) return;
      ck0 = 0;
      if (++cp >= tl.pair_hi) {
        cur = false;
        return;
      }
      cpr = pairs[cp];
      panel_view<T, CA, B2M>(va, tid, A + cpr.a_off, tl.row0, tk.m, cpr.k, cpr.sa_m, cpr.sa_k);
      panel_view<T, CB, B2M>(vb, tid, B + cpr.b_off, tl.col0, tk.n, cpr.k, cpr.sb_n, cpr.sb_k);
    };
    load();
    panel_store<T, CA, B2M, LDP2>(As[0], ra, kla, tid);
    panel_store<T, CB, B2M, LDP2>(Bs[0], rb, klb, tid);
    __syncthreads();
    // the panel loop, compiled twice: without MFMA guards for full tiles, with per-sub-tile guards for edge tiles (one
    // uniform branch per tile, so the MFMA sequence of a full tile is straight-line code)
    auto panels = [&](auto full_tile) {
      constexpr bool FULL = decltype(full_tile)::value;
      int buf = 0;
      while (true) {
        const bool more = cur;  // a next panel (uniform)
        load();                 // zeros past the tile's last panel
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
              if (FULL || (live_r[mi] && live_c[ni])) acc[mi][ni] = Mfma<T>::op(a[mi], b[ni], acc[mi][ni]);
        }
        if (!more) break;
        panel_store<T, CA, B2M, LDP2>(As[buf ^ 1], ra, kla, tid);
        panel_store<T, CB, B2M, LDP2>(Bs[buf ^ 1], rb, klb, tid);
        __syncthreads();
        buf ^= 1;
      }
    };
    if (full) panels(std::true_type{});
    else panels(std::false_type{});

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

template <typename T, typename TC, bool Tab, bool CA, bool CB>
void launch_gemm128_group(const BigTile* tiles, int ntiles, const GemmTask* tasks, const GemmPair* pairs, Operands ops,
                          double alpha, int accumulate, T* part, int max_groups, hipStream_t s) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL((gemm128_kernel<T, TC, Tab, CA, CB>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s, tiles,
                     ntiles, tasks, pairs, ops, alpha, accumulate, part, xcd_chunk());
}

template <typename T, typename TC, bool Tab>
void launch_gemm128_groups(const BigTile* tiles, const int (&n)[4], const GemmTask* tasks, const GemmPair* pairs,
                           Operands ops, double alpha, int accumulate, T* part, int max_groups, hipStream_t s) {
  launch_gemm128_group<T, TC, Tab, true, true>(tiles, n[0], tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  tiles += n[0];
  launch_gemm128_group<T, TC, Tab, true, false>(tiles, n[1], tasks, pairs, ops, alpha, accumulate, part, max_groups,
                                                s);
  tiles += n[1];
  launch_gemm128_group<T, TC, Tab, false, true>(tiles, n[2], tasks, pairs, ops, alpha, accumulate, part, max_groups,
                                                s);
  tiles += n[2];
  launch_gemm128_group<T, TC, Tab, false, false>(tiles, n[3], tasks, pairs, ops, alpha, accumulate, part, max_groups,
                                                 s);
}

template <typename T, typename TC>
void launch_gemm128(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                    const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                    hipStream_t s, int max_groups) {
  if (ntiles[0] + ntiles[1] + ntiles[2] + ntiles[3] <= 0) return;
  T* part = static_cast<T*>(partial);
  if (ops.table)
    launch_gemm128_groups<T, TC, true>(tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  else
    launch_gemm128_groups<T, TC, false>(tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  if (nred <= 0) return;
  if (ops.table)
    hipLaunchKernelGGL((gemm_reduce128_kernel<T, TC, true>), dim3(groups(nred * kRedStrips2, max_groups)), dim3(256),
                       0, s, red, nred, tasks, part, ops, alpha, accumulate);
  else
    hipLaunchKernelGGL((gemm_reduce128_kernel<T, TC, false>), dim3(groups(nred * kRedStrips2, max_groups)), dim3(256),
                       0, s, red, nred, tasks, part, ops, alpha, accumulate);
}

}  // namespace mpx
