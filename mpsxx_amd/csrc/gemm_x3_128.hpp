// 128x128 split-BF16 (x3) grouped-GEMM tiles for large blocks of FP32 operands; device code, included only by
// kernels_x3.hip.
//
// Every FP32 operand element is split on its way into LDS as x = hi + lo, hi = bf16(x), lo = bf16(x - hi), and each
// product is formed as lo*hi + hi*lo + hi*hi by three v_mfma_f32_16x16x32_bf16 with FP32 accumulation (the lo*lo term,
// ~2^-16 relative, is dropped).  One 256-thread work-group per 128x128 C tile; 4 waves each own a 64x64 quarter as
// 4x4 MFMA tiles, so per 32-wide k panel a wave reads 16 operand vectors (hi and lo of 4 A rows-blocks and 4 B
// column blocks) for 48 MFMAs.
//
// LDS holds each panel as [line][k] bf16 (line = row of A or column of B), 64 bytes per line in four 16-byte chunks:
// MFMA lane l reads line l%16 of its sub-tile at k 8*(l/16)..+7 as one 16-byte vector.  Chunk c of line r is stored
// at chunk c ^ ((r >> 2) & 3) (XOR swizzle), so that the 16 lanes of a read, 16 consecutive lines at one chunk, hit 16
// distinct 16-byte bank groups without padding.  Two panel buffers (hi and lo of A and B, 64 KB) leave room for two
// work-groups per CU; the next panel is loaded into registers while the current one is multiplied.
//
// Panel loads: each thread moves 16 elements of A and 16 of B.  For an operand whose unit stride runs along k, a
// thread takes two runs of 8 consecutive k (lines t/4 and t/4 + 64, k 8*(t%4)..+7), read as two 4-wide vector loads
// each; otherwise it takes 16 consecutive k of line t%128 (k half t/128), one element per load, 128 consecutive lines
// per load instruction over the work-group.  Panels wholly inside the block skip the bounds checks.
#pragma once

#include "gemm128.hpp"

namespace mpx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));  // 4-wide load at 4-byte alignment

constexpr int X2K = 32;  // k panel: one bf16 MFMA depth

struct X3View {
  int32_t off0;  // offset of the thread's first element for k0 = 0
  int32_t rstep; // offset of run 1 relative to run 0
  int32_t s_k;   // offset step per unit of k
  int32_t l0, lrun;  // line of run 0 relative to the tile's first line, line offset of run 1
  int32_t kk0, krun; // k of run 0 in the panel, k offset of run 1
  int32_t lim;       // lines of the block from the tile's first line on
  bool contig;       // unit stride along k: vector loads
};

__device__ __forceinline__ void x3_view(X3View& v, bool contig, int tid, int line0, int nline, int64_t s_line,
                                        int64_t s_k) {
  v.contig = contig;
  if (contig) {
    v.l0 = tid >> 2;
    v.lrun = 64;
    v.kk0 = 8 * (tid & 3);
    v.krun = 0;
  } else {
    v.l0 = tid & 127;
    v.lrun = 0;
    v.kk0 = 16 * (tid >> 7);
    v.krun = 8;
  }
  v.off0 = (int32_t)((int64_t)(line0 + v.l0) * s_line + (int64_t)v.kk0 * s_k);
  v.rstep = (int32_t)(v.lrun * s_line + v.krun * s_k);
  v.s_k = (int32_t)s_k;
  v.lim = nline - line0;
}

// the thread's 16 elements of the panel starting at k0 (zero outside the block)
__device__ __forceinline__ void x3_panel_load(float (&r)[16], const float* __restrict__ P, const X3View& v, int k0,
                                              int klen) {
  const float* __restrict__ q = P + (v.off0 + k0 * v.s_k);
  const bool full = v.lim >= B2M && k0 + X2K <= klen;  // uniform over the work-group
  if (full && v.contig) {
#pragma unroll
    for (int run = 0; run < 2; ++run) {
      const f32x4u x0 = *reinterpret_cast<const f32x4u*>(q + run * v.rstep);
      const f32x4u x1 = *reinterpret_cast<const f32x4u*>(q + run * v.rstep + 4);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        r[8 * run + j] = x0[j];
        r[8 * run + 4 + j] = x1[j];
      }
    }
  } else if (full) {
#pragma unroll
    for (int run = 0; run < 2; ++run)
#pragma unroll
      for (int j = 0; j < 8; ++j) r[8 * run + j] = q[run * v.rstep + j * v.s_k];
  } else {
#pragma unroll
    for (int run = 0; run < 2; ++run) {
      const bool lok = v.l0 + run * v.lrun < v.lim;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const bool ok = lok && k0 + v.kk0 + run * v.krun + j < klen;
        r[8 * run + j] = ok ? q[run * v.rstep + j * v.s_k] : 0.f;
      }
    }
  }
}

__device__ __forceinline__ int x3_chunk(int line, int chunk) { return (chunk ^ ((line >> 2) & 3)) * 8; }

// split the thread's elements and store them as two 16-byte vectors per run into the hi and lo panels
__device__ __forceinline__ void x3_panel_store(__bf16 (*Hi)[X2K], __bf16 (*Lo)[X2K], const float (&r)[16],
                                               const X3View& v) {
#pragma unroll
  for (int run = 0; run < 2; ++run) {
    bf16x8 h, l;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float x = r[8 * run + j];
      const __bf16 hb = (__bf16)x;
      h[j] = hb;
      l[j] = (__bf16)(x - (float)hb);
    }
    const int line = v.l0 + run * v.lrun, c = x3_chunk(line, (v.kk0 + run * v.krun) >> 3);
    *reinterpret_cast<bf16x8*>(&Hi[line][c]) = h;
    *reinterpret_cast<bf16x8*>(&Lo[line][c]) = l;
  }
}

template <typename TC, bool Tab>
__global__ __launch_bounds__(256, 2) void gemm_x3_128_kernel(const BigTile* __restrict__ tiles, int ntiles,
                                                             const GemmTask* __restrict__ tasks,
                                                             const GemmPair* __restrict__ pairs, Operands ops,
                                                             double alpha, int accumulate,
                                                             float* __restrict__ partial, int xchunk) {
  __shared__ __attribute__((aligned(16))) __bf16 Ah[2][B2M][X2K];
  __shared__ __attribute__((aligned(16))) __bf16 Al[2][B2M][X2K];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[2][B2N][X2K];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[2][B2N][X2K];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const BigTile tl = tiles[xcd_item(w, gridDim.x, ntiles, xchunk)];
    const GemmTask tk = tasks[tl.task];
    const float* __restrict__ A = opnd<const float, Tab>(ops.a, tk.inst);
    const float* __restrict__ B = opnd<const float, Tab>(ops.b, tk.inst);
    bool live_r[4], live_c[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      live_r[i] = tl.row0 + wr * 64 + i * 16 < tk.m;
      live_c[i] = tl.col0 + wc * 64 + i * 16 < tk.n;
    }
    flt4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = flt4{0, 0, 0, 0};

    int p = tl.pair_lo, k0 = 0;
    GemmPair pr = pairs[p];
    X3View va, vb;
    x3_view(va, pr.sa_k == 1, tid, tl.row0, tk.m, pr.sa_m, pr.sa_k);
    x3_view(vb, pr.sb_k == 1, tid, tl.col0, tk.n, pr.sb_n, pr.sb_k);
    float ra[16], rb[16];
    x3_panel_load(ra, A + pr.a_off, va, 0, pr.k);
    x3_panel_load(rb, B + pr.b_off, vb, 0, pr.k);
    x3_panel_store(Ah[0], Al[0], ra, va);
    x3_panel_store(Bh[0], Bl[0], rb, vb);
    __syncthreads();
    int buf = 0;
    // va, vb describe the panel held in ra, rb (the next one once it is loaded)
    while (true) {
      int np = p, nk0 = k0 + X2K;
      if (nk0 >= pr.k) {
        np = p + 1;
        nk0 = 0;
      }
      const bool more = np < tl.pair_hi;  // uniform over the work-group
      if (more) {
        if (np != p) {
          pr = pairs[np];
          x3_view(va, pr.sa_k == 1, tid, tl.row0, tk.m, pr.sa_m, pr.sa_k);
          x3_view(vb, pr.sb_k == 1, tid, tl.col0, tk.n, pr.sb_n, pr.sb_k);
        }
        x3_panel_load(ra, A + pr.a_off, va, nk0, pr.k);
        x3_panel_load(rb, B + pr.b_off, vb, nk0, pr.k);
      }
      bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ra_ = wr * 64 + i * 16 + lr, rb_ = wc * 64 + i * 16 + lr;
        ah[i] = *reinterpret_cast<const bf16x8*>(&Ah[buf][ra_][x3_chunk(ra_, lk)]);
        al[i] = *reinterpret_cast<const bf16x8*>(&Al[buf][ra_][x3_chunk(ra_, lk)]);
        bh[i] = *reinterpret_cast<const bf16x8*>(&Bh[buf][rb_][x3_chunk(rb_, lk)]);
        bl[i] = *reinterpret_cast<const bf16x8*>(&Bl[buf][rb_][x3_chunk(rb_, lk)]);
      }
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
          if (!(live_r[mi] && live_c[ni])) continue;  // wholly in the padding of an edge tile
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mi], bh[ni], acc[mi][ni], 0, 0, 0);
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bl[ni], acc[mi][ni], 0, 0, 0);
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bh[ni], acc[mi][ni], 0, 0, 0);
        }
      if (!more) break;
      x3_panel_store(Ah[buf ^ 1], Al[buf ^ 1], ra, va);
      x3_panel_store(Bh[buf ^ 1], Bl[buf ^ 1], rb, vb);
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
            const int row = tl.row0 + wr * 64 + mi * 16 + Mfma<float>::row(lk, r);
            const int col = tl.col0 + wc * 64 + ni * 16 + lr;
            if (row < tk.m && col < tk.n) {
              const int64_t idx = tk.c_off + (int64_t)row * tk.n + col;
              const double v = alpha * (double)acc[mi][ni][r];
              C[idx] = accumulate ? (TC)((double)C[idx] + v) : (TC)v;
            }
          }
    } else {
      float* __restrict__ P = partial + (int64_t)tl.slot * (B2M * B2N);
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = wr * 64 + mi * 16 + Mfma<float>::row(lk, r);
            const int col = wc * 64 + ni * 16 + lr;
            P[row * B2N + col] = acc[mi][ni][r];
          }
    }
    __syncthreads();  // every wave is done with the LDS buffers before the next tile's prologue
  }
}

template <typename TC>
void launch_gemm_x3_128(const BigTile* tiles, int ntiles, const RedTile* red, int nred, void* partial,
                        const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                        hipStream_t s, int max_groups) {
  if (ntiles <= 0) return;
  float* part = static_cast<float*>(partial);
  const int xc = xcd_chunk();
  if (ops.table)
    hipLaunchKernelGGL((gemm_x3_128_kernel<TC, true>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s, tiles,
                       ntiles, tasks, pairs, ops, alpha, accumulate, part, xc);
  else
    hipLaunchKernelGGL((gemm_x3_128_kernel<TC, false>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s, tiles,
                       ntiles, tasks, pairs, ops, alpha, accumulate, part, xc);
  if (nred <= 0) return;
  if (ops.table)
    hipLaunchKernelGGL((gemm_reduce128_kernel<float, TC, true>), dim3(groups(nred * kRedStrips2, max_groups)),
                       dim3(256), 0, s, red, nred, tasks, part, ops, alpha, accumulate);
  else
    hipLaunchKernelGGL((gemm_reduce128_kernel<float, TC, false>), dim3(groups(nred * kRedStrips2, max_groups)),
                       dim3(256), 0, s, red, nred, tasks, part, ops, alpha, accumulate);
}

}  // namespace mpx
