// Split-BF16 (x3) grouped-GEMM tiles of FP32 operands, 64x64 (gemm_x3_kernel) and 128x128 (gemm_x3_128_kernel);
// device code, included only by kernels_x3.hip.
//
// Every FP32 operand element is split on its way into LDS as x = hi + lo, hi = bf16(x), lo = bf16(x - hi), and each
// product is formed as lo*hi + hi*lo + hi*hi by three v_mfma_f32_16x16x32_bf16 with FP32 accumulation (the lo*lo term,
// ~2^-16 relative, is dropped).  One 256-thread work-group per C tile; 4 waves each own a quarter of it as NxN MFMA
// tiles (N = 2 for 64x64, 4 for 128x128), so per 32-wide k panel a wave reads 4N operand vectors (hi and lo of its N A
// row blocks and N B column blocks) for 3N^2 MFMAs.
//
// LDS holds each panel as [line][k] bf16 (line = row of A or column of B), 64 bytes per line in four 16-byte chunks:
// MFMA lane l reads line l%16 of its sub-tile at k 8*(l/16)..+7 as one 16-byte vector.  Chunk c of line r is stored
// at chunk c ^ ((r >> 2) & 3) (XOR swizzle), so that the 16 lanes of a read, 16 consecutive lines at one chunk, spread
// over the banks without padding.  Two panel buffers of hi and lo of A and B (64 KB for 128x128 tiles, two work-groups
// per CU; 32 KB for 64x64).
//
// Panels flow global -> registers -> LDS one panel ahead of the MFMAs: while the MFMAs run on panel i from one LDS
// buffer, panel i+1 is loaded into registers and then stored into the other.  All loads are buffer loads against a
// resource covering the operand block: an element outside the block gets an offset beyond the resource's extent and
// reads as zero, so every panel (inside the block or on its edge, and the empty panel past the end of the tile) goes
// through the same loads and no branch surrounds a load; the compiler can then count the loads in flight exactly.  The kernel is specialised per operand
// layout (the host groups the tiles by their task's layout flags, all pairs of a task alike, one launch per group).
// Each thread moves LINES/8 elements of an operand per panel (LINES = tile edge), as runs of 8 consecutive k:
//   unit stride along k: runs on lines t/4 + 64*run (run < LINES/64) at k 8*(t%4)..+7, two 16-byte loads each; k past
//     the pair's end within a run is zeroed when the run is stored;
//   otherwise: runs on line t%LINES at k (LINES/8)*(t/LINES) + 8*run, one element per load, LINES consecutive lines
//     per load instruction over the work-group.
#pragma once

#include "gemm128.hpp"

namespace mpx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float flt4v __attribute__((ext_vector_type(4)));

constexpr int X2K = 32;                      // k panel: one bf16 MFMA depth

// Operand element -> thread map of one layout for panels of LINES lines.
template <bool Contig, int LINES>
struct X3Map {
  static constexpr int RUNS = LINES / 64;  // runs of 8 consecutive k per thread
  static __device__ __forceinline__ int l0(int tid) { return Contig ? tid >> 2 : tid % LINES; }
  static constexpr int lrun = Contig ? 64 : 0;
  static __device__ __forceinline__ int kk0(int tid) { return Contig ? 8 * (tid & 3) : (LINES / 8) * (tid / LINES); }
  static constexpr int krun = Contig ? 0 : 8;
};

// One operand of the cursor's pair: the buffer resource of its block and the thread's offsets.
struct X3View {
  __amdgpu_buffer_rsrc_t rsrc;
  int32_t boff0;  // byte offset of the thread's first element for k0 = 0
  int32_t brun;   // byte offset of run 1 relative to run 0
  int32_t bsk;    // bytes per unit of k
  int32_t lim;    // lines of the block from the tile's first line on
};

// One register set: a panel's elements per operand and, for a unit-stride-k operand, the k left per run.
template <int LINES>
struct X3Set {
  float a[LINES / 8], b[LINES / 8];
  int kla[LINES / 64], klb[LINES / 64];
};

template <bool Contig, int LINES>
__device__ __forceinline__ void x3_view(X3View& v, int tid, const float* __restrict__ block, int line0, int nline,
                                        int klen, int64_t s_line, int64_t s_k) {
  typedef X3Map<Contig, LINES> M;
  // elements spanned by the block (0 for an empty one: then every element reads as zero)
  const int64_t extent = nline > 0 && klen > 0 ? (int64_t)(nline - 1) * s_line + (int64_t)(klen - 1) * s_k + 1 : 0;
  v.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(block), (short)0, (int)(extent * 4), 0x00020000);
  v.boff0 = (int32_t)(4 * ((int64_t)(line0 + M::l0(tid)) * s_line + (int64_t)M::kk0(tid) * s_k));
  v.brun = (int32_t)(4 * (M::lrun * s_line + M::krun * s_k));
  v.bsk = (int32_t)(4 * s_k);
  v.lim = nline - line0;
}

// The thread's elements of the panel at k0; valid = false (past the tile's last panel) reads zeros.  kl[run] = k of
// the run inside the pair (unit-stride-k layout: elements j >= kl[run] are zeroed at the store).
template <bool Contig, int LINES>
__device__ __forceinline__ void x3_load(float (&r)[LINES / 8], int (&kl)[LINES / 64], const X3View& v, int tid,
                                        int k0, int klen, bool valid) {
  typedef X3Map<Contig, LINES> M;
  const int32_t base = v.boff0 + k0 * v.bsk;
#pragma unroll
  for (int run = 0; run < M::RUNS; ++run) {
    const bool lok = valid && M::l0(tid) + run * M::lrun < v.lim;
    const int kleft = klen - k0 - M::kk0(tid) - run * M::krun;
    kl[run] = kleft;
    const int32_t rb = base + run * v.brun;
    if constexpr (Contig) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int32_t off = (lok && 4 * h < kleft) ? rb + 16 * h : kOutOfBlock;
        const flt4v x = __builtin_bit_cast(flt4v, __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, off, 0, 0));
#pragma unroll
        for (int j = 0; j < 4; ++j) r[8 * run + 4 * h + j] = x[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int32_t off = (lok && j < kleft) ? rb + j * v.bsk : kOutOfBlock;
        r[8 * run + j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(v.rsrc, off, 0, 0));
      }
    }
  }
}

__device__ __forceinline__ int x3_chunk(int line, int chunk) { return (chunk ^ ((line >> 2) & 3)) * 8; }

// split the thread's elements and store them as two 16-byte vectors per run into the hi and lo panels
template <bool Contig, int LINES>
__device__ __forceinline__ void x3_store(__bf16 (*Hi)[X2K], __bf16 (*Lo)[X2K], const float (&r)[LINES / 8],
                                         const int (&kl)[LINES / 64], int tid) {
  typedef X3Map<Contig, LINES> M;
#pragma unroll
  for (int run = 0; run < M::RUNS; ++run) {
    bf16x8 h, l;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float x = Contig ? (j < kl[run] ? r[8 * run + j] : 0.f) : r[8 * run + j];
      const __bf16 hb = (__bf16)x;
      h[j] = hb;
      l[j] = (__bf16)(x - (float)hb);
    }
    const int line = M::l0(tid) + run * M::lrun, c = x3_chunk(line, (M::kk0(tid) + run * M::krun) >> 3);
    *reinterpret_cast<bf16x8*>(&Hi[line][c]) = h;
    *reinterpret_cast<bf16x8*>(&Lo[line][c]) = l;
  }
}

template <int LINES>
struct X3Lds {
  __bf16 Ah[2][LINES][X2K];
  __bf16 Al[2][LINES][X2K];
  __bf16 Bh[2][LINES][X2K];
  __bf16 Bl[2][LINES][X2K];
};

// The panel pipeline of one TILE x TILE tile into acc (N x N MFMA tiles per wave), for operand layouts CA, CB.
template <bool CA, bool CB, int TILE, int N = TILE / 32>
__device__ __forceinline__ void x3_tile(flt4 (&acc)[N][N], X3Lds<TILE>& S, const BigTile& tl, const GemmTask& tk,
                                        const GemmPair* __restrict__ pairs, const float* __restrict__ A,
                                        const float* __restrict__ B, int tid, int wr, int wc, int lr, int lk,
                                        const bool (&live_r)[N], const bool (&live_c)[N]) {
  // the load cursor (cp, ck0) walks the tile's (pair, k panel) sequence; cva, cvb view the cursor's pair
  int cp = tl.pair_lo, ck0 = 0;
  GemmPair cpr = pairs[cp];
  X3View cva, cvb;
  x3_view<CA, TILE>(cva, tid, A + cpr.a_off, tl.row0, tk.m, cpr.k, cpr.sa_m, cpr.sa_k);
  x3_view<CB, TILE>(cvb, tid, B + cpr.b_off, tl.col0, tk.n, cpr.k, cpr.sb_n, cpr.sb_k);
  bool cur = true;  // the cursor is on a panel of the tile (uniform)
  auto load = [&](X3Set<TILE>& s) {
    x3_load<CA, TILE>(s.a, s.kla, cva, tid, ck0, cpr.k, cur);
    x3_load<CB, TILE>(s.b, s.klb, cvb, tid, ck0, cpr.k, cur);
    if (!cur) return;
    ck0 += X2K;
    if (ck0 < cpr.k) return;
    ck0 = 0;
    if (++cp >= tl.pair_hi) {
      cur = false;
      return;
    }
    cpr = pairs[cp];
    x3_view<CA, TILE>(cva, tid, A + cpr.a_off, tl.row0, tk.m, cpr.k, cpr.sa_m, cpr.sa_k);
    x3_view<CB, TILE>(cvb, tid, B + cpr.b_off, tl.col0, tk.n, cpr.k, cpr.sb_n, cpr.sb_k);
  };
  auto store = [&](const X3Set<TILE>& s, int b) {
    x3_store<CA, TILE>(S.Ah[b], S.Al[b], s.a, s.kla, tid);
    x3_store<CB, TILE>(S.Bh[b], S.Bl[b], s.b, s.klb, tid);
  };
  auto mfma = [&](int b) {  // the B vectors of the wave's N column blocks stay in registers, A is read per row block
    bf16x8 bh[N], bl[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int rb_ = wc * (TILE / 2) + i * 16 + lr;
      bh[i] = *reinterpret_cast<const bf16x8*>(&S.Bh[b][rb_][x3_chunk(rb_, lk)]);
      bl[i] = *reinterpret_cast<const bf16x8*>(&S.Bl[b][rb_][x3_chunk(rb_, lk)]);
    }
#pragma unroll
    for (int mi = 0; mi < N; ++mi) {
      const int ra_ = wr * (TILE / 2) + mi * 16 + lr;
      const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&S.Ah[b][ra_][x3_chunk(ra_, lk)]);
      const bf16x8 al = *reinterpret_cast<const bf16x8*>(&S.Al[b][ra_][x3_chunk(ra_, lk)]);
#pragma unroll
      for (int ni = 0; ni < N; ++ni) {
        if (!(live_r[mi] && live_c[ni])) continue;  // wholly in the padding of an edge tile
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[ni], acc[mi][ni], 0, 0, 0);
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[ni], acc[mi][ni], 0, 0, 0);
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[ni], acc[mi][ni], 0, 0, 0);
      }
    }
  };
  X3Set<TILE> x;
  load(x);  // panel 0
  store(x, 0);
  __syncthreads();
  int buf = 0;
  while (true) {
    const bool more = cur;  // a next panel: loaded into x while the MFMAs run on the current one (uniform)
    load(x);                // zeros past the tile's last panel
    mfma(buf);
    if (!more) break;
    store(x, buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
}

template <int TILE, typename TC, bool Tab, bool CA, bool CB>
__device__ __forceinline__ void gemm_x3_body(const BigTile* __restrict__ tiles, int ntiles,
                                             const GemmTask* __restrict__ tasks, const GemmPair* __restrict__ pairs,
                                             Operands ops, double alpha, int accumulate, float* __restrict__ partial,
                                             int xchunk) {
  constexpr int N = TILE / 32, H = TILE / 2;  // MFMA tiles per wave and dimension, wave quarter edge
  __shared__ __attribute__((aligned(16))) X3Lds<TILE> S;
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
    bool live_r[N], live_c[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      live_r[i] = tl.row0 + wr * H + i * 16 < tk.m;
      live_c[i] = tl.col0 + wc * H + i * 16 < tk.n;
    }
    flt4 acc[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) acc[i][j] = flt4{0, 0, 0, 0};
    x3_tile<CA, CB, TILE>(acc, S, tl, tk, pairs, A, B, tid, wr, wc, lr, lk, live_r, live_c);

    if (tl.slot < 0) {
      TC* __restrict__ C = opnd<TC, Tab>(ops.c, tk.inst);
#pragma unroll
      for (int mi = 0; mi < N; ++mi)
#pragma unroll
        for (int ni = 0; ni < N; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = tl.row0 + wr * H + mi * 16 + Mfma<float>::row(lk, r);
            const int col = tl.col0 + wc * H + ni * 16 + lr;
            if (row < tk.m && col < tk.n) {
              const int64_t idx = tk.c_off + (int64_t)row * tk.n + col;
              const double v = alpha * (double)acc[mi][ni][r];
              C[idx] = accumulate ? (TC)((double)C[idx] + v) : (TC)v;
            }
          }
    } else {
      float* __restrict__ P = partial + (int64_t)tl.slot * (TILE * TILE);
#pragma unroll
      for (int mi = 0; mi < N; ++mi)
#pragma unroll
        for (int ni = 0; ni < N; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = wr * H + mi * 16 + Mfma<float>::row(lk, r);
            const int col = wc * H + ni * 16 + lr;
            P[row * TILE + col] = acc[mi][ni][r];
          }
    }
    __syncthreads();  // every wave is done with the LDS buffers before the next tile's prologue
  }
}

template <typename TC, bool Tab, bool CA, bool CB>
__global__ __launch_bounds__(256) void gemm_x3_kernel(const BigTile* __restrict__ tiles, int ntiles,
                                                      const GemmTask* __restrict__ tasks,
                                                      const GemmPair* __restrict__ pairs, Operands ops, double alpha,
                                                      int accumulate, float* __restrict__ partial, int xchunk) {
  gemm_x3_body<64, TC, Tab, CA, CB>(tiles, ntiles, tasks, pairs, ops, alpha, accumulate, partial, xchunk);
}

template <typename TC, bool Tab, bool CA, bool CB>
__global__ __launch_bounds__(256, 2) void gemm_x3_128_kernel(const BigTile* __restrict__ tiles, int ntiles,
                                                             const GemmTask* __restrict__ tasks,
                                                             const GemmPair* __restrict__ pairs, Operands ops,
                                                             double alpha, int accumulate,
                                                             float* __restrict__ partial, int xchunk) {
  gemm_x3_body<128, TC, Tab, CA, CB>(tiles, ntiles, tasks, pairs, ops, alpha, accumulate, partial, xchunk);
}

template <int TILE, typename TC, bool Tab, bool CA, bool CB>
void launch_x3_group(const BigTile* tiles, int ntiles, const GemmTask* tasks, const GemmPair* pairs, Operands ops,
                     double alpha, int accumulate, float* part, int max_groups, hipStream_t s) {
  if (ntiles <= 0) return;
  if constexpr (TILE == 128)
    hipLaunchKernelGGL((gemm_x3_128_kernel<TC, Tab, CA, CB>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s,
                       tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part, xcd_chunk());
  else
    hipLaunchKernelGGL((gemm_x3_kernel<TC, Tab, CA, CB>), dim3(groups(ntiles, max_groups)), dim3(256), 0, s, tiles,
                       ntiles, tasks, pairs, ops, alpha, accumulate, part, xcd_chunk());
}

template <int TILE, typename TC, bool Tab>
void launch_x3_groups(const BigTile* tiles, const int (&n)[4], const GemmTask* tasks, const GemmPair* pairs,
                      Operands ops, double alpha, int accumulate, float* part, int max_groups, hipStream_t s) {
  launch_x3_group<TILE, TC, Tab, true, true>(tiles, n[0], tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  tiles += n[0];
  launch_x3_group<TILE, TC, Tab, true, false>(tiles, n[1], tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  tiles += n[1];
  launch_x3_group<TILE, TC, Tab, false, true>(tiles, n[2], tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  tiles += n[2];
  launch_x3_group<TILE, TC, Tab, false, false>(tiles, n[3], tasks, pairs, ops, alpha, accumulate, part, max_groups,
                                               s);
}

template <int TILE, typename TC>
void launch_x3_tiles(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                     const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                     hipStream_t s, int max_groups) {
  if (ntiles[0] + ntiles[1] + ntiles[2] + ntiles[3] <= 0) return;
  float* part = static_cast<float*>(partial);
  if (ops.table)
    launch_x3_groups<TILE, TC, true>(tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  else
    launch_x3_groups<TILE, TC, false>(tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part, max_groups, s);
  if (nred <= 0) return;
  if constexpr (TILE == 128) {
    if (ops.table)
      hipLaunchKernelGGL((gemm_reduce128_kernel<float, TC, true>), dim3(groups(nred * kRedStrips2, max_groups)),
                         dim3(256), 0, s, red, nred, tasks, part, ops, alpha, accumulate);
    else
      hipLaunchKernelGGL((gemm_reduce128_kernel<float, TC, false>), dim3(groups(nred * kRedStrips2, max_groups)),
                         dim3(256), 0, s, red, nred, tasks, part, ops, alpha, accumulate);
  } else {
    if (ops.table)
      hipLaunchKernelGGL((gemm_reduce_kernel<float, TC, true>), dim3(groups(nred * kRedStrips, max_groups)),
                         dim3(256), 0, s, red, nred, tasks, part, ops, alpha, accumulate);
    else
      hipLaunchKernelGGL((gemm_reduce_kernel<float, TC, false>), dim3(groups(nred * kRedStrips, max_groups)),
                         dim3(256), 0, s, red, nred, tasks, part, ops, alpha, accumulate);
  }
}

template <typename TC>
void launch_gemm_x3(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                    const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                    hipStream_t s, int max_groups) {
  launch_x3_tiles<64, TC>(tiles, ntiles, red, nred, partial, tasks, pairs, ops, alpha, accumulate, s, max_groups);
}

template <typename TC>
void launch_gemm_x3_128(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                        const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                        hipStream_t s, int max_groups) {
  launch_x3_tiles<128, TC>(tiles, ntiles, red, nred, partial, tasks, pairs, ops, alpha, accumulate, s, max_groups);
}

}  // namespace mpx
