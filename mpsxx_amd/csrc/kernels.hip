// HIP kernels for gfx950 (CDNA4, wave64).  The 64x64 grouped-GEMM tiles are in gemm_tiles.hpp (their FP64
// instantiation in kernels_f64.hip).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "kernels.hpp"
#include "gemm_tiles.hpp"
#include "gemm128.hpp"

namespace mpx {

// Operand fragment of one (pair, 16-wide k panel) unit of a 32x32 tile, straight from global memory into MFMA source
// registers: lane l takes A[row + l%16][k + l/16] and B[k + l/16][col + l%16] for four 4-wide k steps.
template <typename T>
struct Frag {
  T a0[4], a1[4], b0[4], b1[4];
};

struct TileLanes {
  int r0, r1, c0, c1;
  bool vr0, vr1, vc0, vc1;
  int lk;
};

template <typename T>
__device__ __forceinline__ void load_frag(Frag<T>& f, const T* __restrict__ A, const T* __restrict__ B,
                                          const GemmPair& pr, int k0, const TileLanes& t) {
  const T* a0p = A + pr.a_off + (int64_t)t.r0 * pr.sa_m;
  const T* a1p = A + pr.a_off + (int64_t)t.r1 * pr.sa_m;
  const T* b0p = B + pr.b_off + (int64_t)t.c0 * pr.sb_n;
  const T* b1p = B + pr.b_off + (int64_t)t.c1 * pr.sb_n;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int gk = k0 + 4 * u + t.lk;
    const bool vk = gk < pr.k;
    const int64_t ao = (int64_t)gk * pr.sa_k, bo = (int64_t)gk * pr.sb_k;
    f.a0[u] = (vk && t.vr0) ? a0p[ao] : T(0);
    f.a1[u] = (vk && t.vr1) ? a1p[ao] : T(0);
    f.b0[u] = (vk && t.vc0) ? b0p[bo] : T(0);
    f.b1[u] = (vk && t.vc1) ? b1p[bo] : T(0);
  }
}

template <typename T>
__device__ __forceinline__ void mma_frag(typename Mfma<T>::acc_t (&acc)[2][2], const Frag<T>& f, int kleft) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (4 * u >= kleft) break;  // uniform across the wave
    acc[0][0] = Mfma<T>::op(f.a0[u], f.b0[u], acc[0][0]);
    acc[0][1] = Mfma<T>::op(f.a0[u], f.b1[u], acc[0][1]);
    acc[1][0] = Mfma<T>::op(f.a1[u], f.b0[u], acc[1][0]);
    acc[1][1] = Mfma<T>::op(f.a1[u], f.b1[u], acc[1][1]);
  }
}

// advance a (pair, 16-wide panel) position by `steps` panels across pair boundaries; false past the last pair
__device__ __forceinline__ bool advance_units(const GemmPair* __restrict__ pairs, int pend, int& p, int& panel,
                                              GemmPair& pr, int steps) {
  panel += steps;
  while (true) {
    const int np = (pr.k + 15) / 16;
    if (panel < np) return true;
    panel -= np;
    if (++p >= pend) return false;
    pr = pairs[p];
  }
}

template <typename T, typename TC>
__device__ __forceinline__ void store_tile(TC* __restrict__ C, const GemmTask& tk, const GemmTile& tl,
                                           const typename Mfma<T>::acc_t (&acc)[2][2], int lr, int lk, double alpha,
                                           int accumulate) {
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = tl.row0 + mi * 16 + Mfma<T>::row(lk, r);
        const int col = tl.col0 + ni * 16 + lr;
        if (row < tk.m && col < tk.n) {
          const int64_t idx = tk.c_off + (int64_t)row * tk.n + col;
          const double v = alpha * (double)acc[mi][ni][r];
          C[idx] = accumulate ? (TC)((double)C[idx] + v) : (TC)v;
        }
      }
}

// Small-block variant: one wave per 32x32 C tile (2x2 MFMA 16x16 blocks), operands read straight from global memory
// into the MFMA source registers; the next (pair, panel) unit is loaded while the MFMAs of the current one run.  No
// LDS and no barriers, so thousands of tiny blocks run at full occupancy.  Four independent waves per work-group.
template <typename T, typename TC, bool Tab>
__global__ __launch_bounds__(256) void gemm_small_kernel(const GemmTile* __restrict__ tiles, int ntiles,
                                                         const GemmTask* __restrict__ tasks,
                                                         const GemmPair* __restrict__ pairs, Operands ops,
                                                         double alpha, int accumulate) {
  typedef typename Mfma<T>::acc_t acc_t;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  for (int ti = blockIdx.x * 4 + wave; ti < ntiles; ti += gridDim.x * 4) {
    const GemmTile tl = tiles[ti];
    const GemmTask tk = tasks[tl.task];
    const T* __restrict__ A = opnd<const T, Tab>(ops.a, tk.inst);
    const T* __restrict__ B = opnd<const T, Tab>(ops.b, tk.inst);
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    TileLanes tlanes;
    tlanes.r0 = tl.row0 + lr;
    tlanes.r1 = tl.row0 + 16 + lr;
    tlanes.c0 = tl.col0 + lr;
    tlanes.c1 = tl.col0 + 16 + lr;
    tlanes.vr0 = tlanes.r0 < tk.m;
    tlanes.vr1 = tlanes.r1 < tk.m;
    tlanes.vc0 = tlanes.c0 < tk.n;
    tlanes.vc1 = tlanes.c1 < tk.n;
    tlanes.lk = lk;
    const int pend = tk.pair_begin + tk.pair_count;
    int p = tk.pair_begin, panel = 0;
    GemmPair pr = pairs[p];
    Frag<T> cur;
    load_frag(cur, A, B, pr, 0, tlanes);
    while (true) {
      int np = p, npanel = panel;
      GemmPair npr = pr;
      const bool more = advance_units(pairs, pend, np, npanel, npr, 1);  // uniform across the wave
      Frag<T> nxt;
      if (more) load_frag(nxt, A, B, npr, 16 * npanel, tlanes);
      mma_frag<T>(acc, cur, pr.k - 16 * panel);
      if (!more) break;
      cur = nxt;
      p = np;
      panel = npanel;
      pr = npr;
    }
    store_tile<T, TC>(opnd<TC, Tab>(ops.c, tk.inst), tk, tl, acc, lr, lk, alpha, accumulate);
  }
}

// advance a (pair, 16-wide panel) position by `steps` panels across the pair descriptors staged in LDS
__device__ __forceinline__ bool advance_staged(const GemmPair* sp, int n, int& p, int& panel, GemmPair& pr, int steps) {
  panel += steps;
  while (true) {
    const int np = (pr.k + 15) / 16;
    if (panel < np) return true;
    panel -= np;
    if (++p >= n) return false;
    pr = sp[p];
  }
}

__device__ __forceinline__ int wave_isum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kSplitPairs = 128;  // pair descriptors of one task staged in LDS per round
constexpr int kSplitWaves = 4;    // waves sharing the reduction of one 32x32 tile

// K-split variant for 32x32 tiles with a long reduction (many pairs and/or large k): the four waves of a work-group
// take every fourth (pair, 16-wide k panel) unit of the tile's pair range (each wave loading its next unit while the
// MFMAs of the current one run), then the partial tiles are summed in LDS in a fixed order (deterministic) and stored
// once: into C, or, for a tile whose reduction is cut over several work-groups (slot >= 0), unscaled into the 32x32
// row-major partial tile `slot`, which gemm_reduce32_kernel adds in slot order.  The pair descriptors are first
// staged in LDS by the whole work-group (rounds of kSplitPairs), so the unit loop waits only on operand loads, not on
// a descriptor load before each operand load.
template <typename T, typename TC, bool Tab, int NW>
__global__ __launch_bounds__(64 * NW) void gemm_split_kernel(const BigTile* __restrict__ tiles, int ntiles,
                                                         const GemmTask* __restrict__ tasks,
                                                         const GemmPair* __restrict__ pairs, Operands ops,
                                                         double alpha, int accumulate, T* __restrict__ partial) {
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T red[NW - 1][4][4][64];  // partial tiles of waves 1..NW-1: [wave-1][mi*2+ni][r][lane]
  __shared__ GemmPair sp[kSplitPairs];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const BigTile tl = tiles[w];
    const GemmTask tk = tasks[tl.task];
    const T* __restrict__ A = opnd<const T, Tab>(ops.a, tk.inst);
    const T* __restrict__ B = opnd<const T, Tab>(ops.b, tk.inst);
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    TileLanes tlanes;
    tlanes.r0 = tl.row0 + lr;
    tlanes.r1 = tl.row0 + 16 + lr;
    tlanes.c0 = tl.col0 + lr;
    tlanes.c1 = tl.col0 + 16 + lr;
    tlanes.vr0 = tlanes.r0 < tk.m;
    tlanes.vr1 = tlanes.r1 < tk.m;
    tlanes.vc0 = tlanes.c0 < tk.n;
    tlanes.vc1 = tlanes.c1 < tk.n;
    tlanes.lk = lk;
    const int pend = tl.pair_hi;
    int ubase = 0;  // units of the range before this round, modulo NW
    for (int c0 = tl.pair_lo; c0 < pend; c0 += kSplitPairs) {  // uniform across the work-group
      const int nc = min(kSplitPairs, pend - c0);
      __syncthreads();  // the previous round's descriptors are no longer read
      for (int i = threadIdx.x; i < nc; i += 64 * NW) sp[i] = pairs[c0 + i];
      __syncthreads();
      int u = 0;
      for (int i = lane; i < nc; i += 64) u += (sp[i].k + 15) / 16;
      const int nunits = wave_isum(u);
      // the wave takes the units whose index in the whole task is congruent to its number modulo NW
      int p = 0, panel = 0;
      GemmPair pr = sp[0];
      bool have = advance_staged(sp, nc, p, panel, pr, (wave - ubase + NW) % NW);
      Frag<T> cur;
      if (have) load_frag(cur, A, B, pr, 16 * panel, tlanes);
      while (have) {
        int np = p, npanel = panel;
        GemmPair npr = pr;
        const bool more = advance_staged(sp, nc, np, npanel, npr, NW);
        Frag<T> nxt;
        if (more) load_frag(nxt, A, B, npr, 16 * npanel, tlanes);
        mma_frag<T>(acc, cur, pr.k - 16 * panel);
        if (!more) break;
        cur = nxt;
        p = np;
        panel = npanel;
        pr = npr;
      }
      ubase = (ubase + nunits) % NW;
    }
    if (wave > 0) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) red[wave - 1][mi * 2 + ni][r][lane] = acc[mi][ni][r];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            T v = acc[mi][ni][r];
#pragma unroll
            for (int q = 0; q < NW - 1; ++q) v = v + red[q][mi * 2 + ni][r][lane];
            acc[mi][ni][r] = v;
          }
      if (tl.slot >= 0) {
        T* __restrict__ P = partial + (int64_t)tl.slot * (kSmallTile * kSmallTile);
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
          for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              P[(mi * 16 + Mfma<T>::row(lk, r)) * kSmallTile + ni * 16 + lr] = acc[mi][ni][r];
      } else {
        const GemmTile ct{tl.task, tl.row0, tl.col0, 0};
        store_tile<T, TC>(opnd<TC, Tab>(ops.c, tk.inst), tk, ct, acc, lr, lk, alpha, accumulate);
      }
    }
    __syncthreads();  // red is rewritten by the next tile
  }
}

// Sum of the partial 32x32 tiles of a cut reduction in slot order, scaled and stored (or accumulated) into C: four
// consecutive elements per thread read as one vector, the slot loop unrolled so that several loads are in flight.
template <typename T, typename TC, bool Tab>
__global__ __launch_bounds__(256) void gemm_reduce32_kernel(const RedTile* __restrict__ red, int nred,
                                                            const GemmTask* __restrict__ tasks,
                                                            const T* __restrict__ partial, Operands ops, double alpha,
                                                            int accumulate) {
  typedef typename Mfma<T>::acc_t vec_t;
  constexpr int kTile = kSmallTile * kSmallTile;
  for (int w = blockIdx.x; w < nred; w += gridDim.x) {
    const RedTile rt = red[w];
    const GemmTask tk = tasks[rt.task];
    const int lr = threadIdx.x >> 3, lc = (threadIdx.x & 7) * 4;
    const int row = rt.row0 + lr, col = rt.col0 + lc;
    if (row >= tk.m || col >= tk.n) continue;
    const vec_t* __restrict__ P =
        reinterpret_cast<const vec_t*>(partial + (int64_t)rt.slot0 * kTile + lr * kSmallTile + lc);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll 4
    for (int q = 0; q < rt.nslots; ++q) {
      const vec_t v = P[(int64_t)q * (kTile / 4)];
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

// Scaled accumulation C[r] += sum over pairs of coef * V[r * stride] for contractions with scalar (MPO) blocks.  The
// pair descriptors and their coefficients (S, or the fused S * S2 of two MPO steps) are staged in LDS by the whole
// work-group, 256 pairs per round, so the element loop issues only the independent V loads; pairs are summed in list
// order.
template <typename T, typename TC, bool Tab>
__global__ __launch_bounds__(256) void axpy_grouped_kernel(const Chunk* __restrict__ chunks, int nchunks,
                                                           const AxpyTask* __restrict__ tasks,
                                                           const AxpyPair* __restrict__ pairs, Operands ops,
                                                           double alpha, int accumulate) {
  __shared__ int64_t s_voff[256];
  __shared__ int32_t s_vstr[256];
  __shared__ T s_coef[256];
  for (int w = blockIdx.x; w < nchunks; w += gridDim.x) {
    const Chunk ch = chunks[w];
    const AxpyTask tk = tasks[ch.task];
    const T* __restrict__ V = opnd<const T, Tab>(ops.a, tk.inst);
    const T* __restrict__ S = opnd<const T, Tab>(ops.b, tk.inst);
    const T* __restrict__ S2 = opnd<const T, Tab>(ops.b2, tk.inst);
    TC* __restrict__ C = opnd<TC, Tab>(ops.c, tk.inst);
    const int pend = tk.pair_begin + tk.pair_count;
    constexpr int E = kAxpyChunk / 256;  // elements per thread; the E loads of one pair are independent
    const int end = min(tk.len, ch.start + kAxpyChunk);
    T acc[E];
#pragma unroll
    for (int j = 0; j < E; ++j) acc[j] = T(0);
    for (int c0 = tk.pair_begin; c0 < pend; c0 += 256) {  // uniform across the work-group
      const int nc = min(256, pend - c0);
      __syncthreads();  // the previous round's entries are no longer read
      if ((int)threadIdx.x < nc) {
        const AxpyPair pr = pairs[c0 + threadIdx.x];
        s_voff[threadIdx.x] = pr.v_off;
        s_vstr[threadIdx.x] = pr.v_stride;
        s_coef[threadIdx.x] = pr.s2_off >= 0 ? S[pr.s_off] * S2[pr.s2_off] : S[pr.s_off];
      }
      __syncthreads();
      for (int q = 0; q < nc; ++q) {
        const T coef = s_coef[q];
        const T* vp = V + s_voff[q];
        const int64_t st = s_vstr[q];
#pragma unroll
        for (int j = 0; j < E; ++j) {
          const int r = ch.start + threadIdx.x + 256 * j;
          if (r < end) acc[j] += coef * vp[(int64_t)r * st];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int r = ch.start + threadIdx.x + 256 * j;
      if (r < end) {
        const int64_t idx = tk.c_off + r;
        const double v = alpha * (double)acc[j];
        C[idx] = accumulate ? (TC)((double)C[idx] + v) : (TC)v;
      }
    }
  }
}

// Merged block operator from a scalar-block MPO, one work-group per (sub-block, 1024-element chunk): the column's MPO
// entries and the matching environment block offsets are staged in LDS (256 per round), then every thread sums its
// elements over the entries in order and stores them once at their place in the merged block.
__global__ __launch_bounds__(256) void merged_operator_kernel(const Chunk* __restrict__ chunks, int nchunks,
                                                              const MergedOpTask* __restrict__ tasks,
                                                              const int32_t* __restrict__ slice_blk, int nw,
                                                              const int64_t* __restrict__ env_off,
                                                              const int32_t* __restrict__ col_begin,
                                                              const int32_t* __restrict__ ent_w,
                                                              const int64_t* __restrict__ ent_off,
                                                              const double* __restrict__ env,
                                                              const double* __restrict__ mpo, double* __restrict__ out) {
  __shared__ int64_t s_off[256];
  __shared__ double s_coef[256];
  constexpr int E = kAxpyChunk / 256;
  for (int w = blockIdx.x; w < nchunks; w += gridDim.x) {
    const Chunk ch = chunks[w];
    const MergedOpTask tk = tasks[ch.task];
    const int end = min(tk.rows * tk.cols, ch.start + kAxpyChunk);
    double acc[E];
#pragma unroll
    for (int j = 0; j < E; ++j) acc[j] = 0.0;
    const int cb = col_begin[tk.col], ce = col_begin[tk.col + 1];
    for (int c0 = cb; c0 < ce; c0 += 256) {  // uniform across the work-group
      const int nc = min(256, ce - c0);
      __syncthreads();  // the previous round's entries are no longer read
      if ((int)threadIdx.x < nc) {
        const int blk = slice_blk[(int64_t)tk.slice * nw + ent_w[c0 + threadIdx.x]];
        s_off[threadIdx.x] = blk >= 0 ? env_off[blk] : -1;
        s_coef[threadIdx.x] = mpo[ent_off[c0 + threadIdx.x]];
      }
      __syncthreads();
      for (int q = 0; q < nc; ++q) {
        const int64_t lo = s_off[q];
        if (lo < 0) continue;  // uniform across the work-group
        const double coef = s_coef[q];
        const double* vp = env + lo;
#pragma unroll
        for (int j = 0; j < E; ++j) {
          const int r = ch.start + threadIdx.x + 256 * j;
          if (r < end) acc[j] += coef * vp[r];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int r = ch.start + threadIdx.x + 256 * j;
      if (r < end) out[tk.y_off + (int64_t)(r / tk.cols) * tk.ld + r % tk.cols] = acc[j];
    }
  }
}

template <typename TX, typename TY>
__global__ __launch_bounds__(256) void copy_kernel(const Chunk* __restrict__ chunks,
                                                   const CopySeg* __restrict__ segs, const TX* __restrict__ x,
                                                   TY* __restrict__ y, double alpha, double beta) {
  const Chunk ch = chunks[blockIdx.x];
  const CopySeg sg = segs[ch.task];
  for (int r = ch.start + threadIdx.x; r < sg.len && r < ch.start + kAxpyChunk; r += 256) {
    const int64_t d = sg.dst_off + r;
    const double v = alpha * (double)x[sg.src_off + (int64_t)r * sg.stride];
    y[d] = (beta == 0.0) ? (TY)v : (TY)(beta * (double)y[d] + v);
  }
}

__global__ __launch_bounds__(256) void convert_segments_kernel(const Chunk* __restrict__ chunks, int nchunks,
                                                               const ConvSeg* __restrict__ segs) {
  for (int w = blockIdx.x; w < nchunks; w += gridDim.x) {
    const Chunk ch = chunks[w];
    const ConvSeg sg = segs[ch.task];
    const int64_t end = min(sg.n, (int64_t)ch.start + kAxpyChunk);
    for (int64_t r = ch.start + threadIdx.x; r < end; r += 256) sg.dst[r] = (float)sg.src[r];
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ double block_sum(double v) {
  __shared__ double red[4];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) t = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(256) void dot_partial_kernel(const Chunk* __restrict__ chunks,
                                                          const CopySeg* __restrict__ segs,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ y, double* partial) {
  const Chunk ch = chunks[blockIdx.x];
  const CopySeg sg = segs[ch.task];
  double acc = 0.0;
  for (int r = ch.start + threadIdx.x; r < sg.len && r < ch.start + kAxpyChunk; r += 256)
    acc += x[sg.dst_off + r] * y[sg.src_off + r];
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void reduce_kernel(const double* __restrict__ partial, int n, double* out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[0] = acc;
}

__global__ void scale_kernel(double* x, int64_t n, double alpha) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= alpha;
}

__global__ void fill_kernel(double* x, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

__global__ void axpby_kernel(double a, const double* __restrict__ x, double b, double* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

template <typename TX, typename TY>
__global__ void convert_kernel(const TX* x, TY* y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (TY)x[i];
}

__global__ void combine_kernel(const double* __restrict__ base, int64_t ld, int m, const double* __restrict__ c,
                               double* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += c[j] * base[j * ld + i];
    y[i] = acc;
  }
}

__global__ void combine_sub_kernel(const double* __restrict__ base, int64_t ld, int m, const double* __restrict__ c,
                                   double* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += c[j] * base[j * ld + i];
    y[i] -= acc;
  }
}

__global__ void scale_rnorm_kernel(double* x, int64_t n, const double* __restrict__ nrm2) {
  const double v = nrm2[0], f = v > 0.0 ? 1.0 / sqrt(v) : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= f;
}

constexpr int kRotMax = 10;  // the Davidson subspace bound
__global__ void rotate2_kernel(const double* __restrict__ T, const double* __restrict__ S, int64_t ld, int m,
                               const double* __restrict__ c, int cld, double* __restrict__ TO,
                               double* __restrict__ SO, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double t[kRotMax], sv[kRotMax];
#pragma unroll
    for (int k = 0; k < kRotMax; ++k) {
      t[k] = k < m ? T[k * ld + i] : 0.0;
      sv[k] = k < m ? S[k * ld + i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kRotMax; ++j) {
      if (j < m) {
        double at = 0.0, as = 0.0;
#pragma unroll
        for (int k = 0; k < kRotMax; ++k) {
          const double cf = k < m ? c[j * cld + k] : 0.0;
          at += cf * t[k];
          as += cf * sv[k];
        }
        TO[j * ld + i] = at;
        SO[j * ld + i] = as;
      }
    }
  }
}

__global__ __launch_bounds__(256) void residual_partial_kernel(const double* __restrict__ s0,
                                                               const double* __restrict__ t0, double e,
                                                               double* __restrict__ r, int64_t n, double* partial,
                                                               int npart) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)npart * 256) {
    const double v = s0[i] - e * t0[i];
    r[i] = v;
    acc += v * v;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void precondition_kernel(double* r, const double* __restrict__ d, const unsigned char* __restrict__ pres,
                                    double eval, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (pres[i]) {
      double denm = eval - d[i];
      if (fabs(denm) < 1.0e-8) denm = 1.0e-8;
      r[i] /= denm;
    } else {
      r[i] *= 1.0 / eval;
    }
  }
}

// Gram block: grid (npart, mx*my); each work-group reduces a strided slice into partial[(pair)*npart + part]
__global__ __launch_bounds__(256) void gram_partial_kernel(const double* __restrict__ X, int64_t ldx,
                                                           const double* __restrict__ Y, int64_t ldy, int my,
                                                           int64_t n, double* partial, int npart) {
  const int pair = blockIdx.y;
  const int i = pair / my, j = pair % my;
  const double* x = X + i * ldx;
  const double* y = Y + j * ldy;
  double acc = 0.0;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)npart * 256) acc += x[e] * y[e];
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[(int64_t)pair * npart + blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(const double* __restrict__ partial, int npart, double* G) {
  const int pair = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < npart; i += 256) acc += partial[(int64_t)pair * npart + i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) G[pair] = acc;
}

// ---- batched Davidson: segmented vector operations (one work-group per chunk of one instance's segment) ----------
__global__ __launch_bounds__(256) void seg_gram_partial_kernel(const SegChunk* __restrict__ ch,
                                                               const double* __restrict__ X, int64_t ldx,
                                                               const double* __restrict__ Y, int64_t ldy, int my,
                                                               double* partial, int nch) {
  const SegChunk c = ch[blockIdx.x];
  const int pair = blockIdx.y;
  const double* x = X + (pair / my) * ldx + c.begin;
  const double* y = Y + (pair % my) * ldy + c.begin;
  double acc = 0.0;
  for (int e = threadIdx.x; e < c.len; e += 256) acc += x[e] * y[e];
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[(int64_t)pair * nch + blockIdx.x] = acc;
}

// grid (ninst, npairs): instance i sums the partials of its chunks in chunk order
__global__ __launch_bounds__(256) void seg_reduce_kernel(const double* __restrict__ partial, int nch,
                                                         const int32_t* __restrict__ chunk0, double* G, int gstride) {
  const int inst = blockIdx.x, pair = blockIdx.y;
  const int c0 = chunk0[inst], c1 = chunk0[inst + 1];
  double acc = 0.0;
  for (int c = c0 + threadIdx.x; c < c1; c += 256) acc += partial[(int64_t)pair * nch + c];
  acc = block_sum(acc);
  if (threadIdx.x == 0) G[(int64_t)inst * gstride + pair] = acc;
}

__global__ __launch_bounds__(256) void seg_residual_partial_kernel(const SegChunk* __restrict__ ch,
                                                                   const double* __restrict__ s0,
                                                                   const double* __restrict__ t0,
                                                                   const double* __restrict__ e,
                                                                   double* __restrict__ r, double* partial) {
  const SegChunk c = ch[blockIdx.x];
  const double ev = e[c.inst];
  double acc = 0.0;
  for (int k = threadIdx.x; k < c.len; k += 256) {
    const int64_t i = c.begin + k;
    const double v = s0[i] - ev * t0[i];
    r[i] = v;
    acc += v * v;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void seg_rotate2_kernel(const SegChunk* __restrict__ ch,
                                                          const double* __restrict__ T, const double* __restrict__ S,
                                                          int64_t ld, int m, const double* __restrict__ cf, int cld,
                                                          int cstride, double* __restrict__ TO,
                                                          double* __restrict__ SO) {
  const SegChunk sc = ch[blockIdx.x];
  const double* __restrict__ c = cf + (int64_t)sc.inst * cstride;
  for (int k0 = threadIdx.x; k0 < sc.len; k0 += 256) {
    const int64_t i = sc.begin + k0;
    double t[kRotMax], sv[kRotMax];
#pragma unroll
    for (int k = 0; k < kRotMax; ++k) {
      t[k] = k < m ? T[k * ld + i] : 0.0;
      sv[k] = k < m ? S[k * ld + i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kRotMax; ++j) {
      if (j < m) {
        double at = 0.0, as = 0.0;
#pragma unroll
        for (int k = 0; k < kRotMax; ++k) {
          const double w = k < m ? c[j * cld + k] : 0.0;
          at += w * t[k];
          as += w * sv[k];
        }
        TO[j * ld + i] = at;
        SO[j * ld + i] = as;
      }
    }
  }
}

__global__ __launch_bounds__(256) void seg_combine_sub_kernel(const SegChunk* __restrict__ ch,
                                                              const double* __restrict__ base, int64_t ld, int m,
                                                              const double* __restrict__ cf, int cstride,
                                                              double* __restrict__ y) {
  const SegChunk sc = ch[blockIdx.x];
  const double* __restrict__ c = cf + (int64_t)sc.inst * cstride;
  for (int k0 = threadIdx.x; k0 < sc.len; k0 += 256) {
    const int64_t i = sc.begin + k0;
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += c[j] * base[j * ld + i];
    y[i] -= acc;
  }
}

__global__ __launch_bounds__(256) void seg_scale_rnorm_kernel(const SegChunk* __restrict__ ch, double* x,
                                                              const double* __restrict__ nrm2, int stride) {
  const SegChunk sc = ch[blockIdx.x];
  const double v = nrm2[(int64_t)sc.inst * stride], f = v > 0.0 ? 1.0 / sqrt(v) : 0.0;
  for (int k = threadIdx.x; k < sc.len; k += 256) x[sc.begin + k] *= f;
}

__global__ __launch_bounds__(256) void seg_precondition_kernel(const SegChunk* __restrict__ ch, double* r,
                                                               const double* __restrict__ d,
                                                               const unsigned char* __restrict__ pres,
                                                               const double* __restrict__ eval) {
  const SegChunk sc = ch[blockIdx.x];
  const double ev = eval[sc.inst];
  for (int k = threadIdx.x; k < sc.len; k += 256) {
    const int64_t i = sc.begin + k;
    if (pres[i]) {
      double denm = ev - d[i];
      if (fabs(denm) < 1.0e-8) denm = 1.0e-8;
      r[i] /= denm;
    } else {
      r[i] *= 1.0 / ev;
    }
  }
}

__global__ __launch_bounds__(256) void seg_copy_kernel(const SegChunk* __restrict__ ch,
                                                       const double* __restrict__ src, double* __restrict__ dst,
                                                       const int32_t* __restrict__ mask) {
  const SegChunk sc = ch[blockIdx.x];
  if (mask && mask[sc.inst] == 0) return;
  for (int k = threadIdx.x; k < sc.len; k += 256) dst[sc.begin + k] = src[sc.begin + k];
}

static inline int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}


int xcd_chunk() {
  static const int c = [] {
    const char* e = std::getenv("MPSXX_XCD_CHUNK");
    return e ? std::atoi(e) : 16;
  }();
  return c;
}


template <typename T, typename TC>
void launch_axpy(const Chunk* chunks, int nchunks, const AxpyTask* tasks, const AxpyPair* pairs, Operands ops,
                 double alpha, int accumulate, hipStream_t s, int max_groups) {
  if (nchunks <= 0) return;
  if (ops.table)
    hipLaunchKernelGGL((axpy_grouped_kernel<T, TC, true>), dim3(groups(nchunks, max_groups)), dim3(256), 0, s, chunks,
                       nchunks, tasks, pairs, ops, alpha, accumulate);
  else
    hipLaunchKernelGGL((axpy_grouped_kernel<T, TC, false>), dim3(groups(nchunks, max_groups)), dim3(256), 0, s, chunks,
                       nchunks, tasks, pairs, ops, alpha, accumulate);
}

template <typename TX, typename TY>
void launch_copy(const Chunk* chunks, int nchunks, const CopySeg* segs, const TX* x, TY* y, double alpha,
                 double beta, hipStream_t s) {
  if (nchunks <= 0) return;
  hipLaunchKernelGGL((copy_kernel<TX, TY>), dim3(nchunks), dim3(256), 0, s, chunks, segs, x, y, alpha, beta);
}

void launch_merged_operator(const Chunk* chunks, int nchunks, const MergedOpTask* tasks, const int32_t* slice_blk,
                            int nw, const int64_t* env_off, const int32_t* col_begin, const int32_t* ent_w,
                            const int64_t* ent_off, const double* env, const double* mpo, double* out, hipStream_t s) {
  if (nchunks <= 0) return;
  hipLaunchKernelGGL(merged_operator_kernel, dim3(groups(nchunks, 1 << 20)), dim3(256), 0, s, chunks, nchunks, tasks,
                     slice_blk, nw, env_off, col_begin, ent_w, ent_off, env, mpo, out);
}

void launch_convert_segments(const Chunk* chunks, int nchunks, const ConvSeg* segs, hipStream_t s, int max_groups) {
  if (nchunks <= 0) return;
  hipLaunchKernelGGL(convert_segments_kernel, dim3(groups(nchunks, max_groups)), dim3(256), 0, s, chunks, nchunks,
                     segs);
}

void launch_dot(const Chunk* chunks, int nchunks, const CopySeg* segs, const double* x, const double* y,
                double* partial, double* out, hipStream_t s) {
  if (nchunks > 0)
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nchunks), dim3(256), 0, s, chunks, segs, x, y, partial);
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, s, partial, nchunks, out);
}

void launch_scale(double* x, int64_t n, double alpha, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, n, alpha);
}

void launch_fill(double* x, int64_t n, double v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, n, v);
}

__global__ void svd_scale_kernel(double* X, int batch, int rows, int cols, int64_t stride, const double* __restrict__ S,
                                 int kmax, int by_col) {
  const int64_t per = (int64_t)rows * cols, n = per * batch;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / per, rem = e % per;
    const int i = (int)(rem / cols), c = (int)(rem % cols);
    X[b * stride + rem] *= S[b * kmax + (by_col ? c : i)];
  }
}

void launch_svd_scale(double* X, int batch, int rows, int cols, int64_t stride, const double* S, int kmax, int by_col,
                      hipStream_t s) {
  const int64_t n = (int64_t)rows * cols * batch;
  if (n > 0) hipLaunchKernelGGL(svd_scale_kernel, dim3(grid_for(n)), dim3(256), 0, s, X, batch, rows, cols, stride, S,
                                kmax, by_col);
}

// ---- one-sided Jacobi SVD (Hestenes), blocked over column groups ----
//
// A round kernel runs one work-group per task: the task's nc <= COLS columns Y of X are loaded to LDS, their Gram
// matrix G = Y^T Y is formed with FP64 MFMAs, and one cyclic Jacobi sweep runs on G in the round-robin (circle)
// order, COLS/2 disjoint pairs per step.  For a pair (p, q) with a = G_pp, b = G_qq, g = G_pq the rotation is applied
// when |g| > tol sqrt(a b), tol = eps * max(len, 32): zeta = (b - a) / 2g, t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
// c = 1 / sqrt(1 + t^2), s = c t (Rutishauser's formulas), i.e. y_p <- c y_p - s y_q, y_q <- s y_p + c y_q, which
// zeroes the pair's inner product.  The step's rotations update G two-sidedly (G <- J^T G J, one 2x2 block per
// thread) and are gathered in Q (Q <- Q J); at the end Y <- Y Q and V <- V Q, again with MFMAs.  G is formed afresh
// from the rotated columns by every task, so the inner products of the last sweep, which applies no rotation, are
// exact dot products: a block has converged when one full sweep over all its column pairs applies no rotation.

__global__ void svd_init_kernel(const SvdBlk* __restrict__ blks, double* V) {
  const SvdBlk b = blks[blockIdx.x];
  const int64_t n = (int64_t)b.kk * b.kk;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) V[b.v_off + e] = (e / b.kk == e % b.kk) ? 1.0 : 0.0;
}

// Leading dimension (doubles) of a task's LDS columns of length len; the host sizes LDS with the same formula.
__host__ __device__ __forceinline__ int svd_lds_ld(int len) { return ((len + 15) & ~15) + 4; }

// Columns gcol(0..nc-1) of a column-major matrix (column length len) into LDS columns of stride ld; rows len..ld-1
// and columns nc..cols-1 are zeroed.  Thread tid moves rows tid, tid + nt, ... of every column (consecutive threads,
// consecutive rows of one column), up to 32 independent loads in flight before the LDS writes.
template <typename F>
__device__ __forceinline__ void svd_gather_columns(const double* __restrict__ src, int len, int nc, int cols, int ld,
                                                   F gcol, double* dst, int tid, int nt) {
  for (int i = tid; i < len; i += nt) {
    for (int c0 = 0; c0 < nc; c0 += 32) {
      double tmp[32];
#pragma unroll
      for (int j = 0; j < 32; ++j)
        if (c0 + j < nc) tmp[j] = src[(int64_t)gcol(c0 + j) * len + i];
#pragma unroll
      for (int j = 0; j < 32; ++j)
        if (c0 + j < nc) dst[(c0 + j) * ld + i] = tmp[j];
    }
  }
  const int pad = ld - len;
  for (int e = tid; e < nc * pad; e += nt) dst[(e / pad) * ld + len + e % pad] = 0.0;
  for (int e = tid; e < (cols - nc) * ld; e += nt) dst[nc * ld + e] = 0.0;
}

// out(:, gcol(c)) = Y Q(:, c) for c < nc, rows < rows_out: Y (ld-strided LDS columns, COLS of them), Q (COLS x COLS
// in Qs[k][c]).  Computed as the transposed product Q^T Y^T in 16x16 MFMA tiles, so that consecutive lanes hold
// consecutive rows of one output column (coalesced stores).
template <int COLS, typename F>
__device__ __forceinline__ void svd_apply_q(const double* ys, int ld, int rows_out, const double (*Qs)[33], int nc,
                                            F gcol, double* out, int64_t col_stride, int w, int l) {
  constexpr int NCT = (COLS + 15) / 16;
  const int nrt = (rows_out + 15) / 16;
  for (int tile = w; tile < nrt * NCT; tile += 4) {
    const int row0 = (tile / NCT) * 16, col0 = (tile % NCT) * 16;
    double a[COLS / 4], bb[COLS / 4];
#pragma unroll
    for (int u = 0; u < COLS / 4; ++u) {
      const int k = 4 * u + l / 16, ci = col0 + l % 16;
      a[u] = ci < COLS ? Qs[k][ci] : 0.0;   // A = Q^T: A[i][k] = Q[k][col0 + i]
      bb[u] = ys[k * ld + row0 + l % 16];    // B = Y^T: B[k][j] = Y[row0 + j][k]
    }
    dbl4 acc = {0.0, 0.0, 0.0, 0.0}, acc2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < COLS / 4; u += 2) {
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bb[u], acc, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u + 1], bb[u + 1], acc2, 0, 0, 0);
    }
    acc += acc2;
    const int r = row0 + l % 16;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = col0 + l / 16 + 4 * j;  // D[l/16 + 4j][l%16] = (Y Q)[row0 + l%16][col0 + l/16 + 4j]
      if (c < nc && r < rows_out) out[(int64_t)gcol(c) * col_stride + r] = acc[j];
    }
  }
}

// 1/sqrt(x) and 1/x (x > 0, normal): the hardware estimates refined by one Newton step, y <- y (3 - x y^2) / 2 and
// y <- y (2 - x y), which doubles their ~2^-29 relative accuracy to the double-precision level.
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return y * fma(-0.5 * x * y, y, 1.5);
}
__device__ __forceinline__ double rcp_nr(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return fma(y, fma(-x, y, 1.0), y);
}

// Rutishauser's rotation of a pair with a = |y_p|^2, b = |y_q|^2, g = y_p.y_q (identity unless real and
// |g| > tol sqrt(a b)): t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), zeta = (b - a) / 2g, written as
// 2g sign(b - a) / (|b - a| + sqrt((b - a)^2 + 4g^2)); c = 1 / sqrt(1 + t^2), s = c t.  Branch-free: the formulas
// run on every pair (on a harmless argument where no rotation applies) and the identity is selected afterwards, so
// the two rotations a thread needs per step are one interleaved dependency chain rather than two.
__device__ __forceinline__ void svd_rotation(double a, double b, double g, bool real, double tol2, double& cs,
                                             double& sn) {
  const double d2 = b - a, q2 = d2 * d2 + 4.0 * g * g;
  const bool rot = real && g != 0.0 && g * g > tol2 * (a * b) && q2 > 0.0;
  const double q = rot ? q2 : 1.0;
  const double tt = copysign(1.0, d2) * (2.0 * g) * rcp_nr(fabs(d2) + q * rsqrt_nr(q));
  const double c = rsqrt_nr(1.0 + tt * tt);
  cs = rot ? c : 1.0;
  sn = rot ? c * tt : 0.0;
}

// Pair `t` of step `st` of a task: cross-pairs task (two groups, n0 >= n1): the n1 cross pairs (i, n0 + j) with
// i = (j - st) mod n0, then the n0 - n1 columns of the first group without a partner this step paired among
// themselves (and, if odd in number, with the zero slot n0 + n1) with no rotation; otherwise the circle method over
// all n0 (+ n1) columns rounded up to even.  Every slot is in exactly one pair of a step.
__device__ __forceinline__ void svd_pair(int n0, int n1, bool cross, int st, int t, int& p, int& q, bool& real) {
  if (cross) {
    if (t < n1) {
      p = (t - st + n0) % n0;
      q = n0 + t;
      real = true;
    } else {
      const int u = 2 * (t - n1), idle = n0 - n1;
      p = (n1 + u - st + n0) % n0;
      q = u + 1 < idle ? (n1 + u + 1 - st + n0) % n0 : n0 + n1;
      real = false;
    }
  } else {
    const int nc = n0 + n1, P = nc + (nc & 1);
    p = t == 0 ? P - 1 : (st + t) % (P - 1);
    q = t == 0 ? st : (st - t + P - 1) % (P - 1);
    real = true;
  }
}

template <int COLS>
__global__ __launch_bounds__(256) void svd_round_kernel(const SvdTask* __restrict__ tasks,
                                                        const SvdBlk* __restrict__ blks, double* X, double* V,
                                                        const int32_t* __restrict__ active, int32_t* rot) {
  extern __shared__ double ys[];  // COLS columns of stride ld
  __shared__ double Gs[32][33], Gs2[32][33], Qs[32][33];
  __shared__ int scnt, sneed;
  const SvdTask t = tasks[blockIdx.x];
  if (!active[t.blk]) return;
  const SvdBlk b = blks[t.blk];
  const int nc = t.n0 + t.n1, len = b.len, kk = b.kk, ld = svd_lds_ld(len);
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  auto gcol = [&](int c) { return c < t.n0 ? t.c0 + c : t.c1 + (c - t.n0); };
  // V's columns are only needed at the end: their loads are issued now, into registers, when they fit
  // (thread tid holds row tid of each of the nc <= 32 columns: vectors of at most 256)
  constexpr int kVpf = 32;
  const bool vpf = kk <= 256;
  double vreg[kVpf];
  if (vpf && tid < kk) {
#pragma unroll
    for (int c = 0; c < kVpf; ++c)
      if (c < nc) vreg[c] = V[b.v_off + (int64_t)gcol(c) * kk + tid];
  }
  svd_gather_columns(X + b.x_off, len, nc, COLS, ld, gcol, ys, tid, 256);
  for (int e = tid; e < COLS * COLS; e += 256) Qs[e / COLS][e % COLS] = (e / COLS == e % COLS) ? 1.0 : 0.0;
  if (tid == 0) {
    scnt = 0;
    sneed = 0;
  }
  __syncthreads();
  // G = Y^T Y (zero outside the task's nc columns)
  {
    constexpr int NT = (COLS + 15) / 16;
    if (w < NT * NT) {
      const int c0 = (w / NT) * 16, c1 = (w % NT) * 16, len16 = (len + 15) & ~15;
      dbl4 acc = {0.0, 0.0, 0.0, 0.0}, acc2 = {0.0, 0.0, 0.0, 0.0};
      const int ca = c0 + l % 16, cb = c1 + l % 16;
      const double* ya = ys + (ca < COLS ? ca : 0) * ld + l / 16;
      const double* yb = ys + (cb < COLS ? cb : 0) * ld + l / 16;
      const double ma = ca < COLS ? 1.0 : 0.0, mb = cb < COLS ? 1.0 : 0.0;
      for (int k0 = 0; k0 < len16; k0 += 16) {  // rows len .. len16 - 1 are zero in LDS
        double a[4], bb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[u] = ma * ya[k0 + 4 * u];
          bb[u] = mb * yb[k0 + 4 * u];
        }
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], bb[0], acc, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], bb[1], acc2, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[2], bb[2], acc, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[3], bb[3], acc2, 0, 0, 0);
      }
      acc += acc2;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = c0 + l / 16 + 4 * j, c = c1 + l % 16;
        if (r < COLS && c < COLS) Gs[r][c] = (r < nc && c < nc) ? acc[j] : 0.0;
      }
    }
  }
  __syncthreads();
  const double tol = 2.220446049250313e-16 * (double)max(len, 32), tol2 = tol * tol;
  const bool cross = t.cross != 0;
  const int S = nc + (nc & 1), NP = S / 2, nsteps = cross ? t.n0 : S - 1;
  // thread (i, j) of a step owns the 2x2 block of G at the rows of pair i and the columns of pair j and two rows of Q's
  // columns of pair j.  It computes both pairs' rotations itself from G as it stood before the step (read from one
  // buffer, its block written to the other, so one barrier per step).  Pair indices are kept in registers and
  // advanced in place: the circle method moves every slot but the fixed one (P - 1) by one place per step.
  const int bi = tid / max(NP, 1), bj = tid - bi * max(NP, 1);
  const bool act = tid < NP * NP;
  int pi = 0, qi = 0, pj = 0, qj = 0;
  bool ri = false, rj = false;
  if (act) {
    svd_pair(t.n0, t.n1, cross, 0, bi, pi, qi, ri);
    svd_pair(t.n0, t.n1, cross, 0, bj, pj, qj, rj);
  }
  // a task none of whose pairs would rotate on the G just formed leaves G unchanged (identity rotations), so no pair
  // rotates later in its sweep either: it skips the sweep, exactly as if it had run it
  int need = 0;
  for (int e = tid; e < COLS * COLS; e += 256) {
    const int p = e / COLS, q = e % COLS;
    if (p < q && q < nc && (!cross || (p < t.n0 && q >= t.n0))) {
      const double a = Gs[p][p], bq = Gs[q][q], g = Gs[p][q], d2 = bq - a;
      if (g != 0.0 && g * g > tol2 * (a * bq) && d2 * d2 + 4.0 * g * g > 0.0) need = 1;
    }
  }
  if (need) sneed = 1;
  __syncthreads();
  if (!sneed) return;  // uniform
  double (*Gc)[33] = Gs;
  double (*Gn)[33] = Gs2;
  int myrot = 0;
  for (int st = 0; st < nsteps; ++st) {
    if (act) {
      // every LDS operand of the step is read up front: the G entries of both pairs and of the 2x2 block, and this
      // thread's two rows of Q's columns p_j, q_j (no other thread writes them this step)
      const double ai = Gc[pi][pi], bi2 = Gc[qi][qi], gi = Gc[pi][qi];
      const double aj = Gc[pj][pj], bj2 = Gc[qj][qj], gj = Gc[pj][qj];
      const double b00 = Gc[pi][pj], b01 = Gc[pi][qj], b10 = Gc[qi][pj], b11 = Gc[qi][qj];
      const double qu0 = Qs[2 * bi][pj], qv0 = Qs[2 * bi][qj], qu1 = Qs[2 * bi + 1][pj], qv1 = Qs[2 * bi + 1][qj];
      double ci, si, cj, sj;
      svd_rotation(ai, bi2, gi, ri, tol2, ci, si);
      svd_rotation(aj, bj2, gj, rj, tol2, cj, sj);
      if (bi == bj && si != 0.0) ++myrot;
      // G block (rows p_i, q_i; cols p_j, q_j) <- J_i^T B J_j, J = [[c, s], [-s, c]]
      const double t00 = cj * b00 - sj * b01, t01 = sj * b00 + cj * b01;
      const double t10 = cj * b10 - sj * b11, t11 = sj * b10 + cj * b11;
      Gn[pi][pj] = ci * t00 - si * t10;
      Gn[pi][qj] = ci * t01 - si * t11;
      Gn[qi][pj] = si * t00 + ci * t10;
      Gn[qi][qj] = si * t01 + ci * t11;
      if (sj != 0.0) {  // Q rows 2i, 2i+1, cols p_j, q_j <- B J_j (each Q element has one owner per step)
        Qs[2 * bi][pj] = cj * qu0 - sj * qv0;
        Qs[2 * bi][qj] = sj * qu0 + cj * qv0;
        Qs[2 * bi + 1][pj] = cj * qu1 - sj * qv1;
        Qs[2 * bi + 1][qj] = sj * qu1 + cj * qv1;
      }
      if (cross) {
        svd_pair(t.n0, t.n1, cross, st + 1, bi, pi, qi, ri);
        svd_pair(t.n0, t.n1, cross, st + 1, bj, pj, qj, rj);
      } else {
        const int P1 = S - 1;
        if (bi != 0) pi = pi + 1 == P1 ? 0 : pi + 1;
        qi = qi + 1 == P1 ? 0 : qi + 1;
        if (bj != 0) pj = pj + 1 == P1 ? 0 : pj + 1;
        qj = qj + 1 == P1 ? 0 : qj + 1;
      }
    }
    __syncthreads();
    double (*sw)[33] = Gc;
    Gc = Gn;
    Gn = sw;
  }
  if (myrot) atomicAdd(&scnt, myrot);
  __syncthreads();
  const int total = scnt;
  if (total == 0) return;  // uniform: nothing moved
  svd_apply_q<COLS>(ys, ld, len, Qs, nc, gcol, X + b.x_off, len, w, l);
  __syncthreads();
  if (vpf) {
    const int pad = ld - kk;
    if (tid < kk) {
#pragma unroll
      for (int c = 0; c < kVpf; ++c)
        if (c < nc) ys[c * ld + tid] = vreg[c];
    }
    for (int e = tid; e < nc * pad; e += 256) ys[(e / pad) * ld + kk + e % pad] = 0.0;
    for (int e = tid; e < (COLS - nc) * ld; e += 256) ys[nc * ld + e] = 0.0;
  } else {
    svd_gather_columns(V + b.v_off, kk, nc, COLS, ld, gcol, ys, tid, 256);
  }
  __syncthreads();
  svd_apply_q<COLS>(ys, ld, kk, Qs, nc, gcol, V + b.v_off, kk, w, l);
  if (tid == 0) atomicAdd(rot + t.blk, total);
}

__global__ __launch_bounds__(256) void svd_finish_kernel(const SvdOut* __restrict__ outs,
                                                         const double* __restrict__ Xn, const double* __restrict__ Va,
                                                         double* S, int kmax, double* Ua, int64_t su_a, double* Vta,
                                                         int64_t svt_a, int nmax, int zero_rank) {
  extern __shared__ double sig[];  // kk values, then kk ranks
  const SvdOut o = outs[blockIdx.x];
  const int kk = o.kk, lx = o.u_from_x ? o.m : o.n, lv = o.u_from_x ? o.n : o.m;
  int* rank = reinterpret_cast<int*>(sig + kk);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const double* xn = Xn + o.xn_off;
  const double* va = Va + o.va_off;
  for (int j = w; j < kk; j += 4) {
    double ss = 0.0;
    for (int i = lane; i < lx; i += 64) {
      const double u = xn[(int64_t)j * lx + i];
      ss += u * u;
    }
    ss = wave_sum(ss);
    if (lane == 0) sig[j] = sqrt(ss);
  }
  __syncthreads();
  // numerical rank (zero_rank != 0): values below max(m, n) eps |A|_F are at the rounding level of the decomposition
  // (the backward-error bound of an orthogonal one); they are reported as exactly zero, so noise directions are never
  // kept
  if (w == 0 && zero_rank) {
    double fro = 0.0;
    for (int j = lane; j < kk; j += 64) fro += sig[j] * sig[j];
    fro = wave_sum(fro);
    const double thr = (double)max(o.m, o.n) * 2.220446049250313e-16 * sqrt(fro);
    for (int j = lane; j < kk; j += 64)
      if (sig[j] < thr) sig[j] = 0.0;
  }
  __syncthreads();
  for (int j = tid; j < kk; j += 256) {
    const double v = sig[j];
    int r = 0;
    for (int i = 0; i < kk; ++i) r += (sig[i] > v || (sig[i] == v && i < j)) ? 1 : 0;
    rank[j] = r;
    S[(int64_t)blockIdx.x * kmax + r] = v;
  }
  __syncthreads();
  double* ua = Ua + (int64_t)blockIdx.x * su_a;
  double* vt = Vta + (int64_t)blockIdx.x * svt_a;
  const int64_t nx = (int64_t)lx * kk, nv = (int64_t)lv * kk;
  for (int64_t e = tid; e < nx; e += 256) {
    const int j = (int)(e / lx), i = (int)(e - (int64_t)j * lx);
    const double val = sig[j] > 0.0 ? xn[e] / sig[j] : 0.0;
    if (o.u_from_x) ua[(int64_t)i * kmax + rank[j]] = val;
    else vt[(int64_t)rank[j] * nmax + i] = val;
  }
  for (int64_t e = tid; e < nv; e += 256) {
    const int j = (int)(e / lv), i = (int)(e - (int64_t)j * lv);
    if (o.u_from_x) vt[(int64_t)rank[j] * nmax + i] = va[e];
    else ua[(int64_t)i * kmax + rank[j]] = va[e];
  }
}

// ---- blocked Householder QR of the Jacobi input (one work-group of 256 threads per block) ----
//
// Panel p (columns p0 .. p0+15, rows p0 .. len-1, K = len - p0 rows) is factored in LDS column by column (LAPACK's
// dlarfg: beta = -sign(alpha) |(alpha, x)|, tau = (beta - alpha) / beta, v = x / (alpha - beta), v_0 = 1), its
// triangular factor T is formed (dlarft, forward columnwise: T_ii = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) V^T v_i),
// and the trailing columns are updated in chunks of CW through LDS: C <- (I - V T^T V^T) C, the three products with
// FP64 MFMAs.  Q applied to a matrix Z (Z <- Q Z) runs the panels backwards: Z <- (I - V T V^T) Z.

__host__ __device__ __forceinline__ size_t svd_qr_lds_doubles(int cw, int ld) {
  return (size_t)ld * (kSvdQrPanel + cw) + 256 + 2 * 16 * cw + 32;
}

// Panel/chunk gather: dst[c * ld + i] = src[c * col_stride + i] for c < ncols, i < K; zero elsewhere up to
// ncols_alloc columns of ld rows.
__device__ __forceinline__ void svd_qr_gather(const double* __restrict__ src, int64_t col_stride, int K, int ncols,
                                              int ncols_alloc, int ld, double* dst, int tid) {
  // thread tid moves rows tid, tid + 256, ... of every column, up to 8 loads in flight
  for (int i = tid; i < K; i += 256) {
    for (int c0 = 0; c0 < ncols; c0 += 8) {
      double tmp[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (c0 + j < ncols) tmp[j] = src[(int64_t)(c0 + j) * col_stride + i];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (c0 + j < ncols) dst[(c0 + j) * ld + i] = tmp[j];
    }
  }
  const int pad = ld - K;
  for (int e = tid; e < ncols * pad; e += 256) dst[(e / pad) * ld + K + e % pad] = 0.0;
  for (int e = tid; e < (ncols_alloc - ncols) * ld; e += 256) dst[ncols * ld + e] = 0.0;
}

// C (K x CW in Cs, stride ld) <- C - V (Tt ? T^T : T) (V^T C); V explicit in Vp (16 columns, stride ld).
template <int CW>
__device__ __forceinline__ void svd_qr_block_reflect(const double* Vp, double* Cs, const double* Ts, double* Ws,
                                                     double* W2, int K, int ld, bool transT, int tid) {
  const int l = tid & 63, w = tid >> 6;
  constexpr int NCT = CW / 16;
  const int K4 = (K + 3) & ~3;
  if (w < NCT) {  // Ws = V^T C
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K4; k0 += 4) {
      const double a = Vp[(l % 16) * ld + k0 + l / 16];
      const double bb = Cs[(w * 16 + l % 16) * ld + k0 + l / 16];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) Ws[(l / 16 + 4 * j) * CW + w * 16 + l % 16] = acc[j];
  }
  __syncthreads();
  for (int e = tid; e < 16 * CW; e += 256) {  // W2 = T^T Ws (lower) or T Ws (upper)
    const int i = e / CW, c = e % CW;
    double acc = 0.0;
    if (transT) {
      for (int j = 0; j <= i; ++j) acc += Ts[j * 16 + i] * Ws[j * CW + c];
    } else {
      for (int j = i; j < 16; ++j) acc += Ts[i * 16 + j] * Ws[j * CW + c];
    }
    W2[e] = acc;
  }
  __syncthreads();
  const int nrt = (K + 15) / 16;
  for (int tile = w; tile < nrt * NCT; tile += 4) {  // C -= V W2
    const int row0 = (tile / NCT) * 16, ct = (tile % NCT) * 16;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4) {
      const double a = Vp[(k0 + l / 16) * ld + row0 + l % 16];
      const double bb = W2[(k0 + l / 16) * CW + ct + l % 16];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc, 0, 0, 0);
    }
    // D[l/16 + 4j][l%16] = (V W2)[row0 + l/16 + 4j][ct + l%16]
#pragma unroll
    for (int j = 0; j < 4; ++j) Cs[(ct + l % 16) * ld + row0 + l / 16 + 4 * j] -= acc[j];
  }
  __syncthreads();
}

template <int CW>
__global__ __launch_bounds__(256) void svd_qr_kernel(const SvdQrBlk* __restrict__ blks, double* X, double* tau,
                                                     double* T) {
  extern __shared__ double sm[];
  const SvdQrBlk b = blks[blockIdx.x];
  const int len = b.len, kk = b.kk, ld = svd_lds_ld(len);
  double* Vp = sm;
  double* Cs = Vp + ld * kSvdQrPanel;
  double* Ts = Cs + ld * CW;
  double* Ws = Ts + 256;
  double* W2 = Ws + 16 * CW;
  double* red = W2 + 16 * CW;  // 4 partial sums, tau, scale, beta
  double* tl = red + 16;       // the panel's tau values
  double* x = X + b.x_off;
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  for (int p0 = 0; p0 < kk; p0 += kSvdQrPanel) {
    const int nbp = min(kSvdQrPanel, kk - p0), K = len - p0;
    svd_qr_gather(x + (int64_t)p0 * len + p0, len, K, nbp, kSvdQrPanel, ld, Vp, tid);
    if (tid < 16) tl[tid] = 0.0;
    __syncthreads();
    for (int k = 0; k < nbp; ++k) {
      double* col = Vp + k * ld;
      double ss = 0.0;
      for (int i = k + 1 + tid; i < K; i += 256) ss += col[i] * col[i];
      ss = wave_sum(ss);
      if (l == 0) red[w] = ss;
      __syncthreads();
      if (tid == 0) {
        const double xn2 = red[0] + red[1] + red[2] + red[3], alpha = col[k];
        double tk = 0.0, scale = 0.0, beta = alpha;
        if (xn2 > 0.0) {
          beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
          tk = (beta - alpha) / beta;
          scale = 1.0 / (alpha - beta);
        }
        red[4] = tk;
        red[5] = scale;
        tl[k] = tk;
        tau[b.tau_off + p0 + k] = tk;
        col[k] = beta;
      }
      __syncthreads();
      const double tk = red[4], scale = red[5];
      for (int i = k + 1 + tid; i < K; i += 256) col[i] *= scale;
      __syncthreads();
      if (tk != 0.0) {  // H_k = I - tau v v^T on the panel's later columns, one wave per column
        for (int j = k + 1 + w; j < nbp; j += 4) {
          double* cj = Vp + j * ld;
          double d = 0.0;
          for (int i = k + 1 + l; i < K; i += 64) d += col[i] * cj[i];
          d = wave_sum(d) + cj[k];
          const double f = tk * d;
          for (int i = k + 1 + l; i < K; i += 64) cj[i] -= f * col[i];
          if (l == 0) cj[k] -= f;
        }
      }
      __syncthreads();
    }
    // the factored panel back to X (R on and above the diagonal, v below), then V made explicit in LDS
    for (int e = tid; e < nbp * K; e += 256) {
      const int c = e / K, i = e - c * K;
      x[(int64_t)(p0 + c) * len + p0 + i] = Vp[c * ld + i];
    }
    __syncthreads();
    for (int e = tid; e < nbp * K; e += 256) {
      const int c = e / K, i = e - c * K;
      if (i <= c) Vp[c * ld + i] = (i == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    // S = V^T V (one MFMA tile, wave 0) into Ws, then T column by column
    if (w == 0) {
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
      const int K4 = (K + 3) & ~3;
      for (int k0 = 0; k0 < K4; k0 += 4) {
        const double a = Vp[(l % 16) * ld + k0 + l / 16];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, acc, 0, 0, 0);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) Ws[(l / 16 + 4 * j) * 16 + l % 16] = acc[j];
    }
    __syncthreads();
    for (int i = 0; i < 16; ++i) {
      if (tid < 16) {
        const int j = tid;
        double tv = 0.0;
        if (j < i) {
          double acc = 0.0;
          for (int m = j; m < i; ++m) acc += Ts[j * 16 + m] * Ws[m * 16 + i];
          tv = -tl[i] * acc;
        } else if (j == i) {
          tv = tl[i];
        }
        Ts[j * 16 + i] = tv;
      }
      __syncthreads();
    }
    for (int e = tid; e < 256; e += 256) T[b.t_off + (int64_t)(p0 / kSvdQrPanel) * 256 + e] = Ts[e];
    // trailing columns in chunks of CW: C <- (I - V T^T V^T) C
    for (int c0 = p0 + nbp; c0 < kk; c0 += CW) {
      const int cw = min(CW, kk - c0);
      svd_qr_gather(x + (int64_t)c0 * len + p0, len, K, cw, CW, ld, Cs, tid);
      __syncthreads();
      svd_qr_block_reflect<CW>(Vp, Cs, Ts, Ws, W2, K, ld, true, tid);
      for (int e = tid; e < cw * K; e += 256) {
        const int c = e / K, i = e - c * K;
        x[(int64_t)(c0 + c) * len + p0 + i] = Cs[c * ld + i];
      }
      __syncthreads();
    }
  }
}

// The same factorization split into launches per panel, so that the trailing update, the bulk of the work, runs on
// one work-group per 16-column chunk instead of one per block: svd_qr_panel_kernel factors panel `pi` of every block
// (one work-group per block; blocks with fewer columns return) and writes it back with its tau and T, then
// svd_qr_trail_kernel applies that panel's block reflector to every chunk (block, c0) right of it (other items
// return).  Each column goes through the same operations as in svd_qr_kernel.
__global__ __launch_bounds__(256) void svd_qr_panel_kernel(const SvdQrBlk* __restrict__ blks, double* X, double* tau,
                                                           double* T, int pi) {
  extern __shared__ double sm[];
  const SvdQrBlk b = blks[blockIdx.x];
  const int len = b.len, kk = b.kk, ld = svd_lds_ld(len), p0 = pi * kSvdQrPanel;
  if (p0 >= kk) return;
  double* Vp = sm;
  double* Ts = Vp + ld * kSvdQrPanel;
  double* Ws = Ts + 256;
  double* red = Ws + 256;  // 4 partial sums, tau, scale
  double* tl = red + 16;   // the panel's tau values
  double* x = X + b.x_off;
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int nbp = min(kSvdQrPanel, kk - p0), K = len - p0;
  svd_qr_gather(x + (int64_t)p0 * len + p0, len, K, nbp, kSvdQrPanel, ld, Vp, tid);
  if (tid < 16) tl[tid] = 0.0;
  __syncthreads();
  for (int k = 0; k < nbp; ++k) {
    double* col = Vp + k * ld;
    double ss = 0.0;
    for (int i = k + 1 + tid; i < K; i += 256) ss += col[i] * col[i];
    ss = wave_sum(ss);
    if (l == 0) red[w] = ss;
    __syncthreads();
    if (tid == 0) {
      const double xn2 = red[0] + red[1] + red[2] + red[3], alpha = col[k];
      double tk = 0.0, scale = 0.0, beta = alpha;
      if (xn2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tk = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      red[4] = tk;
      red[5] = scale;
      tl[k] = tk;
      tau[b.tau_off + p0 + k] = tk;
      col[k] = beta;
    }
    __syncthreads();
    const double tk = red[4], scale = red[5];
    for (int i = k + 1 + tid; i < K; i += 256) col[i] *= scale;
    __syncthreads();
    if (tk != 0.0) {  // H_k = I - tau v v^T on the panel's later columns, one wave per column
      for (int j = k + 1 + w; j < nbp; j += 4) {
        double* cj = Vp + j * ld;
        double d = 0.0;
        for (int i = k + 1 + l; i < K; i += 64) d += col[i] * cj[i];
        d = wave_sum(d) + cj[k];
        const double f = tk * d;
        for (int i = k + 1 + l; i < K; i += 64) cj[i] -= f * col[i];
        if (l == 0) cj[k] -= f;
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < nbp * K; e += 256) {
    const int c = e / K, i = e - c * K;
    x[(int64_t)(p0 + c) * len + p0 + i] = Vp[c * ld + i];
  }
  __syncthreads();
  for (int e = tid; e < nbp * K; e += 256) {
    const int c = e / K, i = e - c * K;
    if (i <= c) Vp[c * ld + i] = (i == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (w == 0) {  // S = V^T V (one MFMA tile) into Ws, then T column by column
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    const int K4 = (K + 3) & ~3;
    for (int k0 = 0; k0 < K4; k0 += 4) {
      const double a = Vp[(l % 16) * ld + k0 + l / 16];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) Ws[(l / 16 + 4 * j) * 16 + l % 16] = acc[j];
  }
  __syncthreads();
  for (int i = 0; i < 16; ++i) {
    if (tid < 16) {
      const int j = tid;
      double tv = 0.0;
      if (j < i) {
        double acc = 0.0;
        for (int m = j; m < i; ++m) acc += Ts[j * 16 + m] * Ws[m * 16 + i];
        tv = -tl[i] * acc;
      } else if (j == i) {
        tv = tl[i];
      }
      Ts[j * 16 + i] = tv;
    }
    __syncthreads();
  }
  for (int e = tid; e < 256; e += 256) T[b.t_off + (int64_t)pi * 256 + e] = Ts[e];
}

__global__ __launch_bounds__(256) void svd_qr_trail_kernel(const int2* __restrict__ items,
                                                           const SvdQrBlk* __restrict__ blks, double* X,
                                                           const double* __restrict__ T, int pi) {
  constexpr int CW = kSvdQrPanel;
  extern __shared__ double sm[];
  const int2 it = items[blockIdx.x];
  const SvdQrBlk b = blks[it.x];
  const int len = b.len, kk = b.kk, ld = svd_lds_ld(len), p0 = pi * kSvdQrPanel, c0 = it.y;
  if (c0 < p0 + kSvdQrPanel || c0 >= kk) return;  // uniform: the chunk is not right of this panel
  const int nbp = min(kSvdQrPanel, kk - p0), K = len - p0, cw = min(CW, kk - c0);
  double* Vp = sm;
  double* Cs = Vp + ld * kSvdQrPanel;
  double* Ts = Cs + ld * CW;
  double* Ws = Ts + 256;
  double* W2 = Ws + 16 * CW;
  double* x = X + b.x_off;
  const int tid = threadIdx.x;
  for (int e = tid; e < kSvdQrPanel * ld; e += 256) {
    const int c = e / ld, i = e - c * ld;
    double v = 0.0;
    if (c < nbp && i < K) v = i < c ? 0.0 : (i == c ? 1.0 : x[(int64_t)(p0 + c) * len + p0 + i]);
    Vp[e] = v;
  }
  for (int e = tid; e < 256; e += 256) Ts[e] = T[b.t_off + (int64_t)pi * 256 + e];
  svd_qr_gather(x + (int64_t)c0 * len + p0, len, K, cw, CW, ld, Cs, tid);
  __syncthreads();
  svd_qr_block_reflect<CW>(Vp, Cs, Ts, Ws, W2, K, ld, true, tid);
  for (int i = tid; i < K; i += 256)
    for (int c = 0; c < cw; ++c) x[(int64_t)(c0 + c) * len + p0 + i] = Cs[c * ld + i];
}

__global__ __launch_bounds__(256) void svd_rt_kernel(const SvdQrBlk* __restrict__ blks, const double* __restrict__ X,
                                                     double* Y) {
  const SvdQrBlk b = blks[blockIdx.x];
  const int kk = b.kk;
  for (int64_t e = threadIdx.x; e < (int64_t)kk * kk; e += 256) {
    const int j = (int)(e / kk), i = (int)(e - (int64_t)j * kk);  // Y(i, j) = R(j, i), R(j, i) = X(row j, col i)
    Y[b.y_off + e] = i >= j ? X[b.x_off + (int64_t)i * b.len + j] : 0.0;
  }
}

// Z = Q [W; 0] for the CW-wide column chunk c0 of one block per work-group (the columns are independent, so the
// chunks of one block run in parallel): the chunk starts as [W; 0], then the panels' block reflectors are applied
// backwards, Z <- (I - V T V^T) Z, each through LDS.
template <int CW>
__global__ __launch_bounds__(256) void svd_qr_apply_kernel(const int2* __restrict__ items,
                                                           const SvdQrBlk* __restrict__ blks,
                                                           const double* __restrict__ X, const double* __restrict__ T,
                                                           const double* __restrict__ W, double* Z) {
  extern __shared__ double sm[];
  const int2 it = items[blockIdx.x];
  const SvdQrBlk b = blks[it.x];
  const int len = b.len, kk = b.kk, ld = svd_lds_ld(len), c0 = it.y, cw = min(CW, kk - c0);
  double* Vp = sm;
  double* Cs = Vp + ld * kSvdQrPanel;
  double* Ts = Cs + ld * CW;
  double* Ws = Ts + 256;
  double* W2 = Ws + 16 * CW;
  double* z = Z + b.z_off + (int64_t)c0 * len;
  const int tid = threadIdx.x;
  for (int e = tid; e < len * cw; e += 256) {  // the chunk of Z = [W; 0]
    const int j = e / len, i = e - j * len;
    z[e] = i < kk ? W[b.w_off + (int64_t)(c0 + j) * kk + i] : 0.0;
  }
  __syncthreads();
  const int np = (kk + kSvdQrPanel - 1) / kSvdQrPanel;
  for (int pi = np - 1; pi >= 0; --pi) {
    const int p0 = pi * kSvdQrPanel, nbp = min(kSvdQrPanel, kk - p0), K = len - p0;
    for (int e = tid; e < kSvdQrPanel * ld; e += 256) {
      const int c = e / ld, i = e - c * ld;
      double v = 0.0;
      if (c < nbp && i < K) v = i < c ? 0.0 : (i == c ? 1.0 : X[b.x_off + (int64_t)(p0 + c) * len + p0 + i]);
      Vp[e] = v;
    }
    for (int e = tid; e < 256; e += 256) Ts[e] = T[b.t_off + (int64_t)pi * 256 + e];
    svd_qr_gather(z + p0, len, K, cw, CW, ld, Cs, tid);
    __syncthreads();
    svd_qr_block_reflect<CW>(Vp, Cs, Ts, Ws, W2, K, ld, false, tid);
    for (int i = tid; i < K; i += 256)
      for (int c = 0; c < cw; ++c) z[(int64_t)c * len + p0 + i] = Cs[c * ld + i];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void svd_sort_kernel(const SvdQrBlk* __restrict__ blks, const double* __restrict__ X,
                                                       double* Xs, int32_t* perm) {
  extern __shared__ double nrm[];  // kk squared norms, then kk ranks
  const SvdQrBlk b = blks[blockIdx.x];
  const int len = b.len, kk = b.kk;
  int* rank = reinterpret_cast<int*>(nrm + kk);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const double* x = X + b.x_off;
  for (int j = w; j < kk; j += 4) {
    double ss = 0.0;
    for (int i = lane; i < len; i += 64) {
      const double u = x[(int64_t)j * len + i];
      ss += u * u;
    }
    ss = wave_sum(ss);
    if (lane == 0) nrm[j] = ss == ss ? ss : -1.0;  // NaN last, so the order stays total and rank a permutation
  }
  __syncthreads();
  for (int j = tid; j < kk; j += 256) {
    const double v = nrm[j];
    int r = 0;
    for (int i = 0; i < kk; ++i) r += (nrm[i] > v || (nrm[i] == v && i < j)) ? 1 : 0;
    rank[j] = r;
    perm[b.tau_off + r] = j;
  }
  __syncthreads();
  double* xs = Xs + b.x_off;
  for (int64_t e = tid; e < (int64_t)len * kk; e += 256) {
    const int c = (int)(e / len);
    xs[(int64_t)rank[c] * len + (e - (int64_t)c * len)] = x[e];
  }
}

__global__ __launch_bounds__(256) void svd_unsort_rows_kernel(const SvdQrBlk* __restrict__ blks,
                                                              const int32_t* __restrict__ perm,
                                                              const double* __restrict__ Y, double* Yu) {
  const SvdQrBlk b = blks[blockIdx.x];
  const int kk = b.kk;
  for (int64_t e = threadIdx.x; e < (int64_t)kk * kk; e += 256) {
    const int c = (int)(e / kk), i = (int)(e - (int64_t)c * kk), p = perm[b.tau_off + i];
    if (p >= 0 && p < kk) Yu[b.y_off + (int64_t)c * kk + p] = Y[b.y_off + e];
  }
}

__global__ __launch_bounds__(256) void svd_sweep_end_kernel(int32_t* rot, int32_t* active, int nb, int32_t* any) {
  int a = 0;
  for (int k = threadIdx.x; k < nb; k += 256) {
    const int r = rot[k] > 0 ? 1 : 0;
    active[k] = r;
    rot[k] = 0;
    a |= r;
  }
  a = __syncthreads_or(a);
  if (threadIdx.x == 0) *any = a;
}

void launch_svd_sweep_end(int32_t* rot, int32_t* active, int nb, int32_t* any, hipStream_t s) {
  hipLaunchKernelGGL(svd_sweep_end_kernel, dim3(1), dim3(256), 0, s, rot, active, nb, any);
}

void launch_svd_sort(const SvdQrBlk* blks, int nb, const double* X, double* Xs, int32_t* perm, int kk_max,
                     hipStream_t s) {
  if (nb <= 0) return;
  (void)svd_round_lds_limit();  // raises the kernels' dynamic LDS limit once
  const size_t lds = (size_t)kk_max * (sizeof(double) + sizeof(int));
  hipLaunchKernelGGL(svd_sort_kernel, dim3(nb), dim3(256), lds, s, blks, X, Xs, perm);
}

void launch_svd_unsort_rows(const SvdQrBlk* blks, int nb, const int32_t* perm, const double* Y, double* Yu,
                            hipStream_t s) {
  if (nb > 0) hipLaunchKernelGGL(svd_unsort_rows_kernel, dim3(nb), dim3(256), 0, s, blks, perm, Y, Yu);
}

size_t svd_round_lds_limit() {
  static size_t lim = 0;
  if (!lim) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                                                  hipSuccess || v <= 0)
      v = 65536;
    const void* fns[] = {(const void*)svd_round_kernel<8>,     (const void*)svd_round_kernel<16>,
                         (const void*)svd_round_kernel<32>,    (const void*)svd_finish_kernel,
                         (const void*)svd_qr_kernel<16>,       (const void*)svd_qr_kernel<32>,
                         (const void*)svd_qr_apply_kernel<16>, (const void*)svd_qr_apply_kernel<32>,
                         (const void*)svd_sort_kernel,         (const void*)svd_qr_panel_kernel,
                         (const void*)svd_qr_trail_kernel};
    // the dynamic limit leaves room for the largest static LDS among these kernels (the round kernel's G buffers)
    size_t stat = 0;
    for (const void* f : fns) {
      hipFuncAttributes fa;
      if (hipFuncGetAttributes(&fa, f) == hipSuccess) stat = std::max(stat, (size_t)fa.sharedSizeBytes);
    }
    stat = std::max(stat, (size_t)(3 * 32 * 33 * sizeof(double) + 64));
    lim = (size_t)v - stat;
    for (const void* f : fns) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim);
    (void)hipGetLastError();  // a refused attribute leaves the default limit; launches check their own sizes
  }
  return lim;
}

size_t svd_round_lds_bytes(int cols, int len_max) { return (size_t)cols * svd_lds_ld(len_max) * sizeof(double); }

size_t svd_qr_lds_bytes(int cw, int len_max) { return svd_qr_lds_doubles(cw, svd_lds_ld(len_max)) * sizeof(double); }

void launch_svd_qr(const SvdQrBlk* blks, int nb, double* X, double* tau, double* T, int cw, int len_max,
                   hipStream_t s) {
  if (nb <= 0) return;
  const size_t lds = svd_qr_lds_bytes(cw, len_max);
  if (cw == 32) hipLaunchKernelGGL(svd_qr_kernel<32>, dim3(nb), dim3(256), lds, s, blks, X, tau, T);
  else hipLaunchKernelGGL(svd_qr_kernel<16>, dim3(nb), dim3(256), lds, s, blks, X, tau, T);
}

void launch_svd_qr_split(const SvdQrBlk* blks, int nb, const int2* items, int nitems, int npanels, double* X,
                         double* tau, double* T, int len_max, hipStream_t s) {
  if (nb <= 0) return;
  const size_t lds = svd_qr_lds_bytes(kSvdQrPanel, len_max);
  for (int pi = 0; pi < npanels; ++pi) {
    hipLaunchKernelGGL(svd_qr_panel_kernel, dim3(nb), dim3(256), lds, s, blks, X, tau, T, pi);
    if (nitems > 0 && pi + 1 < npanels)
      hipLaunchKernelGGL(svd_qr_trail_kernel, dim3(nitems), dim3(256), lds, s, items, blks, X, T, pi);
  }
}

void launch_svd_rt(const SvdQrBlk* blks, int nb, const double* X, double* Y, hipStream_t s) {
  if (nb > 0) hipLaunchKernelGGL(svd_rt_kernel, dim3(nb), dim3(256), 0, s, blks, X, Y);
}

void launch_svd_qr_apply(const int2* items, int nitems, const SvdQrBlk* blks, const double* X, const double* T,
                         const double* W, double* Z, int cw, int len_max, hipStream_t s) {
  if (nitems <= 0) return;
  const size_t lds = svd_qr_lds_bytes(cw, len_max);
  if (cw == 32) hipLaunchKernelGGL(svd_qr_apply_kernel<32>, dim3(nitems), dim3(256), lds, s, items, blks, X, T, W, Z);
  else hipLaunchKernelGGL(svd_qr_apply_kernel<16>, dim3(nitems), dim3(256), lds, s, items, blks, X, T, W, Z);
}

void launch_svd_init(const SvdBlk* blks, int nb, double* V, hipStream_t s) {
  if (nb > 0) hipLaunchKernelGGL(svd_init_kernel, dim3(nb), dim3(256), 0, s, blks, V);
}

void launch_svd_round(const SvdTask* tasks, int ntasks, const SvdBlk* blks, double* X, double* V,
                      const int32_t* active, int32_t* rot, int cols, int len_max, hipStream_t s) {
  if (ntasks <= 0) return;
  const size_t lds = svd_round_lds_bytes(cols, len_max);
  if (cols == 32)
    hipLaunchKernelGGL(svd_round_kernel<32>, dim3(ntasks), dim3(256), lds, s, tasks, blks, X, V, active, rot);
  else if (cols == 16)
    hipLaunchKernelGGL(svd_round_kernel<16>, dim3(ntasks), dim3(256), lds, s, tasks, blks, X, V, active, rot);
  else
    hipLaunchKernelGGL(svd_round_kernel<8>, dim3(ntasks), dim3(256), lds, s, tasks, blks, X, V, active, rot);
}

void launch_svd_finish(const SvdOut* outs, int nb, const double* Xn, const double* Va, double* S, int kmax, double* Ua,
                       int64_t su_a, double* Vta, int64_t svt_a, int nmax, int kk_max, int zero_rank, hipStream_t s) {
  if (nb <= 0) return;
  const size_t lds = (size_t)kk_max * (sizeof(double) + sizeof(int));
  hipLaunchKernelGGL(svd_finish_kernel, dim3(nb), dim3(256), lds, s, outs, Xn, Va, S, kmax, Ua, su_a, Vta, svt_a,
                     nmax, zero_rank);
}

void launch_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, s, a, x, b, y, n);
}

template <typename TX, typename TY>
void launch_convert(const TX* x, TY* y, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL((convert_kernel<TX, TY>), dim3(grid_for(n)), dim3(256), 0, s, x, y, n);
}

void launch_combine(const double* base, int64_t ld, int m, const double* c, double* y, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, s, base, ld, m, c, y, n);
}

void launch_precondition(double* r, const double* diag, const unsigned char* present, double eval, int64_t n,
                         hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(precondition_kernel, dim3(grid_for(n)), dim3(256), 0, s, r, diag, present, eval, n);
}

void launch_combine_sub(const double* base, int64_t ld, int m, const double* c, double* y, int64_t n, hipStream_t s) {
  if (n > 0 && m > 0) hipLaunchKernelGGL(combine_sub_kernel, dim3(grid_for(n)), dim3(256), 0, s, base, ld, m, c, y, n);
}

void launch_scale_rnorm(double* x, int64_t n, const double* nrm2, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(scale_rnorm_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, n, nrm2);
}

void launch_rotate2(const double* T, const double* S, int64_t ld, int m, const double* c, int cld, double* TO,
                    double* SO, int64_t n, hipStream_t s) {
  if (n > 0 && m > 0)
    hipLaunchKernelGGL(rotate2_kernel, dim3(grid_for(n)), dim3(256), 0, s, T, S, ld, m, c, cld, TO, SO, n);
}

void launch_residual(const double* s0, const double* t0, double e, double* r, int64_t n, double* G, double* partial,
                     int npart, hipStream_t s) {
  hipLaunchKernelGGL(residual_partial_kernel, dim3(npart), dim3(256), 0, s, s0, t0, e, r, n, partial, npart);
  hipLaunchKernelGGL(gram_reduce_kernel, dim3(1), dim3(256), 0, s, partial, npart, G);
}

void launch_gram(const double* X, int64_t ldx, int mx, const double* Y, int64_t ldy, int my, int64_t n, double* G,
                 double* partial, int npart, hipStream_t s) {
  if (mx <= 0 || my <= 0) return;
  hipLaunchKernelGGL(gram_partial_kernel, dim3(npart, mx * my), dim3(256), 0, s, X, ldx, Y, ldy, my, n, partial,
                     npart);
  hipLaunchKernelGGL(gram_reduce_kernel, dim3(mx * my), dim3(256), 0, s, partial, npart, G);
}

void launch_seg_gram(const SegChunk* ch, int nch, const int32_t* chunk0, int ninst, const double* X, int64_t ldx,
                     int mx, const double* Y, int64_t ldy, int my, double* G, int gstride, double* partial,
                     hipStream_t s) {
  if (nch <= 0 || ninst <= 0 || mx <= 0 || my <= 0) return;
  hipLaunchKernelGGL(seg_gram_partial_kernel, dim3(nch, mx * my), dim3(256), 0, s, ch, X, ldx, Y, ldy, my, partial,
                     nch);
  hipLaunchKernelGGL(seg_reduce_kernel, dim3(ninst, mx * my), dim3(256), 0, s, partial, nch, chunk0, G, gstride);
}

void launch_seg_residual(const SegChunk* ch, int nch, const int32_t* chunk0, int ninst, const double* s0,
                         const double* t0, const double* e, double* r, double* G, int gstride, double* partial,
                         hipStream_t s) {
  if (nch <= 0 || ninst <= 0) return;
  hipLaunchKernelGGL(seg_residual_partial_kernel, dim3(nch), dim3(256), 0, s, ch, s0, t0, e, r, partial);
  hipLaunchKernelGGL(seg_reduce_kernel, dim3(ninst, 1), dim3(256), 0, s, partial, nch, chunk0, G, gstride);
}

void launch_seg_rotate2(const SegChunk* ch, int nch, const double* T, const double* S, int64_t ld, int m,
                        const double* c, int cld, int cstride, double* TO, double* SO, hipStream_t s) {
  if (nch > 0 && m > 0)
    hipLaunchKernelGGL(seg_rotate2_kernel, dim3(nch), dim3(256), 0, s, ch, T, S, ld, m, c, cld, cstride, TO, SO);
}

void launch_seg_combine_sub(const SegChunk* ch, int nch, const double* base, int64_t ld, int m, const double* c,
                            int cstride, double* y, hipStream_t s) {
  if (nch > 0 && m > 0)
    hipLaunchKernelGGL(seg_combine_sub_kernel, dim3(nch), dim3(256), 0, s, ch, base, ld, m, c, cstride, y);
}

void launch_seg_scale_rnorm(const SegChunk* ch, int nch, double* x, const double* nrm2, int stride, hipStream_t s) {
  if (nch > 0) hipLaunchKernelGGL(seg_scale_rnorm_kernel, dim3(nch), dim3(256), 0, s, ch, x, nrm2, stride);
}

void launch_seg_precondition(const SegChunk* ch, int nch, double* r, const double* diag, const unsigned char* present,
                             const double* eval, hipStream_t s) {
  if (nch > 0) hipLaunchKernelGGL(seg_precondition_kernel, dim3(nch), dim3(256), 0, s, ch, r, diag, present, eval);
}

void launch_seg_copy(const SegChunk* ch, int nch, const double* src, double* dst, const int32_t* mask,
                     hipStream_t s) {
  if (nch > 0) hipLaunchKernelGGL(seg_copy_kernel, dim3(nch), dim3(256), 0, s, ch, src, dst, mask);
}

template <typename T, typename TC>
void launch_gemm_small(const GemmTile* tiles, int ntiles, const GemmTask* tasks, const GemmPair* pairs, Operands ops,
                       double alpha, int accumulate, hipStream_t s, int max_groups) {
  if (ntiles <= 0) return;
  if (ops.table)
    hipLaunchKernelGGL((gemm_small_kernel<T, TC, true>), dim3(groups((ntiles + 3) / 4, max_groups)), dim3(256), 0, s,
                       tiles, ntiles, tasks, pairs, ops, alpha, accumulate);
  else
    hipLaunchKernelGGL((gemm_small_kernel<T, TC, false>), dim3(groups((ntiles + 3) / 4, max_groups)), dim3(256), 0, s,
                       tiles, ntiles, tasks, pairs, ops, alpha, accumulate);
}

template <typename T, typename TC>
void launch_gemm_split(const BigTile* tiles, int ntiles, const RedTile* red, int nred, void* partial,
                       const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                       hipStream_t s, int max_groups) {
  if (ntiles <= 0) return;
  T* part = static_cast<T*>(partial);
  if (ops.table) {
    hipLaunchKernelGGL((gemm_split_kernel<T, TC, true, kSplitWaves>), dim3(groups(ntiles, max_groups)),
                       dim3(64 * kSplitWaves), 0, s, tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part);
    if (nred > 0)
      hipLaunchKernelGGL((gemm_reduce32_kernel<T, TC, true>), dim3(groups(nred, max_groups)), dim3(256), 0, s, red,
                         nred, tasks, part, ops, alpha, accumulate);
  } else {
    hipLaunchKernelGGL((gemm_split_kernel<T, TC, false, kSplitWaves>), dim3(groups(ntiles, max_groups)),
                       dim3(64 * kSplitWaves), 0, s, tiles, ntiles, tasks, pairs, ops, alpha, accumulate, part);
    if (nred > 0)
      hipLaunchKernelGGL((gemm_reduce32_kernel<T, TC, false>), dim3(groups(nred, max_groups)), dim3(256), 0, s, red,
                         nred, tasks, part, ops, alpha, accumulate);
  }
}

#define MPX_INST_BIG(T, TC)                                                                                   \
  template void launch_gemm<T, TC>(const BigTile*, const int (&)[4], const RedTile*, int, void*, const GemmTask*,    \
                                   const GemmPair*, \
                                   Operands, double, int, hipStream_t, int);
MPX_INST_BIG(float, float)
MPX_INST_BIG(float, double)
#undef MPX_INST_BIG
template void launch_gemm128<float, float>(const BigTile*, const int (&)[4], const RedTile*, int, void*, const GemmTask*,
                                           const GemmPair*, Operands, double, int, hipStream_t, int);
template void launch_gemm128<float, double>(const BigTile*, const int (&)[4], const RedTile*, int, void*, const GemmTask*,
                                            const GemmPair*, Operands, double, int, hipStream_t, int);
#define MPX_INST_GEMM(FN, T, TC) \
  template void FN<T, TC>(const GemmTile*, int, const GemmTask*, const GemmPair*, Operands, double, int, hipStream_t, \
                           int);
MPX_INST_GEMM(launch_gemm_small, double, double)
MPX_INST_GEMM(launch_gemm_small, float, float)
MPX_INST_GEMM(launch_gemm_small, float, double)
#undef MPX_INST_GEMM
#define MPX_INST_SPLIT(T, TC)                                                                                     \
  template void launch_gemm_split<T, TC>(const BigTile*, int, const RedTile*, int, void*, const GemmTask*,       \
                                         const GemmPair*, Operands, double, int, hipStream_t, int);
MPX_INST_SPLIT(double, double)
MPX_INST_SPLIT(float, float)
MPX_INST_SPLIT(float, double)
#undef MPX_INST_SPLIT
template void launch_axpy<double, double>(const Chunk*, int, const AxpyTask*, const AxpyPair*, Operands, double, int,
                                          hipStream_t, int);
template void launch_axpy<float, float>(const Chunk*, int, const AxpyTask*, const AxpyPair*, Operands, double, int,
                                        hipStream_t, int);
template void launch_axpy<float, double>(const Chunk*, int, const AxpyTask*, const AxpyPair*, Operands, double, int,
                                         hipStream_t, int);
template void launch_copy<double, double>(const Chunk*, int, const CopySeg*, const double*, double*, double, double,
                                          hipStream_t);
template void launch_copy<float, float>(const Chunk*, int, const CopySeg*, const float*, float*, double, double,
                                        hipStream_t);
template void launch_copy<double, float>(const Chunk*, int, const CopySeg*, const double*, float*, double, double,
                                         hipStream_t);
template void launch_copy<float, double>(const Chunk*, int, const CopySeg*, const float*, double*, double, double,
                                         hipStream_t);
template void launch_convert<double, float>(const double*, float*, int64_t, hipStream_t);
template void launch_convert<float, double>(const float*, double*, int64_t, hipStream_t);

}  // namespace mpx
