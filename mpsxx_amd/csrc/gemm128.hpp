// 128x128 grouped-GEMM tiles for large blocks (both sides >= GemmPolicy::big_tile_min), FP64 and FP32 operands; device
// code, included only by .hip files (FP32 instantiated in kernels.hip, FP64 in kernels_f64_128.hip).
//
// One 256-thread work-group per 128x128 C tile, 4 waves each owning a 64x64 quarter computed as 4x4
// v_mfma_{f64,f32}_16x16x4 tiles: per 4-wide k step a wave reads 4 A and 4 B operands from LDS and issues 16 MFMAs,
// four times the MFMAs per operand byte of the 64x64 tiles.  A and B panels (128x16, 16x128) are staged in LDS k-major
// and double-buffered; the next panel is loaded into registers while the MFMAs run on the current one.  Each thread
// moves 8 elements of each operand per panel through buffer loads against a resource covering the operand block: an
// element outside the block gets an offset beyond the extent and reads as zero, so every panel (inside the block or
// on its edge, and the empty panel past the end of the tile) goes through the same branch-free loads and the compiler
// counts the loads in flight exactly.  The kernel is specialised per operand layout (the host groups the tiles by
// their task's layout flags, one launch per group):
//   unit stride along k: a thread takes 4 consecutive k of lines t/4 and t/4 + 64 (k 4*(t%4)..+3), one 16-byte load
//     per line for FP32 and two for FP64; k past the pair's end is zeroed when the panel is stored;
//   otherwise: a thread takes k t/128 + 2*i (i = 0..7) of line t%128, one element per load, 128 consecutive lines
//     per load instruction over the work-group.
// 16x16 sub-tiles that lie wholly in the padding of an edge tile skip their MFMAs.
#pragma once

#include "gemm_tiles.hpp"

namespace mpx {

constexpr int B2M = 128, B2N = 128, LDP2 = 128 + 4;
constexpr int32_t kOutOfBlock = 0x7fffffff;  // a byte offset beyond every block extent (the host keeps blocks < 2^30 B)

// One operand of the current pair: the buffer resource of its block and the thread's offsets.
struct G2View {
  __amdgpu_buffer_rsrc_t rsrc;
  int32_t boff0;  // byte offset of the thread's first element for k0 = 0
  int32_t brun;   // byte offset of the thread's second line (unit-stride-k layout) or second k (otherwise)
  int32_t bsk;    // bytes per unit of k
  int32_t lim;    // lines of the block from the tile's first line on
};

template <bool Contig>
struct G2Map {
  static __device__ __forceinline__ int l0(int tid) { return Contig ? tid >> 2 : tid & 127; }
  static __device__ __forceinline__ int kk0(int tid) { return Contig ? 4 * (tid & 3) : tid >> 7; }
};

template <typename T, bool Contig>
__device__ __forceinline__ void g2_view(G2View& v, int tid, const T* __restrict__ block, int line0, int nline,
                                        int klen, int64_t s_line, int64_t s_k) {
  typedef G2Map<Contig> M;
  // elements spanned by the block (0 for an empty one: then every element reads as zero)
  const int64_t extent = nline > 0 && klen > 0 ? (int64_t)(nline - 1) * s_line + (int64_t)(klen - 1) * s_k + 1 : 0;
  v.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(block), (short)0, (int)(extent * (int64_t)sizeof(T)),
                                             0x00020000);
  v.boff0 = (int32_t)(sizeof(T) * ((int64_t)(line0 + M::l0(tid)) * s_line + (int64_t)M::kk0(tid) * s_k));
  v.brun = (int32_t)(sizeof(T) * (Contig ? 64 * s_line : 2 * s_k));
  v.bsk = (int32_t)(sizeof(T) * s_k);
  v.lim = nline - line0;
}

template <typename T>
__device__ __forceinline__ T g2_load1(__amdgpu_buffer_rsrc_t r, int32_t off) {
  if constexpr (sizeof(T) == 8) return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// The thread's 8 elements of the panel at k0 (valid = false past the tile's last panel: zeros).  kl = k of the
// thread's first element left in the pair (unit-stride-k layout: elements i >= kl of a line are zeroed at the store).
template <typename T, bool Contig>
__device__ __forceinline__ void g2_load(T (&r)[8], int& kl, const G2View& v, int tid, int k0, int klen, bool valid) {
  typedef G2Map<Contig> M;
  const int32_t base = v.boff0 + k0 * v.bsk;
  kl = klen - k0 - M::kk0(tid);
  if constexpr (Contig) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef double d2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int run = 0; run < 2; ++run) {
      const bool ok = valid && M::l0(tid) + 64 * run < v.lim && kl > 0;
      const int32_t off = ok ? base + run * v.brun : kOutOfBlock;
      if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const d2 x = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, ok ? off + 16 * h : off, 0, 0));
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
    for (int i = 0; i < 8; ++i) {
      const int32_t off = (lok && 2 * i < kl) ? base + i * v.brun : kOutOfBlock;
      r[i] = g2_load1<T>(v.rsrc, off);
    }
  }
}

template <typename T, bool Contig>
__device__ __forceinline__ void g2_store(T (*S)[LDP2], const T (&r)[8], int kl, int tid) {
  typedef G2Map<Contig> M;
  const int line = M::l0(tid), k = M::kk0(tid);
  if constexpr (Contig) {
#pragma unroll
    for (int run = 0; run < 2; ++run)
#pragma unroll
      for (int i = 0; i < 4; ++i) S[k + i][line + 64 * run] = i < kl ? r[4 * run + i] : T(0);
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) S[k + 2 * i][line] = r[i];
  }
}

// FP64: one work-group per CU (each wave has the whole 512-entry register file for its 128 accumulator registers and
// the panel prefetch; the 64-cycle FP64 MFMAs of one panel cover the prefetch latency); FP32: two
template <typename T, typename TC, bool Tab, bool CA, bool CB>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 1 : 2) void gemm128_kernel(const BigTile* __restrict__ tiles,
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

    // the load cursor (cp, ck0) walks the tile's (pair, k panel) sequence; va, vb view the cursor's pair
    int cp = tl.pair_lo, ck0 = 0;
    GemmPair cpr = pairs[cp];
    G2View va, vb;
    g2_view<T, CA>(va, tid, A + cpr.a_off, tl.row0, tk.m, cpr.k, cpr.sa_m, cpr.sa_k);
    g2_view<T, CB>(vb, tid, B + cpr.b_off, tl.col0, tk.n, cpr.k, cpr.sb_n, cpr.sb_k);
    bool cur = true;  // the cursor is on a panel of the tile (uniform)
    T ra[8], rb[8];
    int kla, klb;
    auto load = [&]() {
      g2_load<T, CA>(ra, kla, va, tid, ck0, cpr.k, cur);
      g2_load<T, CB>(rb, klb, vb, tid, ck0, cpr.k, cur);
      if (!cur) return;
      ck0 += BK;
      if (ck0 < cpr.k) return;
      ck0 = 0;
      if (++cp >= tl.pair_hi) {
        cur = false;
        return;
      }
      cpr = pairs[cp];
      g2_view<T, CA>(va, tid, A + cpr.a_off, tl.row0, tk.m, cpr.k, cpr.sa_m, cpr.sa_k);
      g2_view<T, CB>(vb, tid, B + cpr.b_off, tl.col0, tk.n, cpr.k, cpr.sb_n, cpr.sb_k);
    };
    load();
    g2_store<T, CA>(As[0], ra, kla, tid);
    g2_store<T, CB>(Bs[0], rb, klb, tid);
    __syncthreads();
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
            if (live_r[mi] && live_c[ni]) acc[mi][ni] = Mfma<T>::op(a[mi], b[ni], acc[mi][ni]);
      }
      if (!more) break;
      g2_store<T, CA>(As[buf ^ 1], ra, kla, tid);
      g2_store<T, CB>(Bs[buf ^ 1], rb, klb, tid);
      __syncthreads();
      buf ^= 1;
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
