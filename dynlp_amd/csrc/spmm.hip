// Row-major SpMM with fused axpy, the masked row operator, and residual reductions.
#include "rowpipe.h"
#include <mutex>
#include <type_traits>

namespace dlp {

// out[o] = add[r] + a[r]*X[r] + b * sum_{j, mask[col_j]} w_j X[col_j]   (o = i or r)
template <typename TW, typename TX, typename TO, int CPL, int U>
__global__ __launch_bounds__(256) void k_row_operator(
    int nrows, const int32_t* __restrict__ rows, int C, int g_shift, int lc_shift,
    const int32_t* __restrict__ indptr, const int32_t* __restrict__ rowlen,
    const int32_t* __restrict__ indices, const TW* __restrict__ w,
    const uint8_t* __restrict__ mask, const TX* __restrict__ X, int64_t ldx,
    const TO* __restrict__ a, TO b, const TO* __restrict__ add, int64_t ldadd,
    TO* __restrict__ out, int64_t ldo, int compact_out) {
  const Lanes L(g_shift, lc_shift);
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> g_shift;
  const bool valid = i < nrows;
  const int r = valid ? (rows ? rows[i] : (int)i) : 0;
  const bool live = valid && (mask == nullptr || mask[r] != 0);
  int beg = 0, end = 0;
  if (live) {
    beg = indptr[r];
    end = rowlen ? beg + rowlen[r] : indptr[r + 1];
  }
  TO acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = 0;
  row_gather<TW, TX, TO, CPL, U>(L, beg, end, indices, w, mask, X, ldx, C, acc);
  group_reduce<TO, CPL>(L, acc);
  if (valid && L.e == 0) {
    const TO ar = a ? a[r] : (TO)0;
    const int64_t o = compact_out ? i : (int64_t)r;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = L.l + k * L.LC;
      if (c < C) {
        TO y = b * acc[k];
        if (a) y += ar * (TO)X[(int64_t)r * ldx + c];
        if (add) y += add[(int64_t)r * ldadd + c];
        out[o * ldo + c] = y;
      }
    }
  }
}

// FP32 operator with 16-byte loads (see row_gather_v4); LC lanes cover ceil(C/4) float4s.
template <int U>
__global__ __launch_bounds__(256) void k_row_operator_v4(
    int nrows, const int32_t* __restrict__ rows, int C, int g_shift, int lc_shift,
    const int32_t* __restrict__ indptr, const int32_t* __restrict__ rowlen,
    const int32_t* __restrict__ indices, const float* __restrict__ w,
    const uint8_t* __restrict__ mask, const float* __restrict__ X, int64_t ldx,
    const float* __restrict__ a, float b, const float* __restrict__ add, int64_t ldadd,
    float* __restrict__ out, int64_t ldo, int compact_out, int xcd_per) {
  const Lanes L(g_shift, lc_shift);
  const int64_t i = (block_order(xcd_per) * blockDim.x + threadIdx.x) >> g_shift;
  const bool valid = i < nrows;
  const int r = valid ? (rows ? rows[i] : (int)i) : 0;
  const bool live = valid && (mask == nullptr || mask[r] != 0);
  int beg = 0, end = 0;
  if (live) {
    beg = indptr[r];
    end = rowlen ? beg + rowlen[r] : indptr[r + 1];
  }
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  row_gather_v4<U>(L, beg, end, indices, w, mask, X, ldx, C, acc);
  for (int off = L.LC; off < L.G; off <<= 1) {
    acc.x += __shfl_xor(acc.x, off, 64);
    acc.y += __shfl_xor(acc.y, off, 64);
    acc.z += __shfl_xor(acc.z, off, 64);
    acc.w += __shfl_xor(acc.w, off, 64);
  }
  if (valid && L.e == 0 && 4 * L.l < C) {
    const float ar = a ? a[r] : 0.f;
    const int64_t o = compact_out ? i : (int64_t)r;
    const float accs[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 4 * L.l + q;
      if (c < C) {
        float y = b * accs[q];
        if (a) y += ar * X[(int64_t)r * ldx + c];
        if (add) y += add[(int64_t)r * ldadd + c];
        out[o * ldo + c] = y;
      }
    }
  }
}

// Packed-lane FP32 operator, one row per wavefront (LanesP).
template <int U>
__global__ __launch_bounds__(256) void k_row_operator_v4p(
    int nrows, const int32_t* __restrict__ rows, int C, const int32_t* __restrict__ indptr,
    const int32_t* __restrict__ rowlen, const int32_t* __restrict__ indices, const float* __restrict__ w,
    const uint8_t* __restrict__ mask, const float* __restrict__ X, int64_t ldx,
    const float* __restrict__ a, float b, const float* __restrict__ add, int64_t ldadd,
    float* __restrict__ out, int64_t ldo, int compact_out, int xcd_per) {
  const LanesP L(C);
  const int64_t i = (block_order(xcd_per) * blockDim.x + threadIdx.x) >> 6;
  const bool valid = i < nrows;
  const int r = valid ? (rows ? rows[i] : (int)i) : 0;
  const bool live = valid && (mask == nullptr || mask[r] != 0);
  int beg = 0, end = 0;
  if (live) {
    beg = indptr[r];
    end = rowlen ? beg + rowlen[r] : indptr[r + 1];
  }
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  row_gather_v4p<U>(L, beg, end, indices, w, mask, X, ldx, acc);
  acc = reduce_v4p(L, acc);
  if (valid && L.e == 0) {
    const float ar = a ? a[r] : 0.f;
    const int64_t o = compact_out ? i : (int64_t)r;
    const float accs[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 4 * L.l + q;
      if (c < C) {
        float y = b * accs[q];
        if (a) y += ar * X[(int64_t)r * ldx + c];
        if (add) y += add[(int64_t)r * ldadd + c];
        out[o * ldo + c] = y;
      }
    }
  }
}

// out[r, 4l..4l+3] = b*acc + ar*xr (columns < C) and dacc += xr .* out.
__device__ __forceinline__ void dot_epilogue(const LanesP& L, int C, int r, const float4& acc, const float4& xr,
                                             float ar, float b, float* __restrict__ out, int64_t ldo,
                                             double (&dacc)[4]) {
  float4 y;
  y.x = b * acc.x + ar * xr.x;
  y.y = b * acc.y + ar * xr.y;
  y.z = b * acc.z + ar * xr.z;
  y.w = b * acc.w + ar * xr.w;
  const int c0 = 4 * L.l;
  float* op = out + (int64_t)r * ldo + c0;
  if (c0 + 3 < C) {
    *reinterpret_cast<float4*>(op) = y;
    dacc[0] += (double)xr.x * (double)y.x;
    dacc[1] += (double)xr.y * (double)y.y;
    dacc[2] += (double)xr.z * (double)y.z;
    dacc[3] += (double)xr.w * (double)y.w;
  } else {
    const float ys[4] = {y.x, y.y, y.z, y.w};
    const float xs[4] = {xr.x, xr.y, xr.z, xr.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (c0 + q < C) {
        op[q] = ys[q];
        dacc[q] += (double)xs[q] * (double)ys[q];
      }
    }
  }
}

// Column sums of dacc: workgroup (LDS) then one FP64 atomic per column per workgroup.
__device__ __forceinline__ void dot_flush(const LanesP& L, int C, bool owner, const double (&dacc)[4], double* sm,
                                          double* __restrict__ dot) {
  __syncthreads();
  if (owner) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 4 * L.l + q;
      if (c < C) atomicAdd(&sm[c], dacc[q]);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&dot[c], sm[c]);
}

// Grid-stride FP32 operator over all rows (one row per wavefront per step) that also
// accumulates dot[c] += X[r,c] * out[r,c] in FP64: the p^T A p column products of PCG, fused
// into the operator's epilogue.  With xcd set (grid a multiple of 8) the blocks an XCD runs
// (blockIdx % 8) walk one contiguous eighth of the rows, so each XCD's L2 serves the
// neighbourhoods of one row range.
template <int U, bool PACK>
__global__ __launch_bounds__(256) void k_row_operator_v4_dot(
    int nrows, int C, const int32_t* __restrict__ indptr, const int32_t* __restrict__ rowlen,
    const int32_t* __restrict__ indices, const float* __restrict__ w, const uint8_t* __restrict__ mask,
    const float* __restrict__ X, int64_t ldx, const float* __restrict__ a, float b, float* __restrict__ out,
    int64_t ldo, double* __restrict__ dot, int xcd, const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sm[256];
  for (int c = threadIdx.x; c < C; c += blockDim.x) sm[c] = 0.0;
  int64_t lo, hi, wx, nwx;
  xcd_rows(nrows, xcd, lo, hi, wx, nwx);
  const LanesP L(C, !PACK);
  const bool owner = L.e == 0 && 4 * L.l < C;
  double dacc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = lo + wx; i < hi; i += nwx) {
    const int r = (int)i;
    const bool live = mask == nullptr || mask[r] != 0;
    int beg = 0, end = 0;
    if (live) {
      beg = indptr[r];
      end = rowlen ? beg + rowlen[r] : indptr[r + 1];
    }
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    row_gather_v4p<U>(L, beg, end, indices, w, mask, X, ldx, acc);
    acc = reduce_v4p(L, acc);
    if (owner) {
      const float ar = a ? a[r] : 0.f;
      const float4 xr = *(reinterpret_cast<const float4*>(X + (int64_t)r * ldx) + L.l);
      dot_epilogue(L, C, r, acc, xr, ar, b, out, ldo, dacc);
    }
  }
  dot_flush(L, C, owner, dacc, sm, dot);
}

// The global solve's operator AP = diag*P - W_UU P with the fused p^T A p, on the pipelined
// row gathers of rowpipe.h, with P in FP32 (W = 4) or BF16 (W = 8; AP stays FP32).  Neighbour
// rows are not masked: every row outside the solved set holds P = 0 during the solve (its
// residual is parked at 0 by the refresh, so P = R/diag + beta P and AP = diag*P keep it 0),
// and its term vanishes.
template <Word K>
struct PcgEpilogue {
  static constexpr int W = WordT<K>::W;
  const LanesP& L;
  int C;
  bool owner;
  const char* __restrict__ X;  // P rows of row_bytes bytes
  unsigned row_bytes;
  const float* __restrict__ a;
  float b;
  float* __restrict__ out;
  int64_t ldo;
  double* sm;  // workgroup column sums of P .* AP (LDS)
  // own P row's lane word, kept raw until post: converting it here would make the row's
  // gathers wait for this load
  typename WordT<K>::raw xw;
  float ar;
  // this lane's partial sums of P .* AP over the rows it finishes (FP32: a few hundred terms
  // of FP16-held directions), added to the workgroup's FP64 column sums once, at the end
  float dacc[W] = {};
  __device__ __forceinline__ void pre(int, int r) {
    if (owner) xw = load_word<K>(X + (int64_t)r * row_bytes + WordT<K>::bytes * L.l);
    else xw = typename WordT<K>::raw{};
    ar = a[r];
  }
  // out[r, W*l ..] = b*acc + ar*P (columns < C), dacc += P .* out; rows outside the mask get
  // AP = diag*P (no neighbour sum), as the plain operator writes them
  __device__ __forceinline__ void post(int, int r, bool, const float (&acc)[W]) {
    if (!owner) return;
    float xr[W];
    unpack<K>(xw, xr);
    const int c0 = W * L.l;
    float y[W];
#pragma unroll
    for (int q = 0; q < W; ++q) y[q] = b * acc[q] + ar * xr[q];
    float* op = out + (int64_t)r * ldo + c0;
    if (c0 + W - 1 < C) {
#pragma unroll
      for (int k = 0; k < W / 4; ++k)
        reinterpret_cast<float4*>(op)[k] = make_float4(y[4 * k], y[4 * k + 1], y[4 * k + 2], y[4 * k + 3]);
#pragma unroll
      for (int q = 0; q < W; ++q) dacc[q] = __builtin_fmaf(xr[q], y[q], dacc[q]);
    } else {
#pragma unroll
      for (int q = 0; q < W; ++q) {
        if (c0 + q < C) {
          op[q] = y[q];
          dacc[q] = __builtin_fmaf(xr[q], y[q], dacc[q]);
        }
      }
    }
  }
  // the lane's sums into the workgroup's FP64 column sums (LDS)
  __device__ __forceinline__ void flush() {
    if (!owner) return;
#pragma unroll
    for (int q = 0; q < W; ++q)
      if (W * L.l + q < C) atomicAdd(&sm[W * L.l + q], (double)dacc[q]);
  }
};

template <int S, Word K, int WPE = 7>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_pcg_operator(
    int nrows, int C, const int32_t* __restrict__ indptr, const int32_t* __restrict__ rowlen,
    const int32_t* __restrict__ indices, const float* __restrict__ w, const uint8_t* __restrict__ rowmask,
    const void* __restrict__ X, unsigned row_bytes, int xbytes, const float* __restrict__ a, float b,
    float* __restrict__ out, int64_t ldo, double* __restrict__ dot, int xcd, const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sm[264];
  for (int c = threadIdx.x; c < 264; c += blockDim.x) sm[c] = 0.0;
  __syncthreads();
  constexpr int W = WordT<K>::W;
  const RowPipe G(nrows, C, W, WordT<K>::bytes, xcd, rowmask, indptr, rowlen, indices, w, X, row_bytes, xbytes);
  PcgEpilogue<K> E{G.L, C, G.L.e == 0 && W * G.L.l < C, (const char*)X, row_bytes, a, b, out, ldo, sm};
  run_rows<S, K, float>(G, E);
  E.flush();
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&dot[c], sm[c]);
}

Tuning& tuning() {
  static Tuning t;
  return t;
}

// Blocks of `kernel` (256 threads) that fit on the current device at once, cached per
// (kernel, device).
int resident_grid(const void* kernel) {
  struct Entry {
    const void* k;
    int dev, blocks;
  };
  static Entry cache[64];
  static int used = 0;
  static std::mutex mu;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  std::lock_guard<std::mutex> lock(mu);
  for (int i = 0; i < used; ++i)
    if (cache[i].k == kernel && cache[i].dev == dev) return cache[i].blocks;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu < 1)
    per_cu = 4;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  const int g = per_cu * cus;
  if (used < 64) cache[used++] = Entry{kernel, dev, g};
  return g;
}

// out = a .* X + b * W_mask X over all rows, and dot[c] += sum_r X[r,c] * out[r,c] (dot not
// cleared here).  FP32, 16-byte rows required; returns DLP_ERR_ARG when the layout does not fit.
int row_operator_dot(int n, int C, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices,
                     const float* w, const uint8_t* mask, const float* X, int64_t ldx, const float* a, float b,
                     float* out, int64_t ldo, double* dot, hipStream_t st, const int32_t* skip) {
  DLP_CHECK_ARG((ldx % 4) == 0 && (ldo % 4) == 0 && (((uintptr_t)X) & 15) == 0 && (((uintptr_t)out) & 15) == 0 &&
                    C <= 256,
                "row_operator_dot: layout");
  if (n <= 0) return DLP_OK;
  const Tuning& T = tuning();
  const bool pack = T.pack != 0;
  const bool u8 = T.unroll == 8;   // auto (0): 4, the measured best of the grid-stride form
  const void* kfn = pack ? (u8 ? reinterpret_cast<const void*>(&k_row_operator_v4_dot<8, true>)
                               : reinterpret_cast<const void*>(&k_row_operator_v4_dot<4, true>))
                         : (u8 ? reinterpret_cast<const void*>(&k_row_operator_v4_dot<8, false>)
                               : reinterpret_cast<const void*>(&k_row_operator_v4_dot<4, false>));
  long long grid = T.dot_grid > 0 ? T.dot_grid : resident_grid(kfn);
  const long long need = ((long long)n + 3) / 4;
  if (grid > need) grid = need;
  const int xcd = T.xcd != 0;
  if (xcd) grid = grid < 8 ? 8 : grid & ~7LL;
#define DLP_DOT(UV, PV)                                                                                      \
  hipLaunchKernelGGL((k_row_operator_v4_dot<UV, PV>), dim3((unsigned)grid), dim3(kBlock), 0, st, n, C, indptr, \
                     rowlen, indices, w, mask, X, ldx, a, b, out, ldo, dot, xcd, skip)
  if (pack) {
    if (u8) DLP_DOT(8, true);
    else DLP_DOT(4, true);
  } else {
    if (u8) DLP_DOT(8, false);
    else DLP_DOT(4, false);
  }
#undef DLP_DOT
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// The global solve's operator (k_pcg_operator): rowmask = the solved set, rowlen required,
// 16-byte rows, C <= 256.  X is FP32 (ldx floats per row) or, with X16, BF16 (x16_kind 1) or
// FP16 (x16_kind 2), ldh values per row, a multiple of 8, or FP8 E4M3 (x16_kind 3), rows of
// 2*ldh bytes (ldh even: 4-byte lane words of 4 columns).  The FP32 form falls back to row_operator_dot when tuning disables the
// pipelined kernel (key 6 = 0) or the block exceeds its 32-bit offsets.
int pcg_operator(int n, int C, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices, const float* w,
                 const uint8_t* rowmask, const float* X, int64_t ldx, const uint16_t* X16, int64_t ldh, int x16_kind,
                 const float* a, float b, float* out, int64_t ldo, double* dot, hipStream_t st, const int32_t* skip) {
  const Tuning& T = tuning();
  const bool bf = X16 != nullptr;
  const bool h16 = bf && x16_kind == 2;
  const bool f8 = bf && x16_kind == 3;
  DLP_CHECK_ARG(!bf || (x16_kind >= 1 && x16_kind <= 3), "pcg_operator: low-precision kind %d", x16_kind);
  const long long xbytes = bf ? (long long)n * ldh * 2 : (long long)n * ldx * 4;
  const bool fits = rowlen && rowmask && a && n < (1 << 24) && xbytes < (1LL << 31);
  if (!bf && (!T.pipe || !fits))
    return row_operator_dot(n, C, indptr, rowlen, indices, w, rowmask, X, ldx, a, b, out, ldo, dot, st, skip);
  DLP_CHECK_ARG(fits && (ldo % 4) == 0 && (((uintptr_t)out) & 15) == 0 && C <= 256 &&
                    (f8 ? ((ldh % 2) == 0 && ldh * 2 >= 4 * ((C + 3) / 4) && (((uintptr_t)X16) & 15) == 0)
                     : bf ? ((ldh % 8) == 0 && (((uintptr_t)X16) & 15) == 0)
                          : ((ldx % 4) == 0 && (((uintptr_t)X) & 15) == 0)),
                "pcg_operator: layout");
  if (n <= 0) return DLP_OK;
  const int S = T.pipe == 8 ? 8 : T.pipe == 6 ? 6 : 4;
  const bool w8 = !h16 && T.bf16_word == 8;
#define DLP_PIPE_FN(SV, KV) reinterpret_cast<const void*>(&k_pcg_operator<SV, KV>)
  // FP8 words take one register per gather: 8 in flight per lane at 6 waves per SIMD measured
  // 0.396 against 0.409 ms at 4 and 7 on wide rows (SBM-1M, 65 columns: 3 edge lanes); on narrow
  // rows (ER-1M, 17 columns: 12 edge lanes) 8 gathers overshoot most rows and depth 4 is
  // faster (1.36M against 1.23M edge-updates/s).  16-bit words spill at depth 6 or 8
  const int f8_lanes_e = 64 / ((C + 3) / 4);
  const int S8 = T.f8_pipe ? T.f8_pipe : (f8_lanes_e <= 4 ? 8 : 4);
  const int f8w = T.f8_waves;
  const void* kfn = f8 ? (S8 == 8   ? (f8w == 5 ? reinterpret_cast<const void*>(&k_pcg_operator<8, Word::F8x4, 5>)
                                               : reinterpret_cast<const void*>(&k_pcg_operator<8, Word::F8x4, 6>))
                          : S8 == 6 ? (f8w == 5 ? reinterpret_cast<const void*>(&k_pcg_operator<6, Word::F8x4, 5>)
                                               : reinterpret_cast<const void*>(&k_pcg_operator<6, Word::F8x4, 6>))
                                   : DLP_PIPE_FN(4, Word::F8x4))
                  : h16 ? (S == 8 ? DLP_PIPE_FN(8, Word::F16x4)
                           : S == 6 ? DLP_PIPE_FN(6, Word::F16x4) : DLP_PIPE_FN(4, Word::F16x4))
                  : bf ? (w8 ? DLP_PIPE_FN(4, Word::BF16x8)
                             : S == 8 ? DLP_PIPE_FN(8, Word::BF16x4)
                             : S == 6 ? DLP_PIPE_FN(6, Word::BF16x4) : DLP_PIPE_FN(4, Word::BF16x4))
                       : S == 8 ? DLP_PIPE_FN(8, Word::F32x4)
                       : S == 6 ? DLP_PIPE_FN(6, Word::F32x4) : DLP_PIPE_FN(4, Word::F32x4);
#undef DLP_PIPE_FN
  long long grid = T.dot_grid > 0 ? T.dot_grid : resident_grid(kfn);
  const long long need = ((long long)n + 3) / 4;
  if (grid > need) grid = need;
  const int xcd = T.xcd != 0;
  if (xcd) grid = grid < 8 ? 8 : grid & ~7LL;
  const void* Xp = bf ? (const void*)X16 : (const void*)X;
  const unsigned rb = bf ? (unsigned)(ldh * 2) : (unsigned)(ldx * 4);
#define DLP_PIPE3(SV, KV, WV)                                                                                 \
  hipLaunchKernelGGL((k_pcg_operator<SV, KV, WV>), dim3((unsigned)grid), dim3(kBlock), 0, st, n, C, indptr,      \
                     rowlen, indices, w, rowmask, Xp, rb, (int)xbytes, a, b, out, ldo, dot, xcd, skip)
#define DLP_PIPE(SV, KV)                                                                                      \
  hipLaunchKernelGGL((k_pcg_operator<SV, KV>), dim3((unsigned)grid), dim3(kBlock), 0, st, n, C, indptr, rowlen, \
                     indices, w, rowmask, Xp, rb, (int)xbytes, a, b, out, ldo, dot, xcd, skip)
  if (f8 && S8 == 8 && f8w == 5) DLP_PIPE3(8, Word::F8x4, 5);
  else if (f8 && S8 == 8) DLP_PIPE3(8, Word::F8x4, 6);
  else if (f8 && S8 == 6 && f8w == 5) DLP_PIPE3(6, Word::F8x4, 5);
  else if (f8 && S8 == 6) DLP_PIPE3(6, Word::F8x4, 6);
  else if (f8) DLP_PIPE(4, Word::F8x4);
  else if (h16 && S == 8) DLP_PIPE(8, Word::F16x4);
  else if (h16 && S == 6) DLP_PIPE(6, Word::F16x4);
  else if (h16) DLP_PIPE(4, Word::F16x4);
  else if (bf && w8) DLP_PIPE(4, Word::BF16x8);
  else if (bf && S == 8) DLP_PIPE(8, Word::BF16x4);
  else if (bf && S == 6) DLP_PIPE(6, Word::BF16x4);
  else if (bf) DLP_PIPE(4, Word::BF16x4);
  else if (S == 8) DLP_PIPE(8, Word::F32x4);
  else if (S == 6) DLP_PIPE(6, Word::F32x4);
  else DLP_PIPE(4, Word::F32x4);
#undef DLP_PIPE
#undef DLP_PIPE3
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

template <typename TW, typename TX, typename TO>
static int launch_row_operator(int nrows, const int32_t* rows, int C, const int32_t* indptr,
                               const int32_t* rowlen, const int32_t* indices, const void* w,
                               const uint8_t* mask, const void* X, int64_t ldx, const void* a,
                               double b, const void* add, int64_t ldadd, void* out, int64_t ldo,
                               int compact_out, double avg_deg, hipStream_t st) {
  const bool v4 = std::is_same<TW, float>::value && std::is_same<TX, float>::value &&
                  std::is_same<TO, float>::value && tuning().vec4 && (ldx % 4) == 0 &&
                  (((uintptr_t)X) & 15) == 0 && C <= 256;
  if (v4) {
    // one row per wavefront; column lanes either packed (ceil(C/4) lanes, 64/that edges) or
    // the next power of two >= ceil(C/4) with edge lanes filling the wavefront
    int lc = ceil_log2((C + 3) / 4);
    if (lc > 6) lc = 6;
    const long long threads = (long long)nrows << 6;
    unsigned g = grid_for(threads);
    int xcd_per = 0;
    if (tuning().xcd) {
      g = (g + 7) & ~7u;
      xcd_per = (int)(g >> 3);
    }
    dim3 grid(g), block(kBlock);
    const float* wf = (const float*)w;
    const float* Xf = (const float*)X;
    const float* af = (const float*)a;
    const float* addf = (const float*)add;
    float* of = (float*)out;
    if (tuning().pack && tuning().unroll == 8)
      hipLaunchKernelGGL((k_row_operator_v4p<8>), grid, block, 0, st, nrows, rows, C, indptr, rowlen, indices, wf,
                         mask, Xf, ldx, af, (float)b, addf, ldadd, of, ldo, compact_out, xcd_per);
    else if (tuning().pack)
      hipLaunchKernelGGL((k_row_operator_v4p<4>), grid, block, 0, st, nrows, rows, C, indptr, rowlen, indices, wf,
                         mask, Xf, ldx, af, (float)b, addf, ldadd, of, ldo, compact_out, xcd_per);
    else if (tuning().unroll != 4)
      hipLaunchKernelGGL((k_row_operator_v4<8>), grid, block, 0, st, nrows, rows, C, 6, lc, indptr, rowlen,
                         indices, wf, mask, Xf, ldx, af, (float)b, addf, ldadd, of, ldo, compact_out, xcd_per);
    else
      hipLaunchKernelGGL((k_row_operator_v4<4>), grid, block, 0, st, nrows, rows, C, 6, lc, indptr, rowlen,
                         indices, wf, mask, Xf, ldx, af, (float)b, addf, ldadd, of, ldo, compact_out, xcd_per);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  RowMap m = choose_map(C, avg_deg);
  const long long threads = (long long)nrows << m.g_shift;
  dim3 grid(grid_for(threads)), block(kBlock);
  const bool u8 = tuning().unroll != 4;
#define DLP_ROWOP(CPLV)                                                                        \
  if (u8) hipLaunchKernelGGL((k_row_operator<TW, TX, TO, CPLV, 8>), grid, block, 0, st, nrows, rows, C, \
                     m.g_shift, m.lc_shift, indptr, rowlen, indices, (const TW*)w, mask,       \
                     (const TX*)X, ldx, (const TO*)a, (TO)b, (const TO*)add, ldadd, (TO*)out,  \
                     ldo, compact_out);                                                        \
  else hipLaunchKernelGGL((k_row_operator<TW, TX, TO, CPLV, 4>), grid, block, 0, st, nrows, rows, C,   \
                     m.g_shift, m.lc_shift, indptr, rowlen, indices, (const TW*)w, mask,       \
                     (const TX*)X, ldx, (const TO*)a, (TO)b, (const TO*)add, ldadd, (TO*)out,  \
                     ldo, compact_out)
  switch (m.cpl) {
    case 1: DLP_ROWOP(1); break;
    case 2: DLP_ROWOP(2); break;
    case 3: DLP_ROWOP(3); break;
    default: DLP_ROWOP(4); break;
  }
#undef DLP_ROWOP
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// rho[c] = max |R[r,c]| / diag[r] over listed rows; one block reduces in LDS, then one
// atomic max per column per block.
template <typename TR>
__global__ __launch_bounds__(256) void k_rho_cols(int nrows, const int32_t* __restrict__ rows,
                                                  int C, const TR* __restrict__ R, int64_t ldr,
                                                  const double* __restrict__ diag,
                                                  double* __restrict__ rho) {
  extern __shared__ double sm[];  // C doubles
  for (int c = threadIdx.x; c < C; c += blockDim.x) sm[c] = 0.0;
  __syncthreads();
  // a block iteration covers rpi whole rows: thread t takes row t / C, column t % C
  const int rpi = C <= (int)blockDim.x ? (int)blockDim.x / C : 1;
  const int sub = C <= (int)blockDim.x ? (int)threadIdx.x / C : 0;
  const int c0 = C <= (int)blockDim.x ? (int)threadIdx.x - sub * C : (int)threadIdx.x;
  const int cstep = C <= (int)blockDim.x ? C : (int)blockDim.x;
  if (sub < rpi) {
    double local = 0.0;
    int lastc = c0;
    for (int64_t i = (int64_t)blockIdx.x * rpi + sub; i < nrows; i += (int64_t)gridDim.x * rpi) {
      const int r = rows ? rows[i] : (int)i;
      const double inv = 1.0 / diag[r];
      for (int c = c0; c < C; c += cstep) {
        const double v = nan_to_inf(fabs((double)R[(int64_t)r * ldr + c]) * inv);
        if (c == c0) local = fmax(local, v);
        else atomic_max_nonneg(&sm[c], v);
      }
    }
    (void)lastc;
    atomic_max_nonneg(&sm[c0], local);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) global_max_nonneg(&rho[c], sm[c]);
}

// Rows with any |R[r,c]| > eps[c]*diag[r] appended to out.  Group of LC lanes per row.
template <typename TR>
__global__ __launch_bounds__(256) void k_violating(int nrows, const int32_t* __restrict__ rows,
                                                   int C, int lc_shift, const TR* __restrict__ R,
                                                   int64_t ldr, const double* __restrict__ diag,
                                                   const double* __restrict__ eps,
                                                   int32_t* __restrict__ out,
                                                   int32_t* __restrict__ count) {
  const int LC = 1 << lc_shift;
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> lc_shift;
  const int l = threadIdx.x & (LC - 1);
  const bool valid = i < nrows;
  const int r = valid ? (rows ? rows[i] : (int)i) : 0;
  bool viol = false;
  if (valid) {
    const double d = diag[r];
    for (int c = l; c < C; c += LC) viol |= !(fabs((double)R[(int64_t)r * ldr + c]) <= eps[c] * d);
  }
  const unsigned long long bal = __ballot(viol);
  const int lane = threadIdx.x & 63;
  const unsigned long long gmask = (LC == 64 ? ~0ull : ((1ull << LC) - 1)) << (lane & ~(LC - 1));
  if (valid && l == 0 && (bal & gmask)) out[atomicAdd(count, 1)] = r;
}

}  // namespace dlp

using namespace dlp;

extern "C" int dlp_row_operator(int nrows, const int32_t* rows, int C, const int32_t* indptr,
                                const int32_t* rowlen, const int32_t* indices, const void* w,
                                int w_dtype, const uint8_t* mask, const void* X, int64_t ldx,
                                int x_dtype, const void* a, double b, const void* add,
                                int64_t ldadd, void* out, int64_t ldo, int out_dtype,
                                int compact_out, double avg_deg, void* stream) {
  DLP_CHECK_ARG(nrows >= 0 && C >= 1 && C <= 256, "row_operator: bad nrows=%d or C=%d", nrows, C);
  if (nrows == 0) return DLP_OK;
  DLP_CHECK_ARG(indptr && indices && w && X && out, "row_operator: null pointer");
  DLP_CHECK_ARG(ldx >= C && ldo >= C && (!add || ldadd >= C), "row_operator: leading dim < C");
  hipStream_t st = (hipStream_t)stream;
#define DLP_RO(TW, TX, TO)                                                                   \
  return launch_row_operator<TW, TX, TO>(nrows, rows, C, indptr, rowlen, indices, w, mask, X, \
                                         ldx, a, b, add, ldadd, out, ldo, compact_out,        \
                                         avg_deg, st)
  if (w_dtype == DLP_F32 && x_dtype == DLP_F32 && out_dtype == DLP_F32) DLP_RO(float, float, float);
  if (w_dtype == DLP_F64 && x_dtype == DLP_F64 && out_dtype == DLP_F64) DLP_RO(double, double, double);
  if (w_dtype == DLP_F32 && x_dtype == DLP_F64 && out_dtype == DLP_F64) DLP_RO(float, double, double);
  if (w_dtype == DLP_F32 && x_dtype == DLP_F32 && out_dtype == DLP_F64) DLP_RO(float, float, double);
#undef DLP_RO
  DLP_CHECK_ARG(false, "row_operator: unsupported dtype combination w=%d x=%d out=%d", w_dtype,
                x_dtype, out_dtype);
  return DLP_ERR_ARG;
}

extern "C" int dlp_row_operator_dot(int n, int C, const int32_t* indptr, const int32_t* rowlen,
                                    const int32_t* indices, const float* w, const uint8_t* mask,
                                    const float* X, int64_t ldx, const float* a, float b, float* out,
                                    int64_t ldo, double* dot, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= 256, "row_operator_dot: bad n=%d or C=%d", n, C);
  if (n == 0) return DLP_OK;
  DLP_CHECK_ARG(indptr && indices && w && X && out && dot, "row_operator_dot: null pointer");
  DLP_CHECK_ARG(ldx >= C && ldo >= C, "row_operator_dot: leading dim < C");
  return row_operator_dot(n, C, indptr, rowlen, indices, w, mask, X, ldx, a, b, out, ldo, dot,
                          (hipStream_t)stream, nullptr);
}

// Y[r] = a[r] * X[r] + b * sum_j w_j X[col_j] (a null: no first term) over every row of a plain
// CSR, FP32, on the pipelined packed-lane gathers; X rows of ldx floats at 4-byte alignment
// (16-byte words when ldx % 4 == 0 and X is 16-byte aligned, else four dword loads per lane).
template <Word K>
struct AxpyEpilogue {
  const LanesP& L;
  int C;
  bool owner;
  const float* __restrict__ X;
  int64_t ldx;
  const float* __restrict__ a;
  float b;
  float* __restrict__ out;
  int64_t ldo;
  bool vec_out;
  float xr[4];
  float ar;
  __device__ __forceinline__ void pre(int, int r) {
    ar = 0.f;
    if (a && owner) {
      load4(X + (int64_t)r * ldx, 4 * L.l, C, xr);
      ar = a[r];
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) xr[q] = 0.f;
    }
  }
  __device__ __forceinline__ void post(int, int r, bool, const float (&acc)[4]) {
    if (!owner) return;
    float y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q] = b * acc[q] + ar * xr[q];
    store4(out + (int64_t)r * ldo, 4 * L.l, C, vec_out, y);
  }
};

template <int S, Word K>
__global__ __launch_bounds__(256) void k_spmm_rows(int nrows, int C, const int32_t* __restrict__ indptr,
                                                   const int32_t* __restrict__ indices, const float* __restrict__ w,
                                                   const float* __restrict__ X, int64_t ldx, int xbytes,
                                                   const float* __restrict__ a, float b, float* __restrict__ out,
                                                   int64_t ldo, int vec_out, int xcd) {
  const RowPipe G(nrows, C, 4, 16, xcd, nullptr, indptr, nullptr, indices, w, X, (unsigned)(ldx * 4), xbytes);
  AxpyEpilogue<K> E{G.L, C, G.L.e == 0 && 4 * G.L.l < C, X, ldx, a, b, out, ldo, vec_out != 0};
  run_rows<S, K, float>(G, E);
}

// One pass of an AMG level on the pipelined row gathers (A = diag - W, FP32 rows of ld floats):
// mode 0 the residual out = R - A X, mode 1 a weighted-Jacobi sweep out = X + omega (R - A X) /
// diag (the smoother's residual and correction in one pass).
template <Word K>
struct LevelEpilogue {
  const LanesP& L;
  int C;
  bool owner;
  const float* __restrict__ X;
  const float* __restrict__ Rr;
  int64_t ld;
  const float* __restrict__ d;
  float omega;
  int mode;
  float* __restrict__ out;
  float xr[4], rr[4];
  float dr;
  __device__ __forceinline__ void pre(int, int r) {
    if (owner) {
      load4(X + (int64_t)r * ld, 4 * L.l, C, xr);
      load4(Rr + (int64_t)r * ld, 4 * L.l, C, rr);
    }
    dr = d[r];
  }
  __device__ __forceinline__ void post(int, int r, bool, const float (&acc)[4]) {
    if (!owner) return;
    float y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float res = rr[q] - dr * xr[q] + acc[q];
      y[q] = mode == 0 ? res : xr[q] + omega * res / dr;
    }
    store4(out + (int64_t)r * ld, 4 * L.l, C, true, y);
  }
};

template <int S>
__global__ __launch_bounds__(256) void k_amg_level_rows(int nrows, int C, const int32_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ indices,
                                                        const float* __restrict__ w, const float* __restrict__ X,
                                                        int64_t ld, int xbytes, const float* __restrict__ R,
                                                        const float* __restrict__ diag, float omega, int mode,
                                                        float* __restrict__ out, int xcd) {
  const RowPipe G(nrows, C, 4, 16, xcd, nullptr, indptr, nullptr, indices, w, X, (unsigned)(ld * 4), xbytes);
  LevelEpilogue<Word::F32x4> E{G.L, C, G.L.e == 0 && 4 * G.L.l < C, X, R, ld, diag, omega, mode, out};
  run_rows<S, Word::F32x4, float>(G, E);
}

// 0 when the packed form does not apply (the caller falls back to the generic row kernel).
static int spmm_rows_f32(int n, int C, const int32_t* indptr, const int32_t* indices, const float* w,
                         const float* X, int64_t ldx, const float* a, float b, float* Y, int64_t ldy,
                         hipStream_t st, bool& done) {
  done = false;
  const long long xbytes = (long long)n * ldx * 4;
  if (!tuning().pipe || C > 256 || ldx < C || n >= (1 << 24) || xbytes >= (1LL << 31) || ldx * 4 >= (1 << 24) ||
      (((uintptr_t)X) & 3) != 0 || (((uintptr_t)Y) & 3) != 0)
    return DLP_OK;
  const bool v16 = (ldx % 4) == 0 && (((uintptr_t)X) & 15) == 0;
  const int vec_out = (ldy % 4) == 0 && (((uintptr_t)Y) & 15) == 0;
  const void* kfn = v16 ? reinterpret_cast<const void*>(&k_spmm_rows<4, Word::F32x4>)
                        : reinterpret_cast<const void*>(&k_spmm_rows<4, Word::F32x4u>);
  long long grid = resident_grid(kfn);
  const long long need = ((long long)n + 3) / 4;
  if (grid > need) grid = need;
  const int xcd = tuning().xcd != 0;
  if (xcd) grid = grid < 8 ? 8 : grid & ~7LL;
  if (v16)
    hipLaunchKernelGGL((k_spmm_rows<4, Word::F32x4>), dim3((unsigned)grid), dim3(kBlock), 0, st, n, C, indptr,
                       indices, w, X, ldx, (int)xbytes, a, b, Y, ldy, vec_out, xcd);
  else
    hipLaunchKernelGGL((k_spmm_rows<4, Word::F32x4u>), dim3((unsigned)grid), dim3(kBlock), 0, st, n, C, indptr,
                       indices, w, X, ldx, (int)xbytes, a, b, Y, ldy, vec_out, xcd);
  DLP_CHECK_LAUNCH();
  done = true;
  return DLP_OK;
}

extern "C" int dlp_spmm_axpy(int n, int C, const int32_t* indptr, const int32_t* indices,
                             const void* w, const void* X, int64_t ldx, const void* a, double b,
                             void* Y, int64_t ldy, int dtype, double avg_deg, void* stream) {
  DLP_CHECK_ARG(dtype == DLP_F32 || dtype == DLP_F64, "spmm_axpy: bad dtype %d", dtype);
  if (dtype == DLP_F32 && n > 0 && C >= 1 && indptr && indices && w && X && Y && ldx >= C && ldy >= C) {
    bool done = false;
    const int rc = spmm_rows_f32(n, C, indptr, indices, (const float*)w, (const float*)X, ldx, (const float*)a,
                                 (float)b, (float*)Y, ldy, (hipStream_t)stream, done);
    if (rc || done) return rc;
  }
  return dlp_row_operator(n, nullptr, C, indptr, nullptr, indices, w, dtype, nullptr, X, ldx,
                          dtype, a, b, nullptr, 0, Y, ldy, dtype, 0, avg_deg, stream);
}

// An AMG level pass (amg.py:114-134's smoother and residual, A = diag - W): mode 0 out = R - A X,
// mode 1 out = X + omega (R - A X) / diag; FP32 n x C blocks of leading dimension ld (a multiple
// of 4, 16-byte aligned, n * ld * 4 < 2^31), plain CSR W.  DLP_ERR_ARG when the layout does not
// fit (the caller then uses dlp_row_operator and dlp_colop).
extern "C" int dlp_amg_level(int n, int C, const int32_t* indptr, const int32_t* indices, const float* w,
                             const float* diag, const float* X, const float* R, int64_t ld, double omega, int mode,
                             float* out, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= 256 && ld >= C && (mode == 0 || mode == 1), "amg_level: bad arguments");
  if (n == 0) return DLP_OK;
  const long long xbytes = (long long)n * ld * 4;
  DLP_CHECK_ARG(indptr && indices && w && diag && X && R && out && (ld % 4) == 0 && (((uintptr_t)X) & 15) == 0 &&
                    (((uintptr_t)R) & 15) == 0 && (((uintptr_t)out) & 15) == 0 && n < (1 << 24) &&
                    xbytes < (1LL << 31) && X != out,
                "amg_level: layout (ld %% 4, 16-byte alignment, n * ld * 4 < 2^31, out apart from X)");
  hipStream_t st = (hipStream_t)stream;
  const void* kfn = reinterpret_cast<const void*>(&k_amg_level_rows<4>);
  long long grid = resident_grid(kfn);
  const long long need = ((long long)n + 3) / 4;
  if (grid > need) grid = need;
  const int xcd = tuning().xcd != 0;
  if (xcd) grid = grid < 8 ? 8 : grid & ~7LL;
  hipLaunchKernelGGL(k_amg_level_rows<4>, dim3((unsigned)grid), dim3(kBlock), 0, st, n, C, indptr, indices, w, X, ld,
                     (int)xbytes, R, diag, (float)omega, mode, out, xcd);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// rho[c] = max_r |R[r,c]| / diag[r] over all rows of a 16-byte-aligned FP32 block (flat quads).
__global__ __launch_bounds__(256) void k_rho_cols4(int n, int Q4, int C, const float4* __restrict__ R,
                                                   const double* __restrict__ diag, double* __restrict__ rho) {
  __shared__ double sm[260];
  for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) sm[c] = 0.0;
  __syncthreads();
  const Quad Qd = quad_of(Q4, C);
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  if (Qd.active && Qd.r0 < n) {
    const int64_t nq = (int64_t)n * Q4;
    int64_t r = Qd.r0;
    for (int64_t i = Qd.i0; i < nq; i += Qd.step, r += Qd.rstep) {
      const double inv = 1.0 / diag[r];
      const float4 v = R[i];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j < Qd.cnt) m[j] = fmax(m[j], nan_to_inf(fabs((double)quad_comp(v, j)) * inv));
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < Qd.cnt) atomic_max_nonneg(&sm[4 * Qd.q4 + j], m[j]);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) global_max_nonneg(rho + c, sm[c]);
}

extern "C" int dlp_rho_cols(int nrows, const int32_t* rows, int C, const void* R, int64_t ldr,
                            int r_dtype, const double* diag, double* rho, void* stream) {
  DLP_CHECK_ARG(nrows >= 0 && C >= 1 && C <= 4096, "rho_cols: bad sizes");
  DLP_CHECK_ARG(R && diag && rho, "rho_cols: null pointer");
  hipStream_t st = (hipStream_t)stream;
  DLP_CHECK_HIP(hipMemsetAsync(rho, 0, sizeof(double) * C, st));
  if (nrows == 0) return DLP_OK;
  // 16-byte quads when most of each row is read (one column of a wide block takes a thread
  // per row instead: the quad layout would leave all but one thread of each row group idle)
  if (!rows && r_dtype == DLP_F32 && (ldr % 4) == 0 && ldr <= 256 && (((uintptr_t)R) & 15) == 0 &&
      2 * C > ldr && tuning().vec4) {
    const int Q4 = (int)(ldr / 4);
    const long long used = Q4 * (kBlock / Q4);
    long long g = ((long long)nrows * Q4 + used - 1) / used;
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_rho_cols4, dim3((unsigned)g), dim3(kBlock), 0, st, nrows, Q4, C, (const float4*)R, diag,
                       rho);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  const long long total = (long long)nrows * C;
  unsigned grid = grid_for(total);
  if (grid > 2048) grid = 2048;
  const size_t shm = sizeof(double) * C;
  if (r_dtype == DLP_F32)
    hipLaunchKernelGGL(k_rho_cols<float>, dim3(grid), dim3(kBlock), shm, st, nrows, rows, C,
                       (const float*)R, ldr, diag, rho);
  else
    hipLaunchKernelGGL(k_rho_cols<double>, dim3(grid), dim3(kBlock), shm, st, nrows, rows, C,
                       (const double*)R, ldr, diag, rho);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_violating_rows(int nrows, const int32_t* rows, int C, const void* R,
                                  int64_t ldr, int r_dtype, const double* diag,
                                  const double* eps, int32_t* out, int32_t* count, void* stream) {
  DLP_CHECK_ARG(nrows >= 0 && C >= 1, "violating_rows: bad sizes");
  DLP_CHECK_ARG(R && diag && eps && out && count, "violating_rows: null pointer");
  hipStream_t st = (hipStream_t)stream;
  DLP_CHECK_HIP(hipMemsetAsync(count, 0, sizeof(int32_t), st));
  if (nrows == 0) return DLP_OK;
  int ls = C <= 1 ? 0 : ceil_log2(C);
  if (ls > 6) ls = 6;
  const long long threads = (long long)nrows << ls;
  if (r_dtype == DLP_F32)
    hipLaunchKernelGGL(k_violating<float>, dim3(grid_for(threads)), dim3(kBlock), 0, st, nrows,
                       rows, C, ls, (const float*)R, ldr, diag, eps, out, count);
  else
    hipLaunchKernelGGL(k_violating<double>, dim3(grid_for(threads)), dim3(kBlock), 0, st, nrows,
                       rows, C, ls, (const double*)R, ldr, diag, eps, out, count);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_tune_set(int key, int value) {
  // key 0: unroll factor of the gather loops (0 = per-kernel default, 4 or 8); key 1: 16-byte
  // FP32 loads on/off;
  // key 2: packed lane groups; key 3: XCD-aware row order; key 4: blocks of the fused-dot
  // operator (0 = resident blocks); key 5: fuse p^T A p into the global solve's operator;
  // key 6: gathers in flight per lane in the pipelined global-solve operator (4, 6 or 8; 0 =
  // the plain fused operator); key 7: BF16 direction values per lane (4: 8-byte words, 8: 16);
  // key 8: workgroups of the flat PCG vector passes (0 = 2048); key 9: non-temporal FP32 streams
  // in the PCG update and direction passes (1 default); key 10: deferred solution updates; key
  // 11: waves per SIMD the FP8 operator is compiled for at depth 6 or 8 (5 or 6); key 12:
  // gathers per lane in flight in the FP8 operator (4, 6 or 8; 0, the default: by row width)
  if (key == 0) {
    DLP_CHECK_ARG(value == 0 || value == 4 || value == 8, "tune: unroll must be 0, 4 or 8");
    tuning().unroll = value;
  } else if (key == 1) {
    tuning().vec4 = value ? 1 : 0;
  } else if (key == 2) {
    tuning().pack = value ? 1 : 0;
  } else if (key == 3) {
    tuning().xcd = value ? 1 : 0;
  } else if (key == 4) {
    DLP_CHECK_ARG(value >= 0 && value <= (1 << 20), "tune: dot grid out of range");
    tuning().dot_grid = value;
  } else if (key == 5) {
    tuning().fuse_dot = value ? 1 : 0;
  } else if (key == 6) {
    DLP_CHECK_ARG(value == 0 || value == 4 || value == 6 || value == 8,
                  "tune: pipelined operator depth must be 0, 4, 6 or 8");
    tuning().pipe = value;
  } else if (key == 7) {
    DLP_CHECK_ARG(value == 4 || value == 8, "tune: BF16 values per lane must be 4 or 8");
    tuning().bf16_word = value;
  } else if (key == 8) {
    DLP_CHECK_ARG(value >= 0 && value <= (1 << 20), "tune: flat pass grid out of range");
    tuning().flat_grid = value;
  } else if (key == 9) {
    tuning().nt_streams = value ? 1 : 0;
  } else if (key == 10) {
    tuning().defer_z = value ? 1 : 0;
  } else if (key == 11) {
    DLP_CHECK_ARG(value == 5 || value == 6, "tune: FP8 operator waves per SIMD at depth 6/8 must be 5 or 6");
    tuning().f8_waves = value;
  } else if (key == 12) {
    DLP_CHECK_ARG(value == 0 || value == 4 || value == 6 || value == 8,
                  "tune: FP8 operator depth must be 0 (by row width), 4, 6 or 8");
    tuning().f8_pipe = value;
  } else {
    DLP_CHECK_ARG(false, "tune: unknown key %d", key);
  }
  return DLP_OK;
}
