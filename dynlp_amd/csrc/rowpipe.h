// Software-pipelined row gathers over every row of a row-major block, one row per wavefront step
// (grid-stride, optionally XCD-aware), shared by the global solve's operator and the full
// residual verification.  Rows are FP32 (W = 4 columns per 16-byte lane word) or BF16 (W = 8 per
// 16-byte word, or W = 4 per 8-byte word).
//
// Lanes are packed (LanesP): ceil(C/W) column lanes times LE edge lanes.  While row
// i's neighbour rows are gathered (S gathers per lane issued before the first multiply-add), row
// i+1's first 64 (column, weight) pairs are already in flight; row ranges come from descriptors
// loaded 64 row steps ahead (one lane per row).  Neighbour rows are
// read with raw buffer loads at 32-bit byte offsets v*row_bytes + word*l (24-bit multiply-add);
// lanes without an edge get an out-of-range offset, for which the buffer load returns zeros, so
// no lane masking or select sits on the gather path.  Rows outside rowmask are skipped; neighbour
// rows are not masked, so the block must hold the values the caller wants summed for every
// neighbour (zero where a neighbour must not count).  Entries [indptr, indptr + rowlen) of a row
// are all live (deletions compact the row).  Callers guarantee n < 2^24 and n*row_bytes < 2^31.
#pragma once
#include "rowgather.h"

namespace dlp {

typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 buf_load4(__amdgpu_buffer_rsrc_t rs, unsigned off) {
  const u4v r = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0);
  return make_float4(__uint_as_float(r.x), __uint_as_float(r.y), __uint_as_float(r.z), __uint_as_float(r.w));
}

// BF16 <-> FP32: a BF16 value is the upper half of an FP32 word; rounding to nearest even.
__device__ __forceinline__ float bf16_lo(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi(unsigned u) { return __uint_as_float(u & 0xffff0000u); }
__device__ __forceinline__ unsigned bf16_bits(float x) {
  const unsigned u = __float_as_uint(x);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ unsigned bf16_pack(float lo, float hi) { return bf16_bits(lo) | (bf16_bits(hi) << 16); }
__device__ __forceinline__ float bf16_round(float x) { return __uint_as_float(bf16_bits(x) << 16); }

// FP16 <-> FP32.  The conversion to FP32 happens inside the multiply-add (v_fma_mix_f32 reads
// either half of a 32-bit register), so FP16 words cost no unpacking instructions.
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float f16_lo(unsigned u) { return (float)__builtin_bit_cast(h2v, u).x; }
__device__ __forceinline__ float f16_hi(unsigned u) { return (float)__builtin_bit_cast(h2v, u).y; }
// round to nearest even, saturating at +-65504 (a direction value never becomes infinite)
__device__ __forceinline__ unsigned f16_pack(float lo, float hi) {
  const float m = 65504.f;
  const h2v h = {(_Float16)fminf(fmaxf(lo, -m), m), (_Float16)fminf(fmaxf(hi, -m), m)};
  return __builtin_bit_cast(unsigned, h);
}

// FP8 E4M3 (OCP: bias 7, largest finite value 448, no infinity), four values per 32-bit word,
// first value in the low byte.  The hardware conversion rounds to nearest even and turns
// magnitudes past 448 into NaN, so values saturate at +-448 first.
__device__ __forceinline__ unsigned f8_pack4(float a, float b, float c, float d) {
  const float m = 448.f;
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(fminf(fmaxf(a, -m), m), fminf(fmaxf(b, -m), m), 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(fminf(fmaxf(c, -m), m), fminf(fmaxf(d, -m), m), w, true);
  return (unsigned)w;
}
__device__ __forceinline__ float4 f8_unpack4(unsigned u) {
  const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)u, false);
  const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)u, true);
  return make_float4(lo[0], lo[1], hi[0], hi[1]);
}

// Lane words of the gathered rows: FP32 x 4 (16 bytes), BF16 x 8 (16 bytes), BF16 x 4 or
// FP16 x 4 (8 bytes), FP32 x 4 at 4-byte alignment (F32x4u: four dword loads, for contiguous
// rows whose length is not a multiple of 4 columns), and FP8 (E4M3) x 4 (4 bytes, converted by
// the hardware's packed FP8 -> FP32 conversion).  W values per lane, column order.
enum class Word { F32x4, BF16x8, BF16x4, F16x4, F32x4u, F8x4 };

template <Word K> struct WordT;
template <> struct WordT<Word::F32x4> {
  static constexpr int W = 4, bytes = 16;
  using raw = u4v;
};
template <> struct WordT<Word::BF16x8> {
  static constexpr int W = 8, bytes = 16;
  using raw = u4v;
};
template <> struct WordT<Word::BF16x4> {
  static constexpr int W = 4, bytes = 8;
  using raw = unsigned int __attribute__((ext_vector_type(2)));
};
template <> struct WordT<Word::F16x4> {
  static constexpr int W = 4, bytes = 8;
  using raw = unsigned int __attribute__((ext_vector_type(2)));
};
template <> struct WordT<Word::F32x4u> {
  static constexpr int W = 4, bytes = 16;
  using raw = u4v;
};
template <> struct WordT<Word::F8x4> {
  static constexpr int W = 4, bytes = 4;
  using raw = unsigned int;
};

template <Word K>
__device__ __forceinline__ typename WordT<K>::raw load_word(__amdgpu_buffer_rsrc_t rs, unsigned off) {
  if constexpr (K == Word::F32x4u) {
    // an out-of-range offset (>= 2^31) stays out of range for all four dwords
    u4v r;
    r.x = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0);
    r.y = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(off + 4u), 0, 0);
    r.z = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(off + 8u), 0, 0);
    r.w = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(off + 12u), 0, 0);
    return r;
  } else if constexpr (WordT<K>::bytes == 16) {
    return __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0);
  } else if constexpr (WordT<K>::bytes == 4) {
    return __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0);
  } else {
    return __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, 0);
  }
}
template <Word K>
__device__ __forceinline__ typename WordT<K>::raw load_word(const void* p) {
  return *reinterpret_cast<const typename WordT<K>::raw*>(p);
}

template <Word K>
__device__ __forceinline__ void unpack(const typename WordT<K>::raw& u, float (&x)[WordT<K>::W]) {
  if constexpr (K == Word::F32x4 || K == Word::F32x4u) {
    x[0] = __uint_as_float(u.x);
    x[1] = __uint_as_float(u.y);
    x[2] = __uint_as_float(u.z);
    x[3] = __uint_as_float(u.w);
  } else if constexpr (K == Word::BF16x8) {
    x[0] = bf16_lo(u.x);
    x[1] = bf16_hi(u.x);
    x[2] = bf16_lo(u.y);
    x[3] = bf16_hi(u.y);
    x[4] = bf16_lo(u.z);
    x[5] = bf16_hi(u.z);
    x[6] = bf16_lo(u.w);
    x[7] = bf16_hi(u.w);
  } else if constexpr (K == Word::BF16x4) {
    x[0] = bf16_lo(u.x);
    x[1] = bf16_hi(u.x);
    x[2] = bf16_lo(u.y);
    x[3] = bf16_hi(u.y);
  } else if constexpr (K == Word::F8x4) {
    const float4 v = f8_unpack4(u);
    x[0] = v.x;
    x[1] = v.y;
    x[2] = v.z;
    x[3] = v.w;
  } else {
    x[0] = f16_lo(u.x);
    x[1] = f16_hi(u.x);
    x[2] = f16_lo(u.y);
    x[3] = f16_hi(u.y);
  }
}

// Rows of this wavefront in a grid-stride pass: [lo, hi) stepping nwx from lo + wx.  With
// xcd set (grid a multiple of 8) the blocks one XCD runs (blockIdx % 8) take one contiguous
// eighth of the rows.
__device__ __forceinline__ void xcd_rows(int nrows, int xcd, int64_t& lo, int64_t& hi, int64_t& wx, int64_t& nwx) {
  const int wave = threadIdx.x >> 6;
  lo = 0;
  hi = nrows;
  if (xcd) {
    const int64_t chunk = ((int64_t)nrows + 7) >> 3;
    lo = (blockIdx.x & 7) * chunk;
    hi = min((int64_t)nrows, lo + chunk);
    wx = (int64_t)(blockIdx.x >> 3) * 4 + wave;
    nwx = (int64_t)(gridDim.x >> 3) * 4;
  } else {
    wx = (int64_t)blockIdx.x * 4 + wave;
    nwx = (int64_t)gridDim.x * 4;
  }
}

// The rows of a pass are ordinals g in [0, nrows): row id rows[g] (rows null: g).  rowmask
// null: every row live; rowlen null: plain CSR (row length indptr[r+1] - indptr[r]).
struct RowPipe {
  LanesP L;
  int64_t lo, hi, wx, nwx;
  const int32_t* __restrict__ rows = nullptr;
  const uint8_t* __restrict__ rowmask;
  const int32_t* __restrict__ indptr;
  const int32_t* __restrict__ rowlen;
  const int32_t* __restrict__ indices;
  const float* __restrict__ w;
  __amdgpu_buffer_rsrc_t rx;
  unsigned rowb, laneq;

  // C columns of W per lane word of word_bytes; X rows of row_bytes bytes (4-byte multiples;
  // 16-byte and 8-byte words need rows that are multiples of the word), the block xbytes long
  __device__ __forceinline__ RowPipe(int nrows, int C, int W, int word_bytes, int xcd, const uint8_t* rm,
                                     const int32_t* ip, const int32_t* rl, const int32_t* ix, const float* wv,
                                     const void* X, unsigned row_bytes, int xbytes)
      : L(W == 4 ? C : (C + 1) / 2), rowmask(rm), indptr(ip), rowlen(rl), indices(ix), w(wv) {
    xcd_rows(nrows, xcd, lo, hi, wx, nwx);
    rx = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(X), (short)0, xbytes, 0x00020000);
    rowb = row_bytes;
    laneq = (unsigned)word_bytes * (unsigned)L.l;
  }
};

// x[q] = p[c0 + q] for c0 + q < C (0 beyond): an own-row read of 4 columns at 4-byte alignment.
__device__ __forceinline__ void load4(const float* __restrict__ p, int c0, int C, float (&x)[4]) {
  if (c0 + 3 < C) {
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = p[c0 + q];
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = c0 + q < C ? p[c0 + q] : 0.f;
  }
}

// p[c0 + q] = y[q] for c0 + q < C; one 16-byte store when vec (p + c0 16-byte aligned) and the
// four columns exist.
__device__ __forceinline__ void store4(float* __restrict__ p, int c0, int C, bool vec, const float (&y)[4]) {
  if (c0 + 3 < C) {
    if (vec) {
      *reinterpret_cast<float4*>(p + c0) = make_float4(y[0], y[1], y[2], y[3]);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) p[c0 + q] = y[q];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c0 + q < C) p[c0 + q] = y[q];
  }
}

template <Word K, typename TA>
__device__ __forceinline__ void fmaw(TA (&acc)[WordT<K>::W], float wv, const typename WordT<K>::raw& u) {
  constexpr int W = WordT<K>::W;
  float x[W];
  unpack<K>(u, x);
#pragma unroll
  for (int q = 0; q < W; ++q) {
    if constexpr (std::is_same<TA, float>::value) acc[q] = __builtin_fmaf(wv, x[q], acc[q]);
    else acc[q] += (TA)wv * (TA)x[q];
  }
}

// Row descriptors of 64 consecutive steps of this wavefront, one lane each: lane j describes
// ordinal base + j*nwx.  mrow = row id, with bit 30 set when the row is dead (outside rowmask) or
// past the end (row 0 then); [mbeg, mend) its entries (empty when dead).  The loads are issued a
// batch ahead of use, so a row step reads its descriptor with a lane read instead of waiting on
// dependent loads of the mask, the row start and the length.
constexpr int kDead = 1 << 30;
__device__ __forceinline__ void meta_load(const RowPipe& G, int64_t base, int& mrow, int& mbeg, int& mend) {
  const int64_t ii = base + (int64_t)(threadIdx.x & 63) * G.nwx;
  mrow = kDead;
  mbeg = mend = 0;
  if (ii < G.hi) {
    const int ri = G.rows ? G.rows[ii] : (int)ii;
    const bool live = G.rowmask ? G.rowmask[ri] != 0 : true;
    mrow = live ? ri : (ri | kDead);
    if (live) {
      mbeg = G.indptr[ri];
      mend = G.rowlen ? mbeg + G.rowlen[ri] : G.indptr[ri + 1];
    }
  }
}

// Lane l of the wavefront: v, wv = the (column, weight) pair beg + l when beg + l < end, else
// (0, 0).  The loads go through buffer resources spanning [beg, min(end, beg + 64)), so lanes past
// the end read zeros without a branch; every row step then issues the same number of vector
// memory operations, which lets the compiler wait for exactly the loads a step consumes instead
// of for everything in flight (stores of the previous row included).
__device__ __forceinline__ void load_pairs(const RowPipe& G, int beg, int end, int& v, float& wv) {
  const int cnt = __builtin_amdgcn_readfirstlane(max(0, min(64, end - beg)));
  const __amdgpu_buffer_rsrc_t ri =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(G.indices + beg), (short)0, cnt * 4, 0x00020000);
  const __amdgpu_buffer_rsrc_t rw =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(G.w + beg), (short)0, cnt * 4, 0x00020000);
  const int off = (threadIdx.x & 63) * 4;
  v = (int)__builtin_amdgcn_raw_buffer_load_b32(ri, off, 0, 0);
  wv = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rw, off, 0, 0));
}

// Runs every row of the pass: epi.pre(g, r) before row r = rows[g]'s gathers are issued
// (own-row loads), epi.post(g, r, live, acc) with live = rowmask[r] and acc = sum over the row's
// entries of w_j * X[col_j, W*l .. W*l+W-1] (zero when !live; valid in edge lane 0), called by
// every lane of the wavefront.
template <int S, Word K, typename TA, typename E>
__device__ __forceinline__ void run_rows(const RowPipe& G, E& epi) {
  constexpr int W = WordT<K>::W;
  using raw = typename WordT<K>::raw;
  const LanesP& L = G.L;
  constexpr unsigned kOOB = 0x80000000u;
  // descriptors: the current batch (c*) and the next one (n*), slot j of the current batch is
  // the row of step i
  int crow, cbeg, cend, nrow, nbeg, nend;
  const int64_t bstep = 64 * G.nwx;
  meta_load(G, G.lo + G.wx, crow, cbeg, cend);
  meta_load(G, G.lo + G.wx + bstep, nrow, nbeg, nend);
  auto take = [&](int vrow, int vbeg, int vend, int j, int& row, int& beg, int& end, bool& live) {
    const int m = __builtin_amdgcn_readlane(vrow, j);
    live = (m & kDead) == 0;
    row = m & (kDead - 1);
    beg = __builtin_amdgcn_readlane(vbeg, j);
    end = __builtin_amdgcn_readlane(vend, j);
  };
  int j = 0;
  // cnt_l: pairs this lane may read (0 on idle edge lanes); the select is branch-free
  auto gather = [&](int v, float wv, int t, int cnt_l, raw& xv, float& ww) {
    const int vv = __shfl(v, t & 63, 64);
    ww = __shfl(wv, t & 63, 64);
    const unsigned o = __umul24((unsigned)vv, G.rowb) + G.laneq;
    xv = load_word<K>(G.rx, t < cnt_l ? o : kOOB);
  };
  auto steps = [&](int v, float wv, int t0, int cnt, TA (&acc)[W]) {
    const int cnt_l = L.on ? cnt : 0;
    for (; t0 < cnt; t0 += 4 * L.LE) {
      raw yv[4];
      float yw[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) gather(v, wv, t0 + u * L.LE + L.e, cnt_l, yv[u], yw[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u) fmaw<K>(acc, yw[u], yv[u]);
    }
  };
  int64_t i = G.lo + G.wx;
  int row0, beg0, end0;
  bool live0;
  take(crow, cbeg, cend, 0, row0, beg0, end0, live0);
  int v0, v1;
  float w0, w1;
  load_pairs(G, beg0, end0, v0, w0);
  // everything issued so far has landed before the loop: otherwise the compiler's wait for the
  // first row's pairs (a load still in flight only on entry) is repeated on every row step
  __builtin_amdgcn_s_waitcnt(0);
  for (; i < G.hi; i += G.nwx) {
    int row1, beg1, end1;
    bool live1;
    if (j < 63) take(crow, cbeg, cend, j + 1, row1, beg1, end1, live1);
    else take(nrow, nbeg, nend, 0, row1, beg1, end1, live1);
    const int g = __builtin_amdgcn_readfirstlane((int)i);
    const int r = row0;
    epi.pre(g, r);
    const int cnt = min(64, end0 - beg0);
    const int cnt_l = L.on ? cnt : 0;
    raw xv[S];
    float ww[S];
#pragma unroll
    for (int u = 0; u < S; ++u) gather(v0, w0, u * L.LE + L.e, cnt_l, xv[u], ww[u]);
    // next row's pairs go out while this row's gathers are in flight
    load_pairs(G, beg1, end1, v1, w1);
    TA acc[W];
#pragma unroll
    for (int q = 0; q < W; ++q) acc[q] = 0;
#pragma unroll
    for (int u = 0; u < S; ++u) fmaw<K>(acc, ww[u], xv[u]);
    steps(v0, w0, S * L.LE, cnt, acc);
    // rows longer than 64 entries: further chunks of 64 pairs
    for (int base = beg0 + 64; base < end0; base += 64) {
      int vc;
      float wc;
      load_pairs(G, base, end0, vc, wc);
      steps(vc, wc, 0, min(64, end0 - base), acc);
    }
    // edge-lane partials of column lane l are at lanes l + k*LC
    TA tot[W];
#pragma unroll
    for (int q = 0; q < W; ++q) tot[q] = acc[q];
    for (int k = 1; k < L.LE; ++k) {
      const int src = L.lane + k * L.LC;
#pragma unroll
      for (int q = 0; q < W; ++q) tot[q] += __shfl(acc[q], src, 64);
    }
    epi.post(g, r, live0, tot);
    v0 = v1;
    w0 = w1;
    row0 = row1;
    beg0 = beg1;
    end0 = end1;
    live0 = live1;
    if (++j == 64) {
      // the next batch becomes current; the one after it is loaded now (first used 63 steps on)
      j = 0;
      crow = nrow;
      cbeg = nbeg;
      cend = nend;
      meta_load(G, i + G.nwx + bstep, nrow, nbeg, nend);
    }
  }
}

}  // namespace dlp
