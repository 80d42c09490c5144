// Vector distance (L2 / cosine) and top-k search.
//
// `oxr_vector_distance` replaces `ops::vector_distance` (crates/oxidelake-
// device/src/cuda/ops.rs:1163-1249, kernels/cuda/vector_distance.cu): one
// wave per row, the query in LDS, 16-byte loads, a wave reduction; the
// distance column is appended to views of the input's columns (no copy).
//
// Top-k is new (the reference sorts distances on the CPU, DataFusion's
// SortExec with a fetch: crates/oxidelake-runtime/tests/embedded.rs:24-25).
// Queries run in batches through four kernels:
//   1. scan: rows x queries on the matrix cores — FP32 (32x32x2), BF16
//      (32x32x16) or FP8 e4m3 with per-row scales (32x32x64) — with FP32
//      accumulation; the epilogue folds row norms and scales into a score
//      that orders rows like the distance (L2: 2 x.q - |x|^2, cosine:
//      x.q / |x|) and writes a tile of scores per query;
//   2. select: a streaming top-C per (segment, query) with a threshold and
//      an LDS candidate buffer, then a merge per query;
//   3. rerank: exact FP32 distances of the C candidates from the FP32 rows,
//      sorted by (distance, row); the first k are the answer.
// With FP32 and C = k this is the plain exact path; BF16/FP8 with C > k is
// the reduced-precision scan with an FP32 rerank.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>

#include "common.h"

struct oxr_vector_index {
  oxr_device *dev = nullptr;
  oxr::StreamCore *owner = nullptr;
  uint64_t n = 0;
  int32_t dim = 0, dimp = 0, metric = 0, precision = 0;
  const float *f32 = nullptr;      // the FP32 rows (the input batch's column)
  const uint32_t *valid = nullptr; // their validity (NULL: all valid)
  std::shared_ptr<void> holder;    // keeps the input's buffers alive
  void *scan = nullptr;            // rows as the scan reads them: [n][dimp]
  bool scan_owned = false;
  float *rowinfo = nullptr;        // L2: |x|^2; cosine: 1/|x| (0: zero norm)
  float *rscale = nullptr;         // FP8: per-row scale
};

namespace oxr {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

enum { P_FP32 = OXR_VEC_FP32, P_BF16 = OXR_VEC_BF16, P_FP8 = OXR_VEC_FP8 };
constexpr float kFp8Max = 448.0f;  // OCP e4m3 (gfx950)
constexpr int kBlock = 256;
constexpr int kMaxDim = 8192;      // the query in LDS (with the rerank keys: < 64 KB)
constexpr int kMaxCandidates = 1024;
constexpr int kMaxScanQueries = 256;  // queries per scan pass (8 tiles of 32)

__device__ inline float wave_sum(float x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}
__device__ inline float wave_max(float x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x = fmaxf(x, __shfl_xor(x, d, 64));
  return x;
}

__device__ inline bool valid_row(const uint32_t *valid, uint64_t r) {
  return !valid || ((valid[r >> 5] >> (r & 31)) & 1u);
}

// ------------------------------------------------------- one-query distance
struct DistArgs {
  const float *x;
  const uint32_t *valid;
  uint64_t n;
  int32_t dim, metric;
  const float *q;
  float *out;
};

// Per-lane partial sums over a row: L2 (sum of squared differences) or
// cosine (dot, |x|^2). 16-byte loads when rows are 16-byte aligned.
__device__ inline void row_partials(const float *x, const float *q, int dim, int metric, int lane,
                                    float *p0, float *p1) {
  float s0 = 0.f, s1 = 0.f;
  if ((dim & 3) == 0) {
    const f32x4 *x4 = (const f32x4 *)x;
    const f32x4 *q4 = (const f32x4 *)q;
    for (int c = lane; c < dim / 4; c += 64) {
      const f32x4 a = x4[c], b = q4[c];
      if (metric == OXR_COSINE) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
          s0 = __fmaf_rn(a[e], b[e], s0);
          s1 = __fmaf_rn(a[e], a[e], s1);
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const float d = a[e] - b[e];
          s0 = __fmaf_rn(d, d, s0);
        }
      }
    }
  } else {
    for (int c = lane; c < dim; c += 64) {
      const float a = x[c], b = q[c];
      if (metric == OXR_COSINE) {
        s0 = __fmaf_rn(a, b, s0);
        s1 = __fmaf_rn(a, a, s1);
      } else {
        const float d = a - b;
        s0 = __fmaf_rn(d, d, s0);
      }
    }
  }
  *p0 = s0;
  *p1 = s1;
}

// row_partials for R rows at once: the same per-row arithmetic in the same
// order (so the same bits), with the R rows' loads in flight together.
template <int R>
__device__ inline void row_partials_multi(const float *x, const uint64_t (&rows)[R], const float *q, int dim,
                                          int metric, int lane, float (&p0)[R], float (&p1)[R]) {
#pragma unroll
  for (int r = 0; r < R; r++) p0[r] = p1[r] = 0.f;
  if ((dim & 3) == 0) {
    const f32x4 *q4 = (const f32x4 *)q;
    for (int c = lane; c < dim / 4; c += 64) {
      const f32x4 b = q4[c];
      f32x4 av[R];
#pragma unroll
      for (int r = 0; r < R; r++) av[r] = ((const f32x4 *)(x + rows[r] * (uint64_t)dim))[c];
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (metric == OXR_COSINE) {
#pragma unroll
          for (int e = 0; e < 4; e++) {
            p0[r] = __fmaf_rn(av[r][e], b[e], p0[r]);
            p1[r] = __fmaf_rn(av[r][e], av[r][e], p1[r]);
          }
        } else {
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const float d = av[r][e] - b[e];
            p0[r] = __fmaf_rn(d, d, p0[r]);
          }
        }
      }
    }
  } else {
    for (int c = lane; c < dim; c += 64) {
      const float b = q[c];
      float av[R];
#pragma unroll
      for (int r = 0; r < R; r++) av[r] = x[rows[r] * (uint64_t)dim + c];
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (metric == OXR_COSINE) {
          p0[r] = __fmaf_rn(av[r], b, p0[r]);
          p1[r] = __fmaf_rn(av[r], av[r], p1[r]);
        } else {
          const float d = av[r] - b;
          p0[r] = __fmaf_rn(d, d, p0[r]);
        }
      }
    }
  }
}

// L2 = sqrt(sum (a-b)^2); cosine = 1 - dot/(|a||b|), NaN when a norm is 0
// (crates/oxidelake-device/src/cpu/kernels.rs:908-927).
__device__ inline float finish_distance(int metric, float s0, float s1, float qnorm2) {
  if (metric == OXR_COSINE) {
    const float na = sqrtf(s1), nb = sqrtf(qnorm2);
    if (na == 0.f || nb == 0.f) return __builtin_nanf("");
    return 1.0f - s0 / (na * nb);
  }
  return sqrtf(s0);
}

__global__ __launch_bounds__(kBlock) void distance_kernel(const DistArgs a) {
  extern __shared__ float s_q[];
  for (int i = threadIdx.x; i < a.dim; i += kBlock) s_q[i] = a.q[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  float qn = 0.f;
  for (int i = lane; i < a.dim; i += 64) qn = __fmaf_rn(s_q[i], s_q[i], qn);
  const float qnorm2 = wave_sum(qn);
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t r = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); r < a.n; r += waves) {
    if (!valid_row(a.valid, r)) {
      if (lane == 0) a.out[r] = 0.f;
      continue;
    }
    float s0, s1;
    row_partials(a.x + r * (uint64_t)a.dim, s_q, a.dim, a.metric, lane, &s0, &s1);
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if (lane == 0) a.out[r] = finish_distance(a.metric, s0, s1, qnorm2);
  }
}

// ------------------------------------------------------------- index build
// One wave per row: |x|^2 (and 1/|x|), and the row in the scan's precision
// (BF16 round-to-nearest-even, FP8 e4m3 with scale max|x| / 448), padded.
struct BuildArgs {
  const float *x;
  uint64_t n;
  int32_t dim, dimp, metric, precision;
  void *scan;
  float *rowinfo, *rscale;
};

__global__ __launch_bounds__(kBlock) void build_kernel(const BuildArgs a) {
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t r = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); r < a.n; r += waves) {
    const float *x = a.x + r * (uint64_t)a.dim;
    float ss = 0.f, mx = 0.f;
    for (int c = lane; c < a.dim; c += 64) {
      const float v = x[c];
      ss = __fmaf_rn(v, v, ss);
      mx = fmaxf(mx, fabsf(v));
    }
    ss = wave_sum(ss);
    mx = wave_max(mx);
    if (lane == 0) {
      if (a.metric == OXR_COSINE) a.rowinfo[r] = ss > 0.f ? 1.0f / sqrtf(ss) : 0.f;
      else a.rowinfo[r] = ss;
    }
    if (a.precision == P_BF16) {
      __bf16 *o = (__bf16 *)a.scan + r * (uint64_t)a.dimp;
      for (int c = lane; c < a.dimp; c += 64) o[c] = (__bf16)(c < a.dim ? x[c] : 0.f);
    } else if (a.precision == P_FP8) {
      const float s = mx > 0.f ? mx / kFp8Max : 1.0f;
      if (lane == 0) a.rscale[r] = s;
      int *o = (int *)((uint8_t *)a.scan + r * (uint64_t)a.dimp);
      for (int c4 = lane; c4 < a.dimp / 4; c4 += 64) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int c = c4 * 4 + e;
          v[e] = c < a.dim ? fminf(fmaxf(x[c] / s, -kFp8Max), kFp8Max) : 0.f;
        }
        const int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
        o[c4] = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
      }
    } else if (a.precision == P_FP32 && a.scan) {  // padded FP32 copy
      float *o = (float *)a.scan + r * (uint64_t)a.dimp;
      for (int c = lane; c < a.dimp; c += 64) o[c] = c < a.dim ? x[c] : 0.f;
    }
  }
}

// Queries in the scan's precision: [qpad][dimp]; FP8 queries get a scale
// each. One block per query row; rows past nq are zero.
struct PrepArgs {
  const float *q;  // [nq][dim]
  int32_t nq, qpad, dim, dimp, precision;
  void *out;
  float *qscale;
};

__global__ __launch_bounds__(kBlock) void prep_kernel(const PrepArgs a) {
  __shared__ float s_max[kBlock / 64];
  const int qi = blockIdx.x;
  const bool real = qi < a.nq;
  const float *q = a.q + (uint64_t)(real ? qi : 0) * a.dim;
  float mx = 0.f;
  if (real)
    for (int c = threadIdx.x; c < a.dim; c += kBlock) mx = fmaxf(mx, fabsf(q[c]));
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = mx;
  __syncthreads();
  mx = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
  if (a.precision == P_FP32) {
    float *o = (float *)a.out + (uint64_t)qi * a.dimp;
    for (int c = threadIdx.x; c < a.dimp; c += kBlock) o[c] = real && c < a.dim ? q[c] : 0.f;
  } else if (a.precision == P_BF16) {
    __bf16 *o = (__bf16 *)a.out + (uint64_t)qi * a.dimp;
    for (int c = threadIdx.x; c < a.dimp; c += kBlock) o[c] = (__bf16)(real && c < a.dim ? q[c] : 0.f);
  } else {
    const float s = mx > 0.f ? mx / kFp8Max : 1.0f;
    if (threadIdx.x == 0) a.qscale[qi] = s;
    int *o = (int *)((uint8_t *)a.out + (uint64_t)qi * a.dimp);
    for (int c4 = threadIdx.x; c4 < a.dimp / 4; c4 += kBlock) {
      float v[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int c = c4 * 4 + e;
        v[e] = real && c < a.dim ? fminf(fmaxf(q[c] / s, -kFp8Max), kFp8Max) : 0.f;
      }
      const int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
      o[c4] = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
    }
  }
}

// -------------------------------------------------------------------- scan
// Rows x queries on the matrix cores. A block is 4 waves; wave w computes
// rows [32*RT*w, 32*RT*(w+1)) of the block's row range against NT tiles of
// 32 queries. The k axis is walked in groups: per group each lane loads LB
// bytes of its row (lane l: row l%32, the half l/32 of the group) straight
// from HBM, and the group's slice of every query is staged through LDS
// (double-buffered). Verified lane layouts (32x32 instructions): A row and
// B column = lane%32, both at the same k set for lane half l/32; result
// element r of lane l is (row 8*(r/4) + 4*(l/32) + r%4, query l%32).
// Keys that order candidates: larger is better. The score's bits made
// order-preserving as an unsigned integer, then the row (smaller row wins a
// tie). 0 = no candidate.
__device__ inline uint32_t ordered(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline uint64_t score_key(float score, uint64_t row) {
  return ((uint64_t)ordered(score) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)row);
}
__device__ inline uint64_t key_row(uint64_t key) { return 0xFFFFFFFFu - (uint32_t)key; }
// The float whose `ordered` is k (k = 0, below every score, comes back NaN,
// which no compare drops on).
__device__ inline float from_ordered(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

struct ScanArgs {
  const void *x;      // [n][dimp] in the scan precision
  const void *qm;     // [qpad][dimp]
  uint64_t n, tile0, tile_rows, nblocks;
  int32_t dimp, groups, nq, metric;
  const float *rowinfo, *rscale, *qscale;
  const uint32_t *valid;
  float *scores;      // [nq][tile_rows]: scores ordered like the distance
  int32_t qoff;       // first query of this launch
  // append mode (list != NULL): no scores; a score at or above its query's
  // bound (the ordered score of tau_keys[q * C + C - 1]) is appended to
  // list[q * cap + atomicAdd(&cnt[q], 1)] as a key (when below cap)
  const uint64_t *tau_keys;
  uint64_t *list;
  uint32_t *cnt;
  uint64_t cap;
  int32_t C;
};

template <int P>
struct Prec;
template <>
struct Prec<P_FP32> {
  static constexpr int EB = 4, KG = 8, LB = 16;  // 4 steps of 32x32x2 per group
};
template <>
struct Prec<P_BF16> {
  static constexpr int EB = 2, KG = 16, LB = 16;  // one 32x32x16
};
template <>
struct Prec<P_FP8> {
  static constexpr int EB = 1, KG = 64, LB = 32;  // one 32x32x64
};

// Byte offset, within a row, of k group g's slice for lane half h. The k
// axis is permuted (a dot product does not care) so that the R k groups of
// one 128-byte line are read by the same two lanes, 64 contiguous bytes
// each: a row's line is consumed by back-to-back loads instead of being
// fetched again for every k group.
template <int P>
__device__ inline uint32_t koff(int g, int h) {
  constexpr int LB = Prec<P>::LB, R = 128 / (Prec<P>::KG * Prec<P>::EB);
  return (uint32_t)((g / R) * 128 + h * 64 + (g % R) * LB);
}

template <int P>
struct Frag {
  uint4 v[Prec<P>::LB / 16];
};

template <int P>
__device__ inline void mfma_group(f32x16 &acc, const Frag<P> &a, const Frag<P> &b) {
  if constexpr (P == P_FP32) {
    f32x4 x, y;
    __builtin_memcpy(&x, &a.v[0], 16);
    __builtin_memcpy(&y, &b.v[0], 16);
#pragma unroll
    for (int j = 0; j < 4; j++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[j], y[j], acc, 0, 0, 0);
  } else if constexpr (P == P_BF16) {
    bf16x8 x, y;
    __builtin_memcpy(&x, &a.v[0], 16);
    __builtin_memcpy(&y, &b.v[0], 16);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, acc, 0, 0, 0);
  } else {
    i32x8 x, y;
    __builtin_memcpy(&x, &a.v[0], 32);
    __builtin_memcpy(&y, &b.v[0], 32);
    acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(x, y, acc, 0, 0, 0, 127, 0, 127);
  }
}

// BRES: the whole query tile (every k group) sits in dynamic LDS, loaded
// once per block, so the k loop has no barriers and each wave streams its
// rows at its own pace (small query tiles, where the scan is HBM-bound).
// QW: waves per row group. The block's 4 waves form (4 / QW) row groups of
// RT row tiles times QW query groups of NT query tiles: QW > 1 shares each
// row's loads between QW waves (through L1) for large query batches.
template <int P, int NT, int RT, bool BRES, int QW>
__global__ __launch_bounds__(kBlock) void scan_kernel(const ScanArgs a) {
  constexpr int LB = Prec<P>::LB, KG = Prec<P>::KG, EB = Prec<P>::EB;
  constexpr int U = LB / 16;                 // 16-byte units per lane per group
  constexpr int QROW = 2 * U;                // units per query per group
  constexpr int QTOT = QW * NT * 32;         // queries of the block's tile
  constexpr int BUNITS = QTOT * QROW;         // units of one group's query slice
  // a group's slice sits in LDS part-major ([part][query]), so the lanes
  // of a fragment read (one query each) read consecutive 16-byte units:
  // query-major, they were a query row apart and shared banks
  constexpr int PER_T = (BUNITS + kBlock - 1) / kBlock;
  constexpr int BROWS = (4 / QW) * RT * 32;  // rows per block
  __shared__ uint4 s_b[2][BRES ? 1 : BUNITS];
  // per row of the block, read once into LDS (so the epilogue holds no row
  // data in registers across the k loop): score = (dot * qscale) * mul -
  // bias, with mul = scale / |x| and bias 0 (cosine) or mul = 2 * scale and
  // bias |x|^2 (L2); bias = +inf marks a row that cannot qualify
  __shared__ __align__(16) float s_mul[BROWS];
  __shared__ __align__(16) float s_bias[BROWS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 31, h = lane >> 5;
  const int rg = wave / QW, qg = wave % QW;  // the wave's row group and query group
  const uint64_t rstride = (uint64_t)a.dimp * EB;  // bytes per row
  const uint8_t *xb = (const uint8_t *)a.x;
  const uint8_t *qb = (const uint8_t *)a.qm;
  const uint64_t tile_end = min(a.n, a.tile0 + a.tile_rows);
  extern __shared__ uint4 s_bres[];  // BRES: [groups][BUNITS]
  if (BRES) {
    const int total = a.groups * BUNITS;
    for (int unit = tid; unit < total; unit += kBlock) {
      const int g = unit / BUNITS, rem = unit % BUNITS;
      const int qrow = rem / QROW, part = rem % QROW;
      s_bres[g * BUNITS + part * QTOT + qrow] =
          *(const uint4 *)(qb + (uint64_t)(a.qoff + qrow) * rstride + koff<P>(g, part / U) + (part % U) * 16);
    }
  }

  for (uint64_t blk = blockIdx.x; blk < a.nblocks; blk += gridDim.x) {
    const uint64_t blockrow = a.tile0 + blk * (uint64_t)BROWS;
    if (blockrow >= tile_end) break;  // never taken: the host sizes nblocks to the tile
    const uint64_t rowbase = blockrow + (uint64_t)rg * RT * 32;
    // the block's rows from one uniform base, each lane's row as a 32-bit
    // offset computed once per block; a row past the tile reads the tile's
    // last row: its scores are dropped (its bias is +inf), so what it
    // multiplies does not matter
    const uint8_t *bbase = xb + blockrow * rstride;
    uint32_t loff[RT];
#pragma unroll
    for (int r = 0; r < RT; r++) {
      const uint64_t row = rowbase + r * 32 + i;
      const uint64_t arow = row < tile_end ? row : tile_end - 1;
      loff[r] = (uint32_t)((arow - blockrow) * rstride) + (uint32_t)(h * 64);
    }
    // every wave has finished the previous block's epilogue (its reads of
    // s_mul and s_bias) before they are rewritten; on the streamed path the
    // k loop's barriers come after this fill, so they do not order it
    __syncthreads();
    for (int t = tid; t < BROWS; t += kBlock) {
      const uint64_t row = blockrow + t;
      float mul = 0.f, bias = INFINITY;  // NULL, zero norm, past the tile
      if (row < tile_end && valid_row(a.valid, row)) {
        const float ri = a.rowinfo[row];
        const float scl = P == P_FP8 ? a.rscale[row] : 1.0f;
        if (a.metric != OXR_COSINE) {
          mul = 2.0f * scl;
          bias = ri;
        } else if (ri > 0.f) {
          mul = scl * ri;
          bias = 0.f;
        }
      }
      s_mul[t] = mul;
      s_bias[t] = bias;
    }
    f32x16 acc[RT][NT];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[r][t] = f32x16{};

    auto load_a = [&](int g, Frag<P> (&fa)[RT]) {
      const uint8_t *gb = bbase + koff<P>(g, 0);  // uniform; the lane half is in loff
#pragma unroll
      for (int r = 0; r < RT; r++) {
        const uint4 *src = (const uint4 *)(gb + loff[r]);
#pragma unroll
        for (int u = 0; u < U; u++) fa[r].v[u] = src[u];
      }
    };
    auto load_b = [&](int g, uint4 (&rb)[PER_T]) {
#pragma unroll
      for (int p = 0; p < PER_T; p++) {
        const int unit = p * kBlock + tid;
        if (unit < BUNITS) {
          const int qrow = unit / QROW, part = unit % QROW;
          rb[p] = *(const uint4 *)(qb + (uint64_t)(a.qoff + qrow) * rstride + koff<P>(g, part / U) + (part % U) * 16);
        }
      }
    };
    auto store_b = [&](int buf, const uint4 (&rb)[PER_T]) {
#pragma unroll
      for (int p = 0; p < PER_T; p++) {
        const int unit = p * kBlock + tid;
        if (unit < BUNITS) s_b[buf][(unit % QROW) * QTOT + unit / QROW] = rb[p];
      }
    };

    Frag<P> fa[RT], fn[RT];
    if (BRES) {
      __syncthreads();  // s_info (and, the first time, s_bres) written
      // three fragment sets in rotation keep two k groups of loads in
      // flight while the third is multiplied
      auto mul = [&](int g, const Frag<P> (&f)[RT]) {
        const uint4 *bg = s_bres + (uint64_t)g * BUNITS;
#pragma unroll
        for (int t = 0; t < NT; t++) {
          Frag<P> fb;
#pragma unroll
          for (int u = 0; u < U; u++) fb.v[u] = bg[(h * U + u) * QTOT + (qg * NT + t) * 32 + i];
#pragma unroll
          for (int r = 0; r < RT; r++) mfma_group<P>(acc[r][t], f[r], fb);
        }
      };
      // D fragment sets in rotation (group x in set x % D): D - 1 k groups
      // of loads stay in flight while one is multiplied. The BF16 wide
      // tile (64 resident queries, one block per CU, so one wave per SIMD)
      // goes six deep: 37.9 ms against 46.5 streamed for 256 queries over
      // 20M x 768. FP8's wide tile measured no gain past three (its
      // fragments are twice as large and spill into the AGPRs).
      constexpr int D = NT * RT >= 8 && U == 1 ? 6 : 3;
      Frag<P> fr[D][RT];
      const int G = a.groups;
#pragma unroll
      for (int d = 0; d < D - 1; d++)
        if (d < G) load_a(d, fr[d]);
      for (int g0 = 0; g0 < G; g0 += D) {
#pragma unroll
        for (int d = 0; d < D; d++) {
          const int g = g0 + d;
          if (g + D - 1 < G) load_a(g + D - 1, fr[(d + D - 1) % D]);
          if (g < G) mul(g, fr[d]);
        }
      }
    }
    uint4 rb[PER_T];
    if (!BRES) {
    load_a(0, fa);
    load_b(0, rb);
    store_b(0, rb);
    __syncthreads();
    }
    for (int g = 0; !BRES && g < a.groups; g++) {
      const bool more = g + 1 < a.groups;
      if (more) {
        load_a(g + 1, fn);
        load_b(g + 1, rb);
      }
      const int buf = g & 1;
#pragma unroll
      for (int t = 0; t < NT; t++) {
        Frag<P> fb;
#pragma unroll
        for (int u = 0; u < U; u++) fb.v[u] = s_b[buf][(h * U + u) * QTOT + (qg * NT + t) * 32 + i];
#pragma unroll
        for (int r = 0; r < RT; r++) mfma_group<P>(acc[r][t], fa[r], fb);
      }
      if (more) {
        store_b(buf ^ 1, rb);
#pragma unroll
        for (int r = 0; r < RT; r++) fa[r] = fn[r];
      }
      __syncthreads();
    }

    // epilogue: scores ordered like the distance (larger is nearer):
    // L2 2 x.q - |x|^2, cosine x.q / |x|; -inf where a row cannot qualify
    // (NULL, zero norm, past the tile). Row factors come from LDS (filled
    // before the k loop; the loop's barriers order it); per query tile the
    // query's scale, and in append mode its bound as a float: a score below
    // it is dropped (the float compare keeps a -0.0 score at a +0.0 bound,
    // which ranks below every pilot key and so never reaches the top C).
    int qv[NT];
    float qsv[NT], tauv[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int q = a.qoff + (qg * NT + t) * 32 + i;
      qv[t] = q < a.nq ? q : -1;
      qsv[t] = (P == P_FP8 && q < a.nq) ? a.qscale[q] : 1.0f;
      tauv[t] = (a.list && q < a.nq) ? from_ordered((uint32_t)(a.tau_keys[(uint64_t)q * a.C + a.C - 1] >> 32))
                                     : -INFINITY;
    }
#pragma unroll
    for (int r = 0; r < RT; r++) {
#pragma unroll
      for (int jg = 0; jg < 4; jg++) {
        // elements 4jg..4jg+3 of the lane are rows 8jg + 4h + 0..3: one
        // 16-byte load of their factors, one 16-byte store of their scores
        // (tile_rows is a multiple of 128, so they are aligned)
        const int m0 = 8 * jg + 4 * h;
        const int br = rg * RT * 32 + r * 32 + m0;
        const float4 mul4 = *(const float4 *)&s_mul[br];
        const float4 bias4 = *(const float4 *)&s_bias[br];
        const float mulv[4] = {mul4.x, mul4.y, mul4.z, mul4.w};
        const float biasv[4] = {bias4.x, bias4.y, bias4.z, bias4.w};
        const uint64_t row0 = rowbase + r * 32 + m0;
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const int q = qv[t];
          if (q < 0) continue;
          float sc[4];
#pragma unroll
          for (int e4 = 0; e4 < 4; e4++) {
            const float s = (acc[r][t][jg * 4 + e4] * qsv[t]) * mulv[e4] - biasv[e4];
            sc[e4] = biasv[e4] == INFINITY ? -INFINITY : s;
          }
          if (a.list) {
#pragma unroll
            for (int e4 = 0; e4 < 4; e4++) {
              if (!(sc[e4] > -INFINITY) || sc[e4] < tauv[t]) continue;
              const uint32_t pos = atomicAdd(&a.cnt[q], 1u);
              if (pos < a.cap) a.list[(uint64_t)q * a.cap + pos] = score_key(sc[e4], row0 + e4);
            }
          } else {
            float *out = a.scores + (uint64_t)q * a.tile_rows + (row0 - a.tile0);
            if (row0 + 4 <= tile_end) {
              *(float4 *)out = make_float4(sc[0], sc[1], sc[2], sc[3]);
            } else {
#pragma unroll
              for (int e4 = 0; e4 < 4; e4++)
                if (row0 + e4 < tile_end) out[e4] = sc[e4];
            }
          }
        }
        // one row group of four at a time: the scheduler may not batch
        // every tile's LDS reads ahead (that would cost occupancy)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
}

// ------------------------------------------------------------- GEMM scan
// Large batches (129-256 queries) as one GEMM tile per block: 256 rows x
// 256 queries, eight waves (two per SIMD) in a 4 x 2 grid, each wave 2 x 4
// tiles of 32 x 32 (64 rows x 128 queries, 128 accumulators). The k axis
// is walked a 128-byte line at a time: the line of each of the block's 256
// rows and of each of the 256 queries is copied into LDS by the direct
// global-to-LDS loads ([256][8] 16-byte units each), and read from there by
// the waves that need it, so every row is read from HBM once per 256
// queries (the resident 128-query kernel reads it twice) and no wave loads
// what another loads. Even lines go to one pair of LDS arrays and odd lines
// to another (the host runs this kernel for an even number of lines), so
// the compiler can tell the line being multiplied from the one in flight
// and waits for the copy only at the line's barrier: the next line's copy
// overlaps this line's products (the block's last line copies the next
// block's first). Unit u of row r sits at u ^ ((r >> 1) & 7): the LDS
// serves a 16-byte read to 16 lanes a clock (256 bytes, 64 banks), and a
// fragment read's 16 lanes (16 consecutive rows at one unit) then fall in
// 16 different 16-byte bank slots (with u ^ (r & 7), rows r and r + 8
// collided: 41% of the LDS's active cycles were bank conflicts). The k
// group loop stays rolled (its
// two unrolled copies' fragments spilled registers for FP8). Row factors
// are double-buffered by block parity; the epilogue is the resident
// kernels' (scores or appends).
constexpr int kGemmRows = 256, kGemmQueries = 256, kGemmThreads = 512;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Row factors of block `blk` into LDS (threads 0..255, one row each):
// scores are (acc * qscale) * mul - bias; bias +inf drops the row (NULL,
// zero norm, past the tile); blo is the bias lowered by 4e-6 of its
// magnitude (the append pre-test of gemm_epilogue).
template <int P>
__device__ __forceinline__ void gemm_fill_rows(const ScanArgs &a, uint64_t blk, int tid, uint64_t tile_end,
                                               float *s_mul, float *s_bias, float *s_blo) {
  if (tid >= kGemmRows) return;
  const uint64_t row = a.tile0 + blk * (uint64_t)kGemmRows + (uint64_t)tid;
  float mul = 0.f, bias = INFINITY;
  if (row < tile_end && valid_row(a.valid, row)) {
    const float ri = a.rowinfo[row];
    const float scl = P == P_FP8 ? a.rscale[row] : 1.0f;
    if (a.metric != OXR_COSINE) {
      mul = 2.0f * scl;
      bias = ri;
    } else if (ri > 0.f) {
      mul = scl * ri;
      bias = 0.f;
    }
  }
  s_mul[tid] = mul;
  s_bias[tid] = bias;
  s_blo[tid] = bias == INFINITY ? INFINITY : bias - 4e-6f * fabsf(bias);
}

// Per query tile t of a wave (queries wq): the query's scale and, in append
// mode, its bound as a float. Fixed for the launch, so loaded once.
template <int P, int NT>
__device__ __forceinline__ void gemm_query_factors(const ScanArgs &a, int wq, int i, float (&qs)[NT],
                                                   float (&tau)[NT]) {
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int q = a.qoff + wq * NT * 32 + t * 32 + i;
    qs[t] = P == P_FP8 && q < a.nq ? a.qscale[q] : 1.0f;
    tau[t] = a.list && q < a.nq ? from_ordered((uint32_t)(a.tau_keys[(uint64_t)q * a.C + a.C - 1] >> 32))
                                : -INFINITY;
  }
}

// The GEMM tile's epilogue for one wave (rows wr, queries wq of the block
// at `blockrow`): scores into a.scores, or appends past each query's bound.
template <int P, int RT, int NT>
__device__ __forceinline__ void gemm_epilogue(const ScanArgs &a, uint64_t blockrow, int wr, int wq, int i, int h,
                                              const f32x16 (&acc)[RT][NT], const float (&qsv)[NT],
                                              const float (&tauv)[NT], const float *s_mul, const float *s_bias,
                                              const float *s_blo) {
  const uint64_t tile_end = min(a.n, a.tile0 + a.tile_rows);
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int q = a.qoff + wq * NT * 32 + t * 32 + i;
    if (q >= a.nq) continue;
    const float qs = qsv[t];
    const float tau = tauv[t];
    // append mode: a score sc = (acc * qs) * mul - bias is kept when
    // sc >= tau; the pre-test acc * (qs * mul) >= tau_lo + blo, with
    // tau and the bias lowered by 4e-6 of their magnitudes, passes
    // every kept score (the two products differ by a few ulps of a
    // value that, near the bound, is at most |tau| + |bias|), so the
    // exact score is computed only for the few that pass it. A NaN
    // bound (a pilot with fewer than C rows) keeps every score.
    const float tau_lo = tau != tau ? -INFINITY : tau - 4e-6f * fabsf(tau);
#pragma unroll
    for (int r = 0; r < RT; r++) {
#pragma unroll
      for (int jg = 0; jg < 4; jg++) {
        const int m0 = 8 * jg + 4 * h;
        const int br = wr * RT * 32 + r * 32 + m0;
        const float4 mul4 = *(const float4 *)&s_mul[br];
        const float4 bias4 = *(const float4 *)&s_bias[br];
        const float mulv[4] = {mul4.x, mul4.y, mul4.z, mul4.w};
        const float biasv[4] = {bias4.x, bias4.y, bias4.z, bias4.w};
        const uint64_t row0 = blockrow + br;
        if (a.list) {
          const float4 blo4 = *(const float4 *)&s_blo[br];
          const float blov[4] = {blo4.x, blo4.y, blo4.z, blo4.w};
#pragma unroll
          for (int e4 = 0; e4 < 4; e4++) {
            const float av = acc[r][t][jg * 4 + e4];
            if (!(av * (qs * mulv[e4]) >= tau_lo + blov[e4])) continue;
            const float sv = (av * qs) * mulv[e4] - biasv[e4];
            const float sce = biasv[e4] == INFINITY ? -INFINITY : sv;
            if (!(sce > -INFINITY) || sce < tau) continue;
            const uint32_t pos = atomicAdd(&a.cnt[q], 1u);
            if (pos < a.cap) a.list[(uint64_t)q * a.cap + pos] = score_key(sce, row0 + e4);
          }
        } else {
          float sc[4];
#pragma unroll
          for (int e4 = 0; e4 < 4; e4++) {
            const float sv = (acc[r][t][jg * 4 + e4] * qs) * mulv[e4] - biasv[e4];
            sc[e4] = biasv[e4] == INFINITY ? -INFINITY : sv;
          }
          float *out = a.scores + (uint64_t)q * a.tile_rows + (row0 - a.tile0);
          if (row0 + 4 <= tile_end) {
            *(float4 *)out = make_float4(sc[0], sc[1], sc[2], sc[3]);
          } else {
#pragma unroll
            for (int e4 = 0; e4 < 4; e4++)
              if (row0 + e4 < tile_end) out[e4] = sc[e4];
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
}

template <int P, int WR, int RT, int NT>
__global__ __launch_bounds__(kGemmThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
void gemm_scan_kernel(const ScanArgs a) {
  constexpr int EB = Prec<P>::EB, KG = Prec<P>::KG, LB = Prec<P>::LB;
  constexpr int U = LB / 16;                // 16-byte units per lane per k group
  constexpr int R = 128 / (KG * EB);        // k groups per line
  constexpr int WQ = 8 / WR;
  static_assert(WR * RT * 32 == kGemmRows && WQ * NT * 32 == kGemmQueries, "block tile");
  __shared__ uint4 s_a0[kGemmRows * 8], s_q0[kGemmRows * 8], s_a1[kGemmRows * 8], s_q1[kGemmRows * 8];
  __shared__ __align__(16) float s_mul[2][kGemmRows];
  __shared__ __align__(16) float s_bias[2][kGemmRows];
  // the bias lowered by 4e-6 of its magnitude (the append pre-test below)
  __shared__ __align__(16) float s_blo[2][kGemmRows];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 31, h = lane >> 5;
  const int wr = wave / WQ, wq = wave % WQ;  // the wave's row slice and query slice
  const uint64_t rstride = (uint64_t)a.dimp * EB;
  const int nlines = (int)(rstride / 128);
  const uint8_t *xb = (const uint8_t *)a.x;
  const uint8_t *qb = (const uint8_t *)a.qm + (uint64_t)a.qoff * rstride;
  const uint64_t tile_end = min(a.n, a.tile0 + a.tile_rows);
  // DMA staging: wave w moves rows [32 w, 32 w + 32) of the rows and of
  // the queries, 8 rows per instruction; lane l writes LDS unit l of the
  // instruction's 64, i.e. slot l & 7 of row r = r0 + (l >> 3), which holds
  // unit (l & 7) ^ ((r >> 1) & 7) of the line; r0 = 32 w + 8 j, so
  // (r >> 1) & 7 = (4 j + (l >> 4)) & 7: two lane offsets, for even and odd j
  uint32_t lane_off[2];
#pragma unroll
  for (int jp = 0; jp < 2; jp++)
    lane_off[jp] = (uint32_t)((lane >> 3) * rstride) + (uint32_t)((((lane & 7) ^ ((4 * jp + (lane >> 4)) & 7))) * 16);
  auto dma_stage = [&](uint64_t blk, int line, uint4 *sa, uint4 *sq) {
    const uint64_t row_w = a.tile0 + blk * (uint64_t)kGemmRows + (uint64_t)wave * (kGemmRows / 8);
#pragma unroll
    for (int j = 0; j < kGemmRows / 8 / 8; j++) {
      const int r0 = wave * (kGemmRows / 8) + j * 8;
      const uint64_t row0 = row_w + (uint64_t)j * 8;
      const uint32_t lo = lane_off[j & 1];
      // rows past the tile read the tile's last row (dropped by its bias)
      const uint8_t *src = row0 + 8 <= tile_end ? xb + row0 * rstride + (uint64_t)line * 128 + lo
                                                 : xb + min(row0 + (uint64_t)(lane >> 3), tile_end - 1) * rstride +
                                                       (uint64_t)line * 128 + (lo % (uint32_t)rstride);
      __builtin_amdgcn_global_load_lds((const void *)src, (void *)(sa + r0 * 8), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void *)(qb + (uint64_t)r0 * rstride + (uint64_t)line * 128 + lo),
                                       (void *)(sq + r0 * 8), 16, 0, 0);
    }
  };
  float qsv[NT], tauv[NT];
  gemm_query_factors<P, NT>(a, wq, i, qsv, tauv);
  uint64_t blk = blockIdx.x;
  if (blk < a.nblocks) {
    gemm_fill_rows<P>(a, blk, tid, tile_end, s_mul[0], s_bias[0], s_blo[0]);
    dma_stage(blk, 0, s_a0, s_q0);
  }
  __syncthreads();
  int par = 0;
  for (; blk < a.nblocks; blk += gridDim.x, par ^= 1) {
    const uint64_t blockrow = a.tile0 + blk * (uint64_t)kGemmRows;
    f32x16 acc[RT][NT];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[r][t] = f32x16{};
    auto compute = [&](const uint4 *sa, const uint4 *sq) {
#pragma unroll 1
      for (int gg = 0; gg < R; gg++) {
        Frag<P> fa[RT];
#pragma unroll
        for (int r = 0; r < RT; r++) {
          const int row = wr * RT * 32 + r * 32 + i;
#pragma unroll
          for (int u = 0; u < U; u++) fa[r].v[u] = sa[row * 8 + ((h * 4 + gg * U + u) ^ ((row >> 1) & 7))];
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
          Frag<P> fq;
          const int qr = wq * NT * 32 + t * 32 + i;
#pragma unroll
          for (int u = 0; u < U; u++) fq.v[u] = sq[qr * 8 + ((h * 4 + gg * U + u) ^ ((qr >> 1) & 7))];
#pragma unroll
          for (int r = 0; r < RT; r++) mfma_group<P>(acc[r][t], fa[r], fq);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    // lines in pairs: even lines in buffer 0, odd lines in buffer 1 (the
    // host runs this kernel only for an even number of lines)
    for (int line = 0; line < nlines; line += 2) {
      dma_stage(blk, line + 1, s_a1, s_q1);
      compute(s_a0, s_q0);
      __syncthreads();
      __builtin_amdgcn_sched_barrier(0);
      const uint64_t nblk = line + 2 < nlines ? blk : blk + gridDim.x;
      if (nblk < a.nblocks) dma_stage(nblk, line + 2 < nlines ? line + 2 : 0, s_a0, s_q0);
      compute(s_a1, s_q1);
      __syncthreads();
    }
    gemm_epilogue<P, RT, NT>(a, blockrow, wr, wq, i, h, acc, qsv, tauv, s_mul[par], s_bias[par], s_blo[par]);
    const uint64_t nblk = blk + gridDim.x;
    if (nblk < a.nblocks) gemm_fill_rows<P>(a, nblk, tid, tile_end, s_mul[par ^ 1], s_bias[par ^ 1], s_blo[par ^ 1]);
  }
}

// ------------------------------------------------------------------ select

constexpr int kBuf = 4096;   // LDS candidate buffer (keys)
constexpr int kChunk = 2048; // keys read per round (8 per thread)

// Bitonic sort of s[0..N) in descending order by the whole block. Each
// stage's compare-exchanges of a thread (N / 2 / kBlock pairs) read all
// their pairs first, then write: one LDS round trip per stage instead of
// one per pair (the network, and so the result, is the plain bitonic one).
template <int N = kBuf>
__device__ void sort_desc(uint64_t *s) {
  constexpr int P = N / 2 / kBlock;
  static_assert(P >= 1 && P * 2 * kBlock == N, "sort size");
  for (int ls1 = 1; (1 << ls1) <= N; ls1++) {  // size = 2^ls1
    for (int ls = ls1 - 1; ls >= 0; ls--) {      // stride = 2^ls
      const int stride = 1 << ls;
      int lo[P];
      uint64_t x[P], y[P];
#pragma unroll
      for (int p = 0; p < P; p++) {
        const int t = threadIdx.x + p * kBlock;
        lo[p] = ((t >> ls) << (ls + 1)) | (t & (stride - 1));
        x[p] = s[lo[p]];
        y[p] = s[lo[p] + stride];
      }
#pragma unroll
      for (int p = 0; p < P; p++) {
        const bool desc = ((lo[p] >> ls1) & 1) == 0;
        if ((x[p] < y[p]) == desc) {
          s[lo[p]] = y[p];
          s[lo[p] + stride] = x[p];
        }
      }
      __syncthreads();
    }
  }
}

// Keeps the best C keys of the buffer: sorts, truncates, sets the threshold.
__device__ void compact(uint64_t *buf, int *count, uint64_t *tau, int C) {
  const int n = *count;
  for (int t = threadIdx.x; t < kBuf; t += kBlock)
    if (t >= n) buf[t] = 0;
  __syncthreads();
  if (n <= kBuf / 2)
    sort_desc<kBuf / 2>(buf);  // the zeros after the keys stay last
  else
    sort_desc(buf);
  if (threadIdx.x == 0) {
    const int keep = n < C ? n : C;
    *count = keep;
    *tau = keep == C ? buf[C - 1] : 0;
  }
  __syncthreads();
}

// Streams `len` keys through the buffer; leaves the best C (sorted,
// descending) in buf[0..C) with zeros after them. Raw values are fetched a
// chunk ahead: the next chunk's loads are issued before this chunk's keys
// are inserted, so a round does not wait for memory from a standing start
// (selection at 256 queries: FP8 2.76 -> 2.22 ms). `fetch(i)` reads value
// i, `conv(v, i)` makes its key (0 = no key).
template <class R, class Fetch, class Conv>
__device__ void top_c_ahead(Fetch fetch, Conv conv, uint64_t len, int C, uint64_t *buf, int *count,
                            uint64_t *tau) {
  constexpr int KPT = kChunk / kBlock;
  if (threadIdx.x == 0) {
    *count = 0;
    *tau = 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1;
  R raw[KPT];
  auto load = [&](uint64_t base) {
#pragma unroll
    for (int j = 0; j < KPT; j++) {
      const uint64_t idx = base + (uint64_t)(j * kBlock + threadIdx.x);
      if (idx < len) raw[j] = fetch(idx);
    }
  };
  if (len) load(0);
  for (uint64_t base = 0; base < len; base += kChunk) {
    uint64_t keys[KPT];
#pragma unroll
    for (int j = 0; j < KPT; j++) {
      const uint64_t idx = base + (uint64_t)(j * kBlock + threadIdx.x);
      keys[j] = idx < len ? conv(raw[j], idx) : 0;
    }
    if (base + kChunk < len) load(base + kChunk);
    const uint64_t th = *tau;
    // wave-aggregated insertion: one LDS atomic per wave and step
#pragma unroll
    for (int j = 0; j < KPT; j++) {
      const bool pass = keys[j] > th;
      const uint64_t m = __ballot(pass);
      if (m == 0) continue;
      int at = 0;
      if (lane == 0) at = atomicAdd(count, __popcll(m));
      at = __shfl(at, 0, 64);
      if (pass) buf[at + __popcll(m & below)] = keys[j];
    }
    __syncthreads();
    const int cnt = *count;
    __syncthreads();  // every thread has read the count before more keys arrive
    if (cnt > kBuf - kChunk) compact(buf, count, tau, C);
  }
  compact(buf, count, tau, C);
}

// The same result as top_c_ahead in about two sorts instead of one per
// kBuf - kChunk keys kept: 2048 keys sampled at even strides are sorted,
// and the sample ranked at 2 C x (2048 / len) bounds the keys worth
// keeping (about 2 C of them pass); every key above the bound is gathered
// in one pass and the buffer sorted once. A bound that keeps fewer than C
// keys, or more than the buffer holds, falls back to top_c_ahead, so the
// kept keys are always the best C of all `len`.
template <class R, class Fetch, class Conv>
__device__ void top_c_sampled(Fetch fetch, Conv conv, uint64_t len, int C, uint64_t *buf, int *count,
                              uint64_t *tau) {
  constexpr int kSample = kBuf / 2;
  __shared__ int s_overflow;
  if (len <= (uint64_t)(4 * kSample)) {  // short inputs: the sample would be most of them
    top_c_ahead<R>(fetch, conv, len, C, buf, count, tau);
    return;
  }
  for (int t = threadIdx.x; t < kBuf; t += kBlock) {
    uint64_t key = 0;
    if (t < kSample) {
      const uint64_t idx = (uint64_t)t * len / kSample;
      key = conv(fetch(idx), idx);
    }
    buf[t] = key;
  }
  __syncthreads();
  sort_desc<kSample>(buf);  // the samples; the zeros after them stay last
  // about 2 C keys pass: a sample rank r keeps r len / 2048 +- sqrt(r) len /
  // 2048 keys, which stays within [C, kBuf] for C up to 1024 and every
  // segment length the host makes (at least 32 C rows)
  const int rank = min(kSample - 1, max(4, (int)((2ull * (uint64_t)C * kSample + len / 2) / len)));
  const uint64_t bound = buf[rank];
  __syncthreads();  // every thread has read the bound before the buffer is reused
  if (threadIdx.x == 0) {
    *count = 0;
    s_overflow = 0;
  }
  __syncthreads();
  if (bound != 0) {
    constexpr int KPT = kChunk / kBlock;
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1;
    // the next chunk's values load while this chunk's keys are placed
    R raw[KPT];
    auto load = [&](uint64_t base) {
#pragma unroll
      for (int j = 0; j < KPT; j++) {
        const uint64_t idx = base + (uint64_t)(j * kBlock + threadIdx.x);
        if (idx < len) raw[j] = fetch(idx);
      }
    };
    load(0);
    for (uint64_t base = 0; base < len; base += kChunk) {
      uint64_t keys[KPT];
#pragma unroll
      for (int j = 0; j < KPT; j++) {
        const uint64_t idx = base + (uint64_t)(j * kBlock + threadIdx.x);
        keys[j] = idx < len ? conv(raw[j], idx) : 0;
      }
      if (base + kChunk < len) load(base + kChunk);
#pragma unroll
      for (int j = 0; j < KPT; j++) {
        const bool pass = keys[j] > bound;
        const uint64_t m = __ballot(pass);
        if (m == 0) continue;
        int at = 0;
        if (lane == 0) at = atomicAdd(count, __popcll(m));
        at = __shfl(at, 0, 64);
        const int pos = at + __popcll(m & below);
        if (pass) {
          if (pos < kBuf) buf[pos] = keys[j];
          else s_overflow = 1;
        }
      }
    }
  }
  __syncthreads();
  const int cnt = *count;
  const bool fallback = bound == 0 || s_overflow != 0 || cnt < C;
  __syncthreads();  // every thread has read the count and flag
  if (fallback) {
    top_c_ahead<R>(fetch, conv, len, C, buf, count, tau);
    return;
  }
  compact(buf, count, tau, C);
}

struct SelectArgs {
  const float *scores;  // [nq][tile_rows]
  uint64_t tile0, tile_len, tile_rows, seg_rows;
  int32_t nq, C, seg_first, segs_total;
  uint64_t *cand;       // [nq][segs_total][C]
  int32_t sampled;      // top_c_sampled for C at least this (knob vector.select_sampled), else top_c_ahead
};
__device__ inline bool tune_get_select_sampled(const SelectArgs &a) { return a.sampled > 0; }



// grid (segments of this tile, queries)
__global__ __launch_bounds__(kBlock) void select_tile_kernel(const SelectArgs a) {
  __shared__ uint64_t s_buf[kBuf];
  __shared__ int s_count;
  __shared__ uint64_t s_tau;
  const int q = blockIdx.y;
  const uint64_t lo = (uint64_t)blockIdx.x * a.seg_rows;
  if (lo >= a.tile_len) return;
  const uint64_t len = min(a.seg_rows, a.tile_len - lo);
  const float *sp = a.scores + (uint64_t)q * a.tile_rows + lo;
  const uint64_t row0 = a.tile0 + lo;
  auto fetch = [&](uint64_t idx) -> float { return sp[idx]; };
  auto conv = [&](float s, uint64_t idx) -> uint64_t { return s > -INFINITY ? score_key(s, row0 + idx) : 0; };
  if (tune_get_select_sampled(a) && a.C >= a.sampled)  // few candidates: the streaming buffer rarely sorts
    top_c_sampled<float>(fetch, conv, len, a.C, s_buf, &s_count, &s_tau);
  else
    top_c_ahead<float>(fetch, conv, len, a.C, s_buf, &s_count, &s_tau);
  uint64_t *out = a.cand + ((uint64_t)q * a.segs_total + a.seg_first + blockIdx.x) * a.C;
  for (int j = threadIdx.x; j < a.C; j += kBlock) out[j] = s_buf[j];
}

// grid (queries): the best C of every segment's C
__global__ __launch_bounds__(kBlock) void merge_kernel(uint64_t *cand, int segs_total, int C,
                                                       uint64_t *best) {
  __shared__ uint64_t s_buf[kBuf];
  __shared__ int s_count;
  __shared__ uint64_t s_tau;
  const int q = blockIdx.x;
  const uint64_t *in = cand + (uint64_t)q * segs_total * C;
  top_c_ahead<uint64_t>([&](uint64_t idx) -> uint64_t { return in[idx]; },
                        [](uint64_t v, uint64_t) -> uint64_t { return v; }, (uint64_t)segs_total * C, C, s_buf,
                        &s_count, &s_tau);
  for (int j = threadIdx.x; j < C; j += kBlock) best[(uint64_t)q * C + j] = s_buf[j];
}

// grid (queries): the best C of the pilot's best C and the query's
// appended keys
__global__ __launch_bounds__(kBlock) void merge_lists_kernel(const uint64_t *pilot_best, const uint64_t *list,
                                                             const uint32_t *cnt, uint64_t cap, int C,
                                                             uint64_t *best) {
  __shared__ uint64_t s_buf[kBuf];
  __shared__ int s_count;
  __shared__ uint64_t s_tau;
  const int q = blockIdx.x;
  const uint64_t *pb = pilot_best + (uint64_t)q * C;
  const uint64_t *ls = list + (uint64_t)q * cap;
  const uint64_t len = (uint64_t)C + min((uint64_t)cnt[q], cap);
  top_c_ahead<uint64_t>([&](uint64_t idx) -> uint64_t { return idx < (uint64_t)C ? pb[idx] : ls[idx - C]; },
                        [](uint64_t v, uint64_t) -> uint64_t { return v; }, len, C, s_buf, &s_count, &s_tau);
  for (int j = threadIdx.x; j < C; j += kBlock) best[(uint64_t)q * C + j] = s_buf[j];
}

// grid (queries): exact FP32 distances of the C candidates, sorted by
// (distance, row); the first k out.
struct RerankArgs {
  const float *x;
  int32_t dim, metric, C, Cpad, k;
  const float *queries;   // [nq][dim] FP32
  const uint64_t *best;   // [nq][C]
  int64_t *out_rows;      // [nq][k]
  float *out_dist;
};

__global__ __launch_bounds__(kBlock) void rerank_kernel(const RerankArgs a) {
  extern __shared__ uint64_t s_dyn[];
  uint64_t *s_key = s_dyn;                      // [Cpad]
  float *s_q = (float *)(s_dyn + a.Cpad);       // [dim]
  const int q = blockIdx.x;
  const float *qv = a.queries + (uint64_t)q * a.dim;
  for (int c = threadIdx.x; c < a.dim; c += kBlock) s_q[c] = qv[c];
  for (int j = threadIdx.x; j < a.Cpad; j += kBlock) s_key[j] = ~0ull;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float qn = 0.f;
  for (int c = lane; c < a.dim; c += 64) qn = __fmaf_rn(s_q[c], s_q[c], qn);
  const float qnorm2 = wave_sum(qn);
  // each wave takes four candidates at a time, their row loads in flight
  // together (one candidate's row at a time left the kernel waiting on
  // memory); an empty slot (key 0) reads row 0 and is not kept
  constexpr int R = 8;
  for (int j0 = wave * R; j0 < a.C; j0 += (kBlock / 64) * R) {
    uint64_t rows[R];
    bool use[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int j = j0 + r;
      const uint64_t key = j < a.C ? a.best[(uint64_t)q * a.C + j] : 0;
      use[r] = key != 0;
      rows[r] = use[r] ? key_row(key) : 0;
    }
    float s0[R], s1[R];
    row_partials_multi<R>(a.x, rows, s_q, a.dim, a.metric, lane, s0, s1);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const float d = finish_distance(a.metric, wave_sum(s0[r]), wave_sum(s1[r]), qnorm2);
      if (use[r] && lane == 0 && !__builtin_isnan(d)) s_key[j0 + r] = ((uint64_t)ordered(d) << 32) | rows[r];
    }
  }
  __syncthreads();
  // ascending bitonic sort of Cpad keys (a power of two)
  for (int size = 2; size <= a.Cpad; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < a.Cpad / 2; t += kBlock) {
        const int lo = 2 * stride * (t / stride) + (t % stride);
        const int hi = lo + stride;
        const bool asc = (lo & size) == 0;
        const uint64_t x = s_key[lo], y = s_key[hi];
        if ((x > y) == asc) {
          s_key[lo] = y;
          s_key[hi] = x;
        }
      }
      __syncthreads();
    }
  }
  for (int j = threadIdx.x; j < a.k; j += kBlock) {
    const uint64_t key = s_key[j];
    int64_t row = -1;
    float d = __builtin_nanf("");
    if (key != ~0ull) {
      row = (int64_t)(uint32_t)key;
      const uint32_t u = (uint32_t)(key >> 32);
      const uint32_t bits = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
      __builtin_memcpy(&d, &bits, 4);
    }
    a.out_rows[(uint64_t)q * a.k + j] = row;
    a.out_dist[(uint64_t)q * a.k + j] = d;
  }
}

// ------------------------------------------------------------------- host
int grid_for(oxr_device *d, uint64_t units, int per_cu) {
  const uint64_t cap = (uint64_t)std::max(1, d->cu_count) * per_cu;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(units, cap));
}

// How a chunk of queries is scanned: per wave RT row tiles x NT query
// tiles, QW query groups per block, the block's query tile resident in LDS
// (bres) or streamed; qpl queries per launch.
struct ScanCfg {
  int nt, rt, qw;
  bool bres;
  int qpl;
  bool gemm = false;  // gemm_scan_kernel: 256 rows x 256 queries per block
  uint64_t rpb() const { return gemm ? (uint64_t)kGemmRows : (uint64_t)(4 / qw) * rt * 32; }
};

// A block's query tile of at most this many bytes (every k group) can sit
// in LDS for the whole scan.
constexpr size_t kResidentQueryBytes = 100 * 1024;

ScanCfg pick_scan(int prec, int nqc, int dimp) {
  const size_t eb = prec == P_FP8 ? 1 : prec == P_BF16 ? 2 : 4;
  const bool res = tune_get("vector.resident_queries", 1) > 0;
  // a query tile of up to kResidentQueryBytes stays in LDS (FP32, 32
  // queries over 768 dims: 96 KB, one block per CU: 14.4 ms a scan against
  // 18.5 streamed)
  auto fits = [&](int queries, size_t lim) { return res && (size_t)queries * dimp * eb <= lim; };
  // small batches: one query tile; fewer row tiles per wave leave
  // registers for more waves per SIMD (one query: FP8 3.15 ms with two
  // against 3.64 with four and 3.23 with one; BF16 6.89 with one against
  // 7.28 with two and 7.84 with four)
  if (nqc <= 32) return {1, prec == P_FP32 ? 4 : prec == P_BF16 ? 1 : 2, 1, fits(32, kResidentQueryBytes), 32};
  // (a 64-query tile resident at 96 KB measured slower than streamed)
  if (nqc <= 64) return {2, 2, 1, fits(64, 64 * 1024), 64};
  // large batches: the GEMM tile (256 rows x 256 queries per block, rows
  // and queries staged through LDS a line at a time), knob vector.gemm
  if (nqc > 128 && tune_get("vector.gemm", 1) > 0 && ((size_t)dimp * eb / 128) % 2 == 0)
    return ScanCfg{8, 4, 1, false, kGemmQueries, true};
  // large FP8 batches: the block's 128 queries resident (96 KB), each wave
  // 64 rows x 128 queries, so every row is loaded by one wave (13.6 ms at
  // 256 queries over 20M x 768). vector.wide = 2: two query groups of 64
  // per block sharing each row's loads through L1 (16.2 ms: the duplicate
  // loads fill the L1's 64 bytes a clock).
  const int64_t wide = tune_get("vector.wide", 1);
  if (prec == P_FP8 && res && wide > 0 && (size_t)128 * dimp * eb <= kResidentQueryBytes)
    return wide == 2 ? ScanCfg{2, 4, 2, true, 128} : ScanCfg{4, 2, 1, true, 128};
  // other large batches: the streamed 4-tile kernel in launches of 128
  // queries (256 queries over 20M x 768: BF16 22.3 ms, FP32 76 ms). The
  // 8-tile kernel in one launch of 256 spills registers (BF16 40.5, FP32
  // 116 ms; vector.nt8 = 1); BF16 with 64 queries resident per launch, 4
  // row tiles per wave and a six-deep load rotation reads the rows twice
  // as often (33.2 ms; vector.wide = 3).
  if (prec == P_BF16 && res && wide == 3 && (size_t)64 * dimp * eb <= kResidentQueryBytes)
    return {2, 4, 1, true, 64};
  if (nqc > 128 && tune_get("vector.nt8", 0) > 0) return {8, 1, 1, false, 256};
  return {4, 1, 1, false, 128};
}

template <int P, int NT, int RT, bool BRES, int QW>
void launch_scan_cfg(oxr_device *dev, uint64_t units, hipStream_t s, const ScanArgs &a) {
  if (BRES) {
    const size_t bres = (size_t)QW * NT * 32 * a.dimp * Prec<P>::EB;
    // as many blocks per CU as LDS holds (a block keeps its query tile)
    const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / (bres + 8 * 1024))));
    hipLaunchKernelGGL((scan_kernel<P, NT, RT, BRES, QW>), dim3(grid_for(dev, units, per_cu)), dim3(kBlock), bres,
                       s, a);
  } else {
    hipLaunchKernelGGL((scan_kernel<P, NT, RT, BRES, QW>), dim3(grid_for(dev, units, NT >= 4 ? 2 : 8)),
                       dim3(kBlock), 0, s, a);
  }
}

template <int P>
void launch_scan_p(const ScanCfg &c, oxr_device *dev, uint64_t units, hipStream_t s, const ScanArgs &a) {
  if (c.gemm) {
    // one block per CU: 128 KB of stage buffers
    hipLaunchKernelGGL((gemm_scan_kernel<P, 4, 2, 4>), dim3(grid_for(dev, units, 1)), dim3(kGemmThreads), 0, s, a);
    return;
  }
  if (P == P_FP8 && c.qw == 2) {
    launch_scan_cfg<P, 2, 4, true, 2>(dev, units, s, a);
    return;
  }
  if (P == P_FP8 && c.nt == 4 && c.rt == 2) {
    launch_scan_cfg<P, 4, 2, true, 1>(dev, units, s, a);
    return;
  }
  if (P != P_FP32 && c.nt == 1 && c.rt == 2 && c.bres) {
    launch_scan_cfg<P, 1, 2, true, 1>(dev, units, s, a);
    return;
  }
  if (P != P_FP32 && c.nt == 1 && c.rt == 1 && c.bres) {
    launch_scan_cfg<P, 1, 1, true, 1>(dev, units, s, a);
    return;
  }
  if (P == P_BF16 && c.nt == 2 && c.rt == 4) {
    launch_scan_cfg<P, 2, 4, true, 1>(dev, units, s, a);
    return;
  }
  switch (c.nt) {
    case 1:
      if (c.bres) launch_scan_cfg<P, 1, 4, true, 1>(dev, units, s, a);
      else launch_scan_cfg<P, 1, 4, false, 1>(dev, units, s, a);
      break;
    case 2:
      if (c.bres) launch_scan_cfg<P, 2, 2, true, 1>(dev, units, s, a);
      else launch_scan_cfg<P, 2, 2, false, 1>(dev, units, s, a);
      break;
    case 4: launch_scan_cfg<P, 4, 1, false, 1>(dev, units, s, a); break;
    default: launch_scan_cfg<P, 8, 1, false, 1>(dev, units, s, a); break;
  }
}

// Every query launch of a chunk over the rows `base` describes.
void launch_scan(int prec, const ScanCfg &c, int nqc, oxr_device *dev, uint64_t units, hipStream_t s,
                 const ScanArgs &base) {
  for (int q0 = 0; q0 < nqc; q0 += c.qpl) {
    ScanArgs a = base;
    a.qoff = q0;
    if (prec == P_FP32) launch_scan_p<P_FP32>(c, dev, units, s, a);
    else if (prec == P_BF16) launch_scan_p<P_BF16>(c, dev, units, s, a);
    else launch_scan_p<P_FP8>(c, dev, units, s, a);
  }
}

// Rows are padded to whole 128-byte lines of the scan precision (the
// scan's k permutation works a line at a time).
int pad_dim(int dim, int prec) {
  const int m = prec == P_FP8 ? 128 : prec == P_BF16 ? 64 : 32;
  return (dim + m - 1) / m * m;
}

// Scratch freed on the stream when the scope ends.
struct Scratch {
  oxr_stream *s;
  std::vector<void *> ptrs;
  explicit Scratch(oxr_stream *st) : s(st) {}
  ~Scratch() {
    for (void *p : ptrs) stream_release(s, p);
  }
  template <class T>
  int get(size_t bytes, T **out) {
    void *p = nullptr;
    OXR_TRY(stream_alloc(s, std::max<size_t>(bytes, 256), &p));
    ptrs.push_back(p);
    *out = (T *)p;
    return OXR_OK;
  }
};

int check_vector_column(const oxr_batch *input, int32_t column, const char *feature) {
  if (!input) return fail(OXR_EPLAN, feature, "null batch");
  if (column < 0 || column >= (int)input->cols.size())
    return fail(OXR_EPLAN, feature, "column index out of range");
  const DevCol &c = input->cols[column];
  if (c.type != OXR_F32_VECTOR)
    return fail(OXR_EPLAN, feature, "column is not a FixedSizeList<Float32> vector");
  if (c.dim <= 0 || c.dim > kMaxDim)
    return fail(OXR_EUNSUPPORTED, feature, "vector size " + std::to_string(c.dim) + " is above " +
                                               std::to_string(kMaxDim));
  return OXR_OK;
}

int search(oxr_stream *s, const oxr_vector_index *ix, const float *queries, int32_t nq, int32_t k,
           int32_t C, int64_t *out_rows, float *out_dist) {
  if (nq <= 0) return OXR_OK;
  if (!queries || !out_rows || !out_dist) return fail(OXR_EPLAN, "rocm.vector", "null argument");
  if (k <= 0) return fail(OXR_EPLAN, "rocm.vector", "k must be positive");
  if (C < k) C = k;
  if (C > kMaxCandidates)
    return fail(OXR_EUNSUPPORTED, "rocm.vector",
                "at most " + std::to_string(kMaxCandidates) + " candidates per query");
  if (ix->n == 0) {  // no rows: every answer is missing
    for (int64_t j = 0; j < (int64_t)nq * k; j++) {
      out_rows[j] = -1;
      out_dist[j] = NAN;
    }
    return OXR_OK;
  }
  hipStream_t st = s->s;
  Scratch sc(s);
  const size_t qbytes = (size_t)nq * ix->dim * 4;
  float *dq;
  OXR_TRY(sc.get(qbytes, &dq));
  OXR_HIP(hipMemcpyAsync(dq, queries, qbytes, hipMemcpyHostToDevice, st));
  int Cpad = 64;  // a power of two; >= 2 keeps the query after the keys 16-byte aligned
  while (Cpad < C) Cpad <<= 1;
  int64_t *drows;
  float *ddist;
  OXR_TRY(sc.get((size_t)nq * k * 8, &drows));
  OXR_TRY(sc.get((size_t)nq * k * 4, &ddist));
  const int eb = ix->precision == P_FP8 ? 1 : ix->precision == P_BF16 ? 2 : 4;
  const void *xscan = ix->scan ? ix->scan : (const void *)ix->f32;

  for (int q0 = 0; q0 < nq; q0 += kMaxScanQueries) {
    const int nqc = std::min(kMaxScanQueries, nq - q0);
    const ScanCfg cfg = pick_scan(ix->precision, nqc, ix->dimp);
    const int qpad = (nqc + cfg.qpl - 1) / cfg.qpl * cfg.qpl;
    void *qm;
    float *qscale;
    OXR_TRY(sc.get((size_t)qpad * ix->dimp * eb, &qm));
    OXR_TRY(sc.get((size_t)qpad * 4, &qscale));
    PrepArgs pa{dq + (size_t)q0 * ix->dim, nqc, qpad, ix->dim, ix->dimp, ix->precision, qm, qscale};
    hipLaunchKernelGGL(prep_kernel, dim3(qpad), dim3(kBlock), 0, st, pa);
    OXR_HIP(hipGetLastError());

    const uint64_t rpb = cfg.rpb();
    ScanArgs base{};
    base.x = xscan;
    base.qm = qm;
    base.dimp = ix->dimp;
    base.groups = ix->dimp / (ix->precision == P_FP8 ? 64 : ix->precision == P_BF16 ? 16 : 8);
    base.nq = nqc;
    base.metric = ix->metric;
    base.rowinfo = ix->rowinfo;
    base.rscale = ix->rscale;
    base.qscale = qscale;
    base.valid = ix->valid;
    base.C = C;

    // The best C keys per query of rows [begin, end), into out ([nqc][C]):
    // score tiles of at most ~2 GB, each selected in segments, then merged.
    auto scan_select = [&](uint64_t begin, uint64_t end, uint64_t *out) -> int {
      const uint64_t rows = end - begin;
      uint64_t tile = std::max<uint64_t>(rpb, ((2ull << 30) / (4ull * nqc)) / rpb * rpb);
      if (tile > rows) tile = (rows + rpb - 1) / rpb * rpb;
      const uint64_t ntiles = (rows + tile - 1) / tile;
      // segments: about 2048 select blocks per tile, each at least 8K rows
      // and 32 C (each keeps its best C, and one block per query merges
      // them all: at C = 1000 over 1M rows, 8K-row segments left 125K keys
      // to merge, 0.61 ms of a 5.3 ms one-query search)
      const uint64_t want = std::max<uint64_t>(1, 2048 / (uint64_t)nqc);
      const uint64_t seg_rows =
          std::max<uint64_t>(std::max<uint64_t>(8192, 32ull * (uint64_t)C), (tile + want - 1) / want);
      const uint64_t segs_per_tile = (tile + seg_rows - 1) / seg_rows;
      const uint64_t segs_total = segs_per_tile * ntiles;
      float *scores;
      uint64_t *cand;
      OXR_TRY(sc.get((size_t)nqc * tile * 4, &scores));
      OXR_TRY(sc.get((size_t)nqc * segs_total * C * 8, &cand));
      // segments past the end of the last tile stay empty
      OXR_TRY(fill_zero(s, cand, (size_t)nqc * segs_total * C * 8));
      for (uint64_t ti = 0; ti < ntiles; ti++) {
        const uint64_t t0 = begin + ti * tile;
        const uint64_t tlen = std::min(tile, end - t0);
        ScanArgs a = base;
        a.n = end;
        a.tile0 = t0;
        a.tile_rows = tile;
        a.nblocks = (tlen + rpb - 1) / rpb;
        a.scores = scores;
        {
          ProfScope prof(st, "vector_scan");
          launch_scan(ix->precision, cfg, nqc, ix->dev, a.nblocks, st, a);
          OXR_HIP(hipGetLastError());
        }
        SelectArgs sa{scores, t0, tlen, tile, seg_rows, nqc, C, (int)(ti * segs_per_tile), (int)segs_total, cand,
                      (int32_t)tune_get("vector.select_sampled", 128)};
        const uint64_t segs_here = (tlen + seg_rows - 1) / seg_rows;
        {
          ProfScope prof(st, "vector_select");
          hipLaunchKernelGGL(select_tile_kernel, dim3((unsigned)segs_here, nqc), dim3(kBlock), 0, st, sa);
          OXR_HIP(hipGetLastError());
        }
      }
      ProfScope prof(st, "vector_merge");
      hipLaunchKernelGGL(merge_kernel, dim3(nqc), dim3(kBlock), 0, st, cand, (int)segs_total, C, out);
      OXR_HIP(hipGetLastError());
      return OXR_OK;
    };

    uint64_t *best;
    OXR_TRY(sc.get((size_t)nqc * C * 8, &best));
    // Pilot (knob vector.pilot): the best C of the first rows bound each
    // query's final C-th best from below, so the other rows need no score
    // array: the scan appends only the keys at or above the bound. A list
    // that overflows sends the batch down the plain path. About 2048 C
    // rows, at least 1M: a looser bound floods the scan with appends (at
    // C = 1000 over 20M rows, 256 queries: 256K-row pilot 44 ms of scan,
    // 1M 13.7, 2M 12.4), a larger pilot costs more selection (4M: 5.5 ms
    // of select against 2.8).
    const int64_t pilot_knob = tune_get("vector.pilot_rows", -1);
    const uint64_t pilot_want = pilot_knob > 0 ? (uint64_t)pilot_knob : std::max<uint64_t>(1ull << 20, (uint64_t)C * 2048);
    const uint64_t pilot = std::min<uint64_t>(ix->n, std::max<uint64_t>(rpb, pilot_want / rpb * rpb));
    bool done = false;
    if (tune_get("vector.pilot", 1) > 0 && ix->n >= 4 * pilot) {
      uint64_t *pbest, *list;
      uint32_t *cnt;
      const uint64_t rest = ix->n - pilot;
      const int64_t cap_knob = tune_get("vector.pilot_cap", -1);  // tests force the overflow path with it
      const uint64_t cap = cap_knob > 0 ? (uint64_t)cap_knob
                                        : std::max<uint64_t>(65536, 4 * (uint64_t)C * ((rest + pilot - 1) / pilot));
      OXR_TRY(sc.get((size_t)nqc * C * 8, &pbest));
      OXR_TRY(sc.get((size_t)nqc * cap * 8, &list));
      OXR_TRY(sc.get((size_t)nqc * 4, &cnt));
      OXR_TRY(fill_zero(s, cnt, (size_t)nqc * 4));
      OXR_TRY(scan_select(0, pilot, pbest));
      ScanArgs a = base;
      a.n = ix->n;
      a.tile0 = pilot;
      a.tile_rows = (rest + rpb - 1) / rpb * rpb;
      a.nblocks = a.tile_rows / rpb;
      a.tau_keys = pbest;
      a.list = list;
      a.cnt = cnt;
      a.cap = cap;
      {
        ProfScope prof(st, "vector_scan");
        launch_scan(ix->precision, cfg, nqc, ix->dev, a.nblocks, st, a);
        OXR_HIP(hipGetLastError());
      }
      uint32_t *hcnt = nullptr;  // the list lengths, read back through pinned memory
      OXR_TRY(stream_pinned(s, (size_t)nqc * 4, (void **)&hcnt));
      OXR_HIP(hipMemcpyAsync(hcnt, cnt, (size_t)nqc * 4, hipMemcpyDeviceToHost, st));
      OXR_HIP(hipStreamSynchronize(st));
      bool overflow = false;
      for (int c = 0; c < nqc; c++) overflow |= hcnt[c] > cap;
      if (!overflow) {
        ProfScope prof(st, "vector_merge");
        hipLaunchKernelGGL(merge_lists_kernel, dim3(nqc), dim3(kBlock), 0, st, pbest, list, cnt, cap, C, best);
        OXR_HIP(hipGetLastError());
        done = true;
      }
    }
    if (!done) OXR_TRY(scan_select(0, ix->n, best));
    RerankArgs ra{ix->f32, ix->dim, ix->metric, C, Cpad, k, dq + (size_t)q0 * ix->dim, best,
                  drows + (size_t)q0 * k, ddist + (size_t)q0 * k};
    const size_t shm = (size_t)Cpad * 8 + (size_t)ix->dim * 4;
    {
      ProfScope prof(st, "vector_rerank");
      hipLaunchKernelGGL(rerank_kernel, dim3(nqc), dim3(kBlock), shm, st, ra);
      OXR_HIP(hipGetLastError());
    }
  }
  OXR_HIP(hipMemcpyAsync(out_rows, drows, (size_t)nq * k * 8, hipMemcpyDeviceToHost, st));
  OXR_HIP(hipMemcpyAsync(out_dist, ddist, (size_t)nq * k * 4, hipMemcpyDeviceToHost, st));
  OXR_HIP(hipStreamSynchronize(st));
  return OXR_OK;
}

int make_index(oxr_stream *s, const oxr_batch *input, int32_t column, int32_t metric, int32_t precision,
               oxr_vector_index **out) {
  OXR_TRY(check_vector_column(input, column, "rocm.vector"));
  if (metric != OXR_L2 && metric != OXR_COSINE) return fail(OXR_EPLAN, "rocm.vector", "unknown metric");
  if (precision != P_FP32 && precision != P_BF16 && precision != P_FP8)
    return fail(OXR_EPLAN, "rocm.vector", "unknown precision");
  if (input->rows >= 0xFFFFFFFFull)
    return fail(OXR_EUNSUPPORTED, "rocm.vector", "at most 2^32 - 2 rows per index");
  const DevCol &c = input->cols[column];
  std::unique_ptr<oxr_vector_index> ix(new oxr_vector_index());
  ix->dev = s->dev;
  ix->n = input->rows;
  ix->dim = c.dim;
  ix->dimp = pad_dim(c.dim, precision);
  ix->metric = metric;
  ix->precision = precision;
  ix->f32 = (const float *)c.values;
  ix->valid = (const uint32_t *)c.validity;
  ix->holder = share_columns(input);
  ix->owner = retain(s->core);
  auto cleanup = [&]() {
    stream_release(s, ix->scan_owned ? ix->scan : nullptr);
    stream_release(s, ix->rowinfo);
    stream_release(s, ix->rscale);
  };
  const uint64_t n = std::max<uint64_t>(ix->n, 1);
  int rc = stream_alloc(s, n * 4, (void **)&ix->rowinfo);
  if (rc == OXR_OK && precision == P_FP8) rc = stream_alloc(s, n * 4, (void **)&ix->rscale);
  // FP32 rows are scanned in place unless they need padding
  const bool copy = precision != P_FP32 || ix->dimp != ix->dim;
  if (rc == OXR_OK && copy) {
    const int eb = precision == P_FP8 ? 1 : precision == P_BF16 ? 2 : 4;
    rc = stream_alloc(s, n * (uint64_t)ix->dimp * eb, &ix->scan);
    ix->scan_owned = rc == OXR_OK;
  }
  if (rc == OXR_OK && ix->n) {
    BuildArgs b{ix->f32, ix->n, ix->dim, ix->dimp, metric, precision, copy ? ix->scan : nullptr, ix->rowinfo,
                ix->rscale};
    ProfScope prof(s->s, "vector_build");
    hipLaunchKernelGGL(build_kernel, dim3(grid_for(s->dev, (ix->n + 3) / 4, 16)), dim3(kBlock), 0, s->s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail_hip(e, "vector build kernel");
  }
  if (rc != OXR_OK) {
    cleanup();
    release(ix->owner);
    return rc;
  }
  *out = ix.release();
  return OXR_OK;
}

}  // namespace
}  // namespace oxr

using namespace oxr;

extern "C" int oxr_vector_distance(oxr_stream *s, const oxr_batch *input, int32_t column,
                                   const float *query, int32_t dim, int32_t metric,
                                   const char *output_name, oxr_batch **out) {
  clear_error();
  if (!s || !out || !query) return fail(OXR_EPLAN, "rocm.vector", "null argument");
  OXR_TRY(check_vector_column(input, column, "rocm.vector"));
  const DevCol &c = input->cols[column];
  if (dim != c.dim)
    return fail(OXR_EPLAN, "rocm.vector",
                "query has " + std::to_string(dim) + " elements, the column " + std::to_string(c.dim));
  if (metric != OXR_L2 && metric != OXR_COSINE) return fail(OXR_EPLAN, "rocm.vector", "unknown metric");
  const uint64_t n = input->rows;
  std::unique_ptr<oxr_batch> ob(new_batch(s, n));
  ob->holders.push_back(share_columns(input));
  for (size_t i = 0; i < input->cols.size(); i++) {
    DevCol v = input->cols[i];
    v.owned = false;
    ob->cols.push_back(v);
    ob->names.push_back(input->names[i]);
  }
  DevCol d;
  OXR_TRY(alloc_column(s, OXR_FLOAT32, 0, n, c.validity != nullptr, &d));
  ob->cols.push_back(d);
  ob->names.push_back(output_name ? output_name : "distance");
  if (n == 0) {
    *out = ob.release();
    return OXR_OK;
  }
  float *dq = nullptr;
  OXR_TRY(stream_alloc(s, (size_t)dim * 4, (void **)&dq));
  hipError_t e = hipMemcpyAsync(dq, query, (size_t)dim * 4, hipMemcpyHostToDevice, s->s);
  if (e == hipSuccess && c.validity)
    e = hipMemcpyAsync(d.validity, c.validity, validity_bytes(n), hipMemcpyDeviceToDevice, s->s);
  if (e == hipSuccess) {
    DistArgs a{(const float *)c.values, (const uint32_t *)c.validity, n, dim, metric, dq, (float *)d.values};
    ProfScope prof(s->s, "vector_distance");
    hipLaunchKernelGGL(distance_kernel, dim3(grid_for(s->dev, (n + 3) / 4, 16)), dim3(kBlock),
                       (size_t)dim * 4, s->s, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->s);  // the host query may go away
  stream_release(s, dq);
  if (e != hipSuccess) {
    oxr_batch_free(ob.release());
    return fail_hip(e, "vector distance");
  }
  *out = ob.release();
  return OXR_OK;
}

extern "C" int oxr_vector_index_create(oxr_stream *s, const oxr_batch *input, int32_t column,
                                       int32_t metric, int32_t precision, oxr_vector_index **out) {
  clear_error();
  if (!s || !out) return fail(OXR_EPLAN, "rocm.vector", "null argument");
  return make_index(s, input, column, metric, precision, out);
}

extern "C" int oxr_vector_index_search(oxr_stream *s, const oxr_vector_index *index,
                                       const float *queries, int32_t nq, int32_t k, int32_t candidates,
                                       int64_t *out_rows, float *out_dist) {
  clear_error();
  if (!s || !index) return fail(OXR_EPLAN, "rocm.vector", "null argument");
  if (index->dev != s->dev) return fail(OXR_EDEVICE, "rocm.vector", "index belongs to another device");
  return search(s, index, queries, nq, k, candidates, out_rows, out_dist);
}

extern "C" void oxr_vector_index_free(oxr_vector_index *index) {
  if (!index) return;
  hipStream_t st = index->owner ? index->owner->s : nullptr;
  if (index->scan_owned) index->dev->arena->release(index->scan, st);
  index->dev->arena->release(index->rowinfo, st);
  index->dev->arena->release(index->rscale, st);
  release(index->owner);
  delete index;
}

extern "C" int oxr_vector_topk(oxr_stream *s, const oxr_batch *input, int32_t column, const float *queries,
                               int32_t nq, int32_t dim, int32_t metric, int32_t k, int64_t *out_rows,
                               float *out_dist) {
  clear_error();
  if (!s) return fail(OXR_EPLAN, "rocm.vector", "null argument");
  OXR_TRY(check_vector_column(input, column, "rocm.vector"));
  if (dim != input->cols[column].dim) return fail(OXR_EPLAN, "rocm.vector", "query size differs from the column's");
  oxr_vector_index *ix = nullptr;
  OXR_TRY(make_index(s, input, column, metric, P_FP32, &ix));
  const int rc = search(s, ix, queries, nq, k, k, out_rows, out_dist);
  oxr_vector_index_free(ix);
  return rc;
}
