// Synthetic tables generated directly in HBM.
//
// Counter-based: every value is a pure function of (seed, column, row), from
// SplitMix64 (the reference's table generator uses the same mixer,
// crates/oxidelake-storage/src/datagen.rs:89-95), so a table of any size is
// reproducible and a batch of rows can be regenerated independently of the
// rest. Distributions: uniform integers and reals, normal (Box-Muller),
// lognormal, Zipf (rejection-inversion, Hormann and Derflinger 1996), a
// sequence, a random permutation (a Feistel bijection with cycle walking) and
// Gaussian-mixture vectors (optionally L2-normalised).
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>

#include "common.h"

namespace oxr {
namespace {

__host__ __device__ inline uint64_t mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// The k-th 64-bit draw of (stream, row).
__device__ inline uint64_t draw(uint64_t stream, uint64_t row, uint32_t k) {
  return mix(mix(stream ^ (row * 0xD1B54A32D192ED03ull)) + k * 0x8CB92BA72F3D8DD7ull);
}

// Uniform double in [0, 1) with 53 random bits.
__device__ inline double unit(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

__device__ inline double normal(uint64_t stream, uint64_t row, uint32_t k) {
  // Box-Muller; u1 in (0, 1]
  const double u1 = 1.0 - unit(draw(stream, row, 2 * k));
  const double u2 = unit(draw(stream, row, 2 * k + 1));
  return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// Zipf over {1..n} with exponent s > 0 (s != 1 handled by the helper forms).
struct ZipfParams {
  double s, n, hx1, hn, skip;
};

__host__ __device__ inline double helper1(double x) {  // log1p(x)/x
  return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x));
}
__host__ __device__ inline double helper2(double x) {  // expm1(x)/x
  return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x));
}
__host__ __device__ inline double zh(double s, double x) { return exp(-s * log(x)); }
__host__ __device__ inline double zH(double s, double x) {
  const double lx = log(x);
  return helper2((1.0 - s) * lx) * lx;
}
__host__ __device__ inline double zHinv(double s, double x) {
  double t = x * (1.0 - s);
  if (t < -1.0) t = -1.0;
  return exp(helper1(t) * x);
}

ZipfParams zipf_params(double s, double n) {
  ZipfParams z;
  z.s = s;
  z.n = n;
  z.hx1 = zH(s, 1.5) - 1.0;
  z.hn = zH(s, n + 0.5);
  z.skip = 2.0 - zHinv(s, zH(s, 2.5) - zh(s, 2.0));
  return z;
}

__device__ inline int64_t zipf(const ZipfParams &z, uint64_t stream, uint64_t row) {
  for (uint32_t it = 0; it < 64; it++) {
    const double u = z.hn + unit(draw(stream, row, it)) * (z.hx1 - z.hn);
    const double x = zHinv(z.s, u);
    double k = floor(x + 0.5);
    if (k < 1) k = 1;
    else if (k > z.n) k = z.n;
    if (k - x <= z.skip || u >= zH(z.s, k + 0.5) - zh(z.s, k)) return (int64_t)k;
  }
  return 1;  // acceptance is near-certain per draw; 64 rejections in a row do not occur
}

// A bijection on [0, 4^half_bits) (balanced Feistel, 4 rounds); cycle-walked
// into [0, n) it is a permutation of 0..n-1.
__device__ inline uint64_t feistel(uint64_t x, int half_bits, uint64_t key) {
  const uint64_t mask = (1ull << half_bits) - 1;
  uint64_t l = x >> half_bits, r = x & mask;
  for (int i = 0; i < 4; i++) {
    const uint64_t f = mix(r ^ (key + i * 0x9E3779B97F4A7C15ull)) & mask;
    const uint64_t nl = r;
    r = l ^ f;
    l = nl;
  }
  return (l << half_bits) | r;
}

__device__ inline uint64_t permute(uint64_t row, uint64_t n, int half_bits, uint64_t key) {
  uint64_t x = row;
  do {
    x = feistel(x, half_bits, key);
  } while (x >= n);
  return x;
}

struct GenCol {
  void *values;
  uint32_t *valid;
  int32_t type, dist, dim;
  int32_t half_bits;
  double a, b, null_frac;
  uint64_t stream;
  ZipfParams z;
};

__global__ void gen_scalar_kernel(GenCol c, uint64_t n) {
  for (uint64_t row = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; row < n;
       row += (uint64_t)gridDim.x * blockDim.x) {
    double v = 0;
    int64_t iv = 0;
    bool is_int = false;
    switch (c.dist) {
      case OXR_GEN_UNIFORM_INT: {
        const uint64_t span = (uint64_t)((int64_t)c.b - (int64_t)c.a);
        iv = (int64_t)c.a + (int64_t)(span ? draw(c.stream, row, 0) % span : 0);
        is_int = true;
        break;
      }
      case OXR_GEN_UNIFORM_REAL: v = c.a + (c.b - c.a) * unit(draw(c.stream, row, 0)); break;
      case OXR_GEN_NORMAL: v = c.a + c.b * normal(c.stream, row, 0); break;
      case OXR_GEN_LOGNORMAL: v = exp(c.a + c.b * normal(c.stream, row, 0)); break;
      case OXR_GEN_ZIPF: iv = zipf(c.z, c.stream, row) - 1; is_int = true; break;
      case OXR_GEN_SEQUENCE: iv = (int64_t)c.a + (int64_t)row; is_int = true; break;
      case OXR_GEN_PERMUTATION:
        iv = (int64_t)c.a + (int64_t)permute(row, n, c.half_bits, c.stream);
        is_int = true;
        break;
    }
    switch (c.type) {
      case OXR_INT32: ((int32_t *)c.values)[row] = (int32_t)(is_int ? iv : (int64_t)v); break;
      case OXR_INT64: ((int64_t *)c.values)[row] = is_int ? iv : (int64_t)v; break;
      case OXR_FLOAT32: ((float *)c.values)[row] = (float)(is_int ? (double)iv : v); break;
      default: ((double *)c.values)[row] = is_int ? (double)iv : v; break;
    }
  }
}

// Validity: bit set with probability 1 - null_frac, one 32-row word per thread.
__global__ void gen_valid_kernel(GenCol c, uint64_t n) {
  const uint64_t words = (n + 31) / 32;
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < words;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t bits = 0;
    for (int i = 0; i < 32; i++) {
      const uint64_t row = w * 32 + i;
      if (row >= n) break;
      if (unit(draw(c.stream ^ 0x5A5A5A5A5A5A5A5Aull, row, 0)) >= c.null_frac) bits |= 1u << i;
    }
    c.valid[w] = bits;
  }
}

// Mixture vectors: one wave per row; centroid j's coordinate d is
// N(0,1) from (stream', j, d); point = centroid + b * N(0,1); optional
// L2 normalisation (a = centroid count, null_frac unused).
__global__ void gen_vector_kernel(GenCol c, uint64_t n, int normalize) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const int dim = c.dim;
  const uint64_t k = (uint64_t)c.a;
  const uint64_t cstream = c.stream ^ 0xC3A5C85C97CB3127ull;
  for (uint64_t row = wave; row < n; row += waves) {
    const uint64_t cid = k ? draw(c.stream, row, 0x7FFFFFFF) % k : 0;
    float *out = (float *)c.values + row * (uint64_t)dim;
    float ss = 0.f;
    for (int d = lane; d < dim; d += 64) {
      const double cen = k ? normal(cstream, cid * (uint64_t)dim + d, 0) : 0.0;
      const float x = (float)(cen + c.b * normal(c.stream, row * (uint64_t)dim + d, 1));
      out[d] = x;
      ss += x * x;
    }
    if (normalize) {
      for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
      const float inv = ss > 0.f ? 1.0f / sqrtf(ss) : 0.f;
      for (int d = lane; d < dim; d += 64) out[d] *= inv;
    }
  }
}

}  // namespace
}  // namespace oxr

using namespace oxr;

extern "C" int oxr_generate(oxr_stream *s, const oxr_gen_spec *cols, const char *const *names,
                            int32_t ncols, uint64_t rows, oxr_batch **out) {
  clear_error();
  if (!s || !out || (ncols > 0 && !cols)) return fail(OXR_EPLAN, "rocm.generate", "null argument");
  std::unique_ptr<oxr_batch> b(new_batch(s, rows));
  auto cleanup = [&]() {
    for (auto &c : b->cols) {
      stream_release(s, c.values);
      stream_release(s, c.validity);
    }
  };
  const unsigned blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((rows + 255) / 256, 16384));
  for (int32_t i = 0; i < ncols; i++) {
    const oxr_gen_spec &g = cols[i];
    if (g.type < OXR_INT32 || g.type > OXR_F32_VECTOR || (g.type == OXR_F32_VECTOR && g.dim <= 0)) {
      cleanup();
      return fail(OXR_EPLAN, "rocm.generate", "bad column type");
    }
    DevCol c;
    int rc = alloc_column(s, g.type, g.type == OXR_F32_VECTOR ? g.dim : 0, rows, g.null_frac > 0, &c);
    if (rc != OXR_OK) {
      cleanup();
      return rc;
    }
    b->cols.push_back(c);
    b->names.push_back(names && names[i] ? names[i] : "c" + std::to_string(i));
    GenCol gc{};
    gc.values = c.values;
    gc.valid = (uint32_t *)c.validity;
    gc.type = g.type;
    gc.dist = g.dist;
    gc.dim = g.dim;
    gc.a = g.a;
    gc.b = g.b;
    gc.null_frac = g.null_frac;
    gc.stream = mix(g.seed * 0x632BE59BD9B4E019ull + (uint64_t)i);
    if (g.dist == OXR_GEN_ZIPF) {
      if (!(g.a > 0) || g.a == 1.0 || !(g.b >= 1)) {
        cleanup();
        return fail(OXR_EPLAN, "rocm.generate", "Zipf needs an exponent > 0 other than 1 and n >= 1");
      }
      gc.z = zipf_params(g.a, std::floor(g.b));
    }
    if (g.dist == OXR_GEN_PERMUTATION) {
      int hb = 1;
      while ((1ull << (2 * hb)) < rows) hb++;
      gc.half_bits = hb;
    }
    if (rows) {
      if (g.type == OXR_F32_VECTOR) {
        hipLaunchKernelGGL(gen_vector_kernel, dim3(std::min<unsigned>((unsigned)((rows + 3) / 4), 65535u)),
                           dim3(256), 0, s->s, gc, rows, g.dist == OXR_GEN_MIXTURE_NORMALIZED ? 1 : 0);
      } else {
        hipLaunchKernelGGL(gen_scalar_kernel, dim3(blocks), dim3(256), 0, s->s, gc, rows);
      }
      if (g.null_frac > 0)
        hipLaunchKernelGGL(gen_valid_kernel, dim3(blocks), dim3(256), 0, s->s, gc, rows);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        cleanup();
        return fail_hip(e, "generator launch");
      }
    }
  }
  hipError_t e = hipStreamSynchronize(s->s);
  if (e != hipSuccess) {
    cleanup();
    return fail_hip(e, "generate");
  }
  *out = b.release();
  return OXR_OK;
}
