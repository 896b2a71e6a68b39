// Warm-start network on MI355X (gfx950): the C-ABI of include/fvs2d_mlp.h.
//
// Training step for a batch of M rows (M a multiple of 64), depth L hidden
// layers of width 256:
//   k_features       cell records -> bfloat16 input rows X0[M][128] (68 used)
//   k_gemm_nt<FWD>   Z = X W + b, H = GELU(Z), one launch per hidden layer
//   k_head           y = H W + b (float32), loss, dy, dZ of the last hidden
//                    layer, output-layer weight/bias gradient partials
//   k_gemm_tn        weight-gradient partials X^T dZ (and bias partials)
//   k_gemm_nt<BWD>   dZ_prev = (dZ W^T) * GELU'(Z_prev)
//   k_adam           sums the partials in a fixed order, Adam update, and
//                    refreshes the bfloat16 copies of the weights
// The GEMMs run on v_mfma_f32_16x16x32_bf16: a workgroup of 4 wavefronts
// owns 64 rows x all 256 columns (A read once), each wavefront a 64 x 64 tile
// of 4 x 4 MFMA blocks.  Operands are staged through LDS with the reduction
// index contiguous (rows padded to 40 elements, 80 bytes, so 16-byte
// fragment reads hit distinct banks).  The weight gradient reduces over the
// batch rows: each workgroup sums a contiguous slice of rows into one of S
// partial copies, the Adam kernel adds the S copies in slice order, so the
// result does not depend on scheduling (no float atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/fvs2d.h"
#include "../../include/fvs2d_mlp.h"

namespace fvs {
int set_error(int code, const std::string &msg);
}

namespace {

using u16 = unsigned short;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kW = FVS_MLP_WIDTH;
constexpr int kOut = FVS_MLP_OUT;
constexpr int kFeat = FVS_MLP_FEATURES;
constexpr int kRaw = FVS_MLP_RAW;
constexpr int kInPad = 128;
constexpr int kMaxDepth = 8;
constexpr int kMaxLayers = kMaxDepth + 1;
constexpr int kMaxSplit = 64;
constexpr int kBK = 32;        // reduction depth of one MFMA
constexpr int kLds = kBK + 8;  // padded LDS row (elements)
constexpr int kRows = 64;      // GEMM rows per workgroup

static_assert(kW == 256, "the GEMM tiling assumes width 256");

int fail(int code, const std::string &msg) { return fvs::set_error(code, msg); }
// an exception thrown by host code inside an entry point ends as an error
// code and message, never as a C++ exception crossing into the caller
int host_exception(const char *fn) {
  try {
    throw;
  } catch (const std::bad_alloc &) {
    return fail(FVS_ENOMEM, std::string(fn) + ": host memory exhausted");
  } catch (const std::exception &e) {
    return fail(FVS_ESTATE, std::string(fn) + ": " + e.what());
  } catch (...) {
    return fail(FVS_ESTATE, std::string(fn) + ": host exception");
  }
}

#define MLPC(x)                                                                           \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(FVS_EHIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

// float -> bfloat16, round to nearest even (NaN stays NaN)
__host__ __device__ inline u16 f2bf(float x) {
  unsigned u = __builtin_bit_cast(unsigned, x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u16(u >> 16);
}
__host__ __device__ inline float bf2f(u16 h) { return __builtin_bit_cast(float, unsigned(h) << 16); }

// erf by Abramowitz & Stegun 7.1.26 (Hastings' rational approximation,
// |error| <= 1.5e-7): branch-free, so the 64 evaluations per thread of a GEMM
// epilogue do not split the wavefront; far below bfloat16 resolution (3.9e-3)
__device__ inline float erf_as(float x) {
  const float ax = fabsf(x);
  const float t = 1.0f / (1.0f + 0.3275911f * ax);
  const float poly = ((((1.061405429f * t - 1.453152027f) * t + 1.421413741f) * t - 0.284496736f) * t +
                      0.254829592f) * t;
  return copysignf(1.0f - poly * __expf(-ax * ax), x);
}
__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erf_as(z * 0.70710678118654752f)) + z * __expf(-0.5f * z * z) * 0.39894228040143268f;
}

// splitmix64 finaliser (Steele, Lea, Flood 2014): row sampling and weight init
__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------- features
// 16 threads per row, thread q writes input columns 8q..8q+7 (16 bytes):
// q = 0..3 the radial harmonics 4q+1..4q+4, q = 4..7 the angular harmonics
// 4(q-4)+1..+4 (one sincos of the base angle and one of the first harmonic,
// then the angle-addition recurrence), q = 8 the four scalar channels.
// sample != 0: row m takes record hash(seed, step, m) mod ndata; otherwise
// record m, and rows m >= nvalid are zero.
__global__ __launch_bounds__(256) void k_features(int M, long nvalid, const float *__restrict__ raw,
                                                  const float *__restrict__ tgt, long ndata,
                                                  unsigned long long seed, long step, int sample,
                                                  u16 *__restrict__ X, float *__restrict__ T,
                                                  float *__restrict__ F32) {
  const long g = long(blockIdx.x) * 256 + threadIdx.x;
  const long m = g >> 4;
  const int q = int(g & 15);
  if (m >= M) return;
  bool valid;
  long src;
  if (sample) {
    valid = true;
    src = long(mix64(seed ^ mix64((unsigned long long)step * 0x100000001ull + (unsigned long long)m)) %
               (unsigned long long)ndata);
  } else {
    valid = m < nvalid;
    src = m;
  }
  float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (valid && q <= 8) {
    const float *r = raw + src * kRaw;
    if (q < 8) {
      const double dx = double(r[0]) - 0.5, y = r[1];
      double a;
      if (q < 4) {
        const double kPi = 3.14159265358979323846;
        a = kPi * (log1p(4.0 * sqrt(dx * dx + y * y)) / log1p(200.0));
      } else {
        a = atan2(y, dx);
      }
      const int k0 = 4 * (q & 3) + 1;
      double s1, c1, sk, ck;
      sincos(a, &s1, &c1);
      sincos(double(k0) * a, &sk, &ck);
      for (int h = 0; h < 4; ++h) {
        f[2 * h] = float(sk);
        f[2 * h + 1] = float(ck);
        const double sn = sk * c1 + ck * s1, cn = ck * c1 - sk * s1;
        sk = sn;
        ck = cn;
      }
    } else {
      f[0] = float(log1p(100.0 * double(r[2])) / log1p(5000.0));
      f[1] = float((double(r[3]) - 0.675) / 0.175);
      f[2] = float((double(r[4]) - 1.0) / 3.0);
      f[3] = float((double(r[5]) - 0.125) / 0.025);
    }
  }
  unsigned pk[4];
  for (int c = 0; c < 4; ++c) pk[c] = unsigned(f2bf(f[2 * c])) | (unsigned(f2bf(f[2 * c + 1])) << 16);
  *reinterpret_cast<uint4 *>(X + m * kInPad + 8 * q) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
  if (F32 && valid)
    for (int c = 0; c < 8; ++c)
      if (8 * q + c < kFeat) F32[m * kFeat + 8 * q + c] = f[c];
  if (T && q == 0)
    for (int o = 0; o < kOut; ++o) T[m * kOut + o] = valid ? tgt[src * kOut + o] : 0.0f;
}

// ------------------------------------------------------------- GEMM core
// One kBK slab of C[m][n] += sum_k A[m][k] B[n][k] with A_lds rows 16i..
// (4 blocks) and B_lds rows 64w + 16j.. (4 blocks).  Fragment layout of
// v_mfma_f32_16x16x32_bf16 for lane l: first operand row l%16, k
// 8(l/16)..+7; second operand column l%16, same k; result rows
// 4(l/16)..+3, column l%16.  The B fragment goes first, so the result is
// C transposed: lane l holds C[16i + l%16][64w + 16j + 4(l/16) + e], four
// consecutive columns, stored as one 8- or 16-byte vector.
__device__ __forceinline__ void mma_slab(const u16 *__restrict__ As, const u16 *__restrict__ Bs, int lane, int w,
                                         f32x4 (&acc)[4][4]) {
  bf16x8 a[4], b[4];
  const int r = lane & 15, kq = 8 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8 *>(As + (16 * i + r) * kLds + kq);
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const bf16x8 *>(Bs + (64 * w + 16 * j + r) * kLds + kq);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
}

__device__ __forceinline__ unsigned pack2(float a, float b) {
  return unsigned(f2bf(a)) | (unsigned(f2bf(b)) << 16);
}

// C[M][256] = A[M][K] * B[256][K]^T (both row-major, K a multiple of 32).
// EPI 0 (forward): z = C + bias; Z = bf16(z), H = bf16(GELU(z)).
// EPI 1 (backward data): dZ = bf16(C * GELU'(Zp)).
// The next slab's global data is loaded into registers while the current
// one is multiplied.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(int K, const u16 *__restrict__ A, const u16 *__restrict__ B,
                                                 const float *__restrict__ bias, const u16 *__restrict__ Zp,
                                                 u16 *__restrict__ O1, u16 *__restrict__ O2) {
  __shared__ __attribute__((aligned(16))) u16 As[kRows * kLds];
  __shared__ __attribute__((aligned(16))) u16 Bs[kW * kLds];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long r0 = long(blockIdx.x) * kRows;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int arow = tid >> 2, ac = (tid & 3) * 8;
  const u16 *Ap = A + (r0 + arow) * K + ac;
  uint4 ra = *reinterpret_cast<const uint4 *>(Ap), rb[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ch = tid + 256 * q;
    rb[q] = *reinterpret_cast<const uint4 *>(B + long(ch >> 2) * K + (ch & 3) * 8);
  }
  for (int k0 = 0; k0 < K; k0 += kBK) {
    *reinterpret_cast<uint4 *>(As + arow * kLds + ac) = ra;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ch = tid + 256 * q;
      *reinterpret_cast<uint4 *>(Bs + (ch >> 2) * kLds + (ch & 3) * 8) = rb[q];
    }
    __syncthreads();
    if (k0 + kBK < K) {
      ra = *reinterpret_cast<const uint4 *>(Ap + k0 + kBK);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ch = tid + 256 * q;
        rb[q] = *reinterpret_cast<const uint4 *>(B + long(ch >> 2) * K + k0 + kBK + (ch & 3) * 8);
      }
    }
    mma_slab(As, Bs, lane, w, acc);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long row = r0 + 16 * i + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n0 = 64 * w + 16 * j + 4 * (lane >> 4);
      const long idx = row * kW + n0;
      if (EPI == 0) {
        const float4 bv = *reinterpret_cast<const float4 *>(bias + n0);
        const float z0 = acc[i][j][0] + bv.x, z1 = acc[i][j][1] + bv.y, z2 = acc[i][j][2] + bv.z,
                    z3 = acc[i][j][3] + bv.w;
        *reinterpret_cast<uint2 *>(O1 + idx) = make_uint2(pack2(z0, z1), pack2(z2, z3));
        *reinterpret_cast<uint2 *>(O2 + idx) = make_uint2(pack2(gelu(z0), gelu(z1)), pack2(gelu(z2), gelu(z3)));
      } else {
        const uint2 zz = *reinterpret_cast<const uint2 *>(Zp + idx);
        const float z0 = bf2f(u16(zz.x & 0xffffu)), z1 = bf2f(u16(zz.x >> 16)), z2 = bf2f(u16(zz.y & 0xffffu)),
                    z3 = bf2f(u16(zz.y >> 16));
        *reinterpret_cast<uint2 *>(O1 + idx) =
            make_uint2(pack2(acc[i][j][0] * gelu_grad(z0), acc[i][j][1] * gelu_grad(z1)),
                       pack2(acc[i][j][2] * gelu_grad(z2), acc[i][j][3] * gelu_grad(z3)));
      }
    }
  }
}

// Large-tile form of k_gemm_nt for M a multiple of 256: 16 wavefronts own
// 256 rows x 256 columns (wave w: rows 64(w/4).., columns 64(w%4)..), so a
// workgroup reads each weight slab once per 256 rows instead of per 64
// (weights are a quarter of the L2 traffic), and the slabs are double
// buffered in LDS: one barrier per slab.  Same arithmetic as k_gemm_nt.
constexpr int kBigRows = 256, kBigThreads = 1024;

template <int EPI>
__global__ __launch_bounds__(kBigThreads) void k_gemm_nt_big(int K, const u16 *__restrict__ A,
                                                             const u16 *__restrict__ B, const float *__restrict__ bias,
                                                             const u16 *__restrict__ Zp, u16 *__restrict__ O1,
                                                             u16 *__restrict__ O2) {
  __shared__ __attribute__((aligned(16))) u16 As[2][kBigRows * kLds];
  __shared__ __attribute__((aligned(16))) u16 Bs[2][kW * kLds];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wm = wv >> 2, wn = wv & 3;
  const long r0 = long(blockIdx.x) * kBigRows;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // one 16-byte chunk of A and one of B per thread per slab (256 rows x 32)
  const int row = tid >> 2, c8 = (tid & 3) * 8;
  const u16 *Ap = A + (r0 + row) * K + c8;
  const u16 *Bp = B + long(row) * K + c8;
  uint4 ra = *reinterpret_cast<const uint4 *>(Ap);
  uint4 rb = *reinterpret_cast<const uint4 *>(Bp);
  *reinterpret_cast<uint4 *>(&As[0][row * kLds + c8]) = ra;
  *reinterpret_cast<uint4 *>(&Bs[0][row * kLds + c8]) = rb;
  __syncthreads();
  const int ns = K / kBK;
  for (int st = 0; st < ns; ++st) {
    const int cur = st & 1;
    const bool more = st + 1 < ns;
    if (more) {
      ra = *reinterpret_cast<const uint4 *>(Ap + (st + 1) * kBK);
      rb = *reinterpret_cast<const uint4 *>(Bp + (st + 1) * kBK);
    }
    mma_slab(&As[cur][(64 * wm) * kLds], &Bs[cur][0], lane, wn, acc);
    if (more) {
      *reinterpret_cast<uint4 *>(&As[cur ^ 1][row * kLds + c8]) = ra;
      *reinterpret_cast<uint4 *>(&Bs[cur ^ 1][row * kLds + c8]) = rb;
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long rr = r0 + 64 * wm + 16 * i + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n0 = 64 * wn + 16 * j + 4 * (lane >> 4);
      const long idx = rr * kW + n0;
      if (EPI == 0) {
        const float4 bv = *reinterpret_cast<const float4 *>(bias + n0);
        const float z0 = acc[i][j][0] + bv.x, z1 = acc[i][j][1] + bv.y, z2 = acc[i][j][2] + bv.z,
                    z3 = acc[i][j][3] + bv.w;
        *reinterpret_cast<uint2 *>(O1 + idx) = make_uint2(pack2(z0, z1), pack2(z2, z3));
        *reinterpret_cast<uint2 *>(O2 + idx) = make_uint2(pack2(gelu(z0), gelu(z1)), pack2(gelu(z2), gelu(z3)));
      } else {
        const uint2 zz = *reinterpret_cast<const uint2 *>(Zp + idx);
        const float z0 = bf2f(u16(zz.x & 0xffffu)), z1 = bf2f(u16(zz.x >> 16)), z2 = bf2f(u16(zz.y & 0xffffu)),
                    z3 = bf2f(u16(zz.y >> 16));
        *reinterpret_cast<uint2 *>(O1 + idx) =
            make_uint2(pack2(acc[i][j][0] * gelu_grad(z0), acc[i][j][1] * gelu_grad(z1)),
                       pack2(acc[i][j][2] * gelu_grad(z2), acc[i][j][3] * gelu_grad(z3)));
      }
    }
  }
}

// move rows (2mp, 2mp+1) x 8 columns, held as v0/v1, into a transposed LDS
// tile T[col][m] as 32-bit pairs
__device__ __forceinline__ void put_pairs(u16 *__restrict__ T, int col0, int mp, uint4 v0, uint4 v1) {
  const unsigned a0[4] = {v0.x, v0.y, v0.z, v0.w}, a1[4] = {v1.x, v1.y, v1.z, v1.w};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const unsigned lo = (a0[c >> 1] >> (16 * (c & 1))) & 0xffffu, hi = (a1[c >> 1] >> (16 * (c & 1))) & 0xffffu;
    *reinterpret_cast<unsigned *>(T + (col0 + c) * kLds + 2 * mp) = lo | (hi << 16);
  }
}

// Workgroup -> (k-tile, row slice) of the weight-gradient kernels.  Workgroups
// go to the eight XCDs round-robin by index; the nkt k-tiles of one slice read
// the same D rows, so they are given indices eight apart (one XCD, whose L2
// then serves D to all of them) and dispatched close together.  Needs the
// slice count to be a multiple of 8, else the plain order is used.
__device__ __forceinline__ void tn_tile(int nkt, int nsplit, int &kt, int &s) {
  const int b = int(blockIdx.x);
  if (nsplit % 8 == 0) {
    const int r = b & 7, q = b >> 3;
    kt = q % nkt;
    s = (q / nkt) * 8 + r;
  } else {
    kt = b % nkt;
    s = b / nkt;
  }
}

// Weight-gradient partials: for k-tile kt (64 rows of dW) and row slice s,
// dW[k][n] = sum over m in slice of A[m][k] * D[m][n], A[M][K], D[M][256].
// Both operands are transposed into LDS (m contiguous); a thread moves two
// rows' worth of 8 columns and writes 32-bit pairs; the next slab's global
// data is loaded into registers while the current one is multiplied.
// Workgroups of k-tile 0 also sum D's columns (the bias gradient).  Output
// rows k >= Kreal (input padding) are not written.
__global__ __launch_bounds__(256) void k_gemm_tn(int K, int Kreal, int chunk, const u16 *__restrict__ A,
                                                 const u16 *__restrict__ D, float *__restrict__ G, long P,
                                                 long offW, long offb) {
  __shared__ __attribute__((aligned(16))) u16 As[kRows * kLds];
  __shared__ __attribute__((aligned(16))) u16 Bs[kW * kLds];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nkt = K / kRows;
  int kt, s;
  tn_tile(nkt, int(gridDim.x) / nkt, kt, s);
  const long m0 = long(s) * chunk;
  const int kb = kt * kRows;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float db = 0.f;
  const int amp = tid & 15, akg = tid >> 4;  // A: threads < 128, rows 2amp, 2amp+1, columns 8akg..
  const u16 *Ap = A + (m0 + 2 * amp) * K + kb + 8 * akg;
  const int bmp0 = tid & 15, bng0 = tid >> 4, bng1 = (tid + 256) >> 4;  // D: two chunks per thread
  const u16 *Dp0 = D + (m0 + 2 * bmp0) * kW + 8 * bng0;
  const u16 *Dp1 = D + (m0 + 2 * bmp0) * kW + 8 * bng1;
  uint4 ra0{}, ra1{}, rb00, rb01, rb10, rb11;
  if (tid < 128) {
    ra0 = *reinterpret_cast<const uint4 *>(Ap);
    ra1 = *reinterpret_cast<const uint4 *>(Ap + K);
  }
  rb00 = *reinterpret_cast<const uint4 *>(Dp0);
  rb01 = *reinterpret_cast<const uint4 *>(Dp0 + kW);
  rb10 = *reinterpret_cast<const uint4 *>(Dp1);
  rb11 = *reinterpret_cast<const uint4 *>(Dp1 + kW);
  for (int mm = 0; mm < chunk; mm += kBK) {
    if (tid < 128) put_pairs(As, 8 * akg, amp, ra0, ra1);
    put_pairs(Bs, 8 * bng0, bmp0, rb00, rb01);
    put_pairs(Bs, 8 * bng1, bmp0, rb10, rb11);
    __syncthreads();
    if (mm + kBK < chunk) {
      const long o = long(mm + kBK);
      if (tid < 128) {
        ra0 = *reinterpret_cast<const uint4 *>(Ap + o * K);
        ra1 = *reinterpret_cast<const uint4 *>(Ap + (o + 1) * K);
      }
      rb00 = *reinterpret_cast<const uint4 *>(Dp0 + o * kW);
      rb01 = *reinterpret_cast<const uint4 *>(Dp0 + (o + 1) * kW);
      rb10 = *reinterpret_cast<const uint4 *>(Dp1 + o * kW);
      rb11 = *reinterpret_cast<const uint4 *>(Dp1 + (o + 1) * kW);
    }
    if (kt == 0) {
      const u16 *row = Bs + tid * kLds;
      for (int c = 0; c < kBK; ++c) db += bf2f(row[c]);
    }
    mma_slab(As, Bs, lane, w, acc);
    __syncthreads();
  }
  float *Gs = G + long(s) * P;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = kb + 16 * i + (lane & 15);
    if (k >= Kreal) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n0 = 64 * w + 16 * j + 4 * (lane >> 4);
      *reinterpret_cast<float4 *>(Gs + offW + long(k) * kW + n0) =
          make_float4(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
    }
  }
  if (kt == 0) Gs[offb + tid] = db;
}

// The same weight-gradient partials with four wave groups per workgroup
// (1024 threads): group g multiplies slabs g, g + 4, g + 8, ... of the row
// slice into its own accumulators and LDS buffers, so a CU keeps 16
// wavefronts in flight instead of 4; the groups' sums are then added into
// group 0's in group order (((g0 + g1) + g2) + g3), through LDS, one row
// block at a time.  Needs chunk to be a multiple of 4 kBK.
constexpr int kTnGroups = 4;
constexpr int kTnBuf = (kRows + kW) * kLds;  // one group's A and D slabs (elements)
__global__ __launch_bounds__(256 * kTnGroups) void k_gemm_tn_g(int K, int Kreal, int chunk, const u16 *__restrict__ A,
                                                               const u16 *__restrict__ D, float *__restrict__ G,
                                                               long P, long offW, long offb) {
  __shared__ __attribute__((aligned(16))) u16 smem[kTnGroups * kTnBuf];
  const int tid = threadIdx.x & 255, grp = threadIdx.x >> 8, lane = tid & 63, w = tid >> 6;
  u16 *As = smem + grp * kTnBuf, *Bs = As + kRows * kLds;
  const int nkt = K / kRows;
  int kt, s;
  tn_tile(nkt, int(gridDim.x) / nkt, kt, s);
  const long m0 = long(s) * chunk + grp * kBK;
  const int kb = kt * kRows;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float db = 0.f;
  const int amp = tid & 15, akg = tid >> 4;
  const u16 *Ap = A + (m0 + 2 * amp) * K + kb + 8 * akg;
  const int bmp0 = tid & 15, bng0 = tid >> 4, bng1 = (tid + 256) >> 4;
  const u16 *Dp0 = D + (m0 + 2 * bmp0) * kW + 8 * bng0;
  const u16 *Dp1 = D + (m0 + 2 * bmp0) * kW + 8 * bng1;
  uint4 ra0{}, ra1{}, rb00, rb01, rb10, rb11;
  if (tid < 128) {
    ra0 = *reinterpret_cast<const uint4 *>(Ap);
    ra1 = *reinterpret_cast<const uint4 *>(Ap + K);
  }
  rb00 = *reinterpret_cast<const uint4 *>(Dp0);
  rb01 = *reinterpret_cast<const uint4 *>(Dp0 + kW);
  rb10 = *reinterpret_cast<const uint4 *>(Dp1);
  rb11 = *reinterpret_cast<const uint4 *>(Dp1 + kW);
  constexpr int stride = kTnGroups * kBK;
  for (int mm = 0; mm < chunk; mm += stride) {
    if (tid < 128) put_pairs(As, 8 * akg, amp, ra0, ra1);
    put_pairs(Bs, 8 * bng0, bmp0, rb00, rb01);
    put_pairs(Bs, 8 * bng1, bmp0, rb10, rb11);
    __syncthreads();
    if (mm + stride < chunk) {
      const long o = long(mm + stride);
      if (tid < 128) {
        ra0 = *reinterpret_cast<const uint4 *>(Ap + o * K);
        ra1 = *reinterpret_cast<const uint4 *>(Ap + (o + 1) * K);
      }
      rb00 = *reinterpret_cast<const uint4 *>(Dp0 + o * kW);
      rb01 = *reinterpret_cast<const uint4 *>(Dp0 + (o + 1) * kW);
      rb10 = *reinterpret_cast<const uint4 *>(Dp1 + o * kW);
      rb11 = *reinterpret_cast<const uint4 *>(Dp1 + (o + 1) * kW);
    }
    if (kt == 0) {
      const u16 *row = Bs + tid * kLds;
      for (int c = 0; c < kBK; ++c) db += bf2f(row[c]);
    }
    mma_slab(As, Bs, lane, w, acc);
    __syncthreads();
  }
  // groups 1..3 hand their sums to group 0 through LDS, one row block i at a time
  float *red = reinterpret_cast<float *>(smem);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (grp > 0)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        *reinterpret_cast<f32x4 *>(red + (long((grp - 1) * 256 + tid) * 16 + 4 * j)) = acc[i][j];
    __syncthreads();
    if (grp == 0)
      for (int g2 = 1; g2 < kTnGroups; ++g2)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const f32x4 o = *reinterpret_cast<const f32x4 *>(red + (long((g2 - 1) * 256 + tid) * 16 + 4 * j));
          acc[i][j] = acc[i][j] + o;
        }
    __syncthreads();
  }
  if (kt == 0) {
    if (grp > 0) red[(grp - 1) * 256 + tid] = db;
    __syncthreads();
    if (grp == 0)
      for (int g2 = 1; g2 < kTnGroups; ++g2) db += red[(g2 - 1) * 256 + tid];
  }
  if (grp != 0) return;
  float *Gs = G + long(s) * P;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = kb + 16 * i + (lane & 15);
    if (k >= Kreal) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n0 = 64 * w + 16 * j + 4 * (lane >> 4);
      *reinterpret_cast<float4 *>(Gs + offW + long(k) * kW + n0) =
          make_float4(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
    }
  }
  if (kt == 0) Gs[offb + tid] = db;
}

// ------------------------------------------------------------ output layer
constexpr int kHeadRows = 128;                 // rows per workgroup
constexpr int kHeadPart = kW * kOut + kOut;    // output-layer parameters

// Workgroup b handles rows [128b, min(128b + 128, M)), 8 rows at a time (M is
// a multiple of 64, so every lane of a row pair runs the same trips), 32 lanes per
// row, 8 hidden columns per lane.  y = H Wh + bh in float32.  With dZ set:
// dy = 2 (y - t) / (nvalid * 4) for rows < nvalid (0 otherwise), the last
// hidden layer's dZ = bf16((dy Wh^T) * GELU'(Z)), and this workgroup's
// partial gradients of Wh and bh (Gh[b][...]) and loss sum (lossp[b]).
__global__ __launch_bounds__(256) void k_head(long M, long nvalid, float inv_no, const u16 *__restrict__ H,
                                              const u16 *__restrict__ Z, const float *__restrict__ Wh,
                                              const float *__restrict__ bh, const float *__restrict__ T,
                                              float *__restrict__ Y, u16 *__restrict__ dZ, float *__restrict__ Gh,
                                              float *__restrict__ lossp) {
  __shared__ float Ws[kW * kOut];
  __shared__ float red[4][32][8 * kOut + kOut + 1];
  const int tid = threadIdx.x, q = tid & 31, rs = tid >> 5, lane = tid & 63, wv = tid >> 6;
  for (int e = tid; e < kW * kOut; e += 256) Ws[e] = Wh[e];
  __syncthreads();
  float gw[8][kOut], gb[kOut], ls = 0.f;
  for (int c = 0; c < 8; ++c)
    for (int o = 0; o < kOut; ++o) gw[c][o] = 0.f;
  for (int o = 0; o < kOut; ++o) gb[o] = 0.f;
  const long mbeg = long(blockIdx.x) * kHeadRows, mend = mbeg + kHeadRows < M ? mbeg + kHeadRows : M;
  // the next row's H and Z are in flight while this row is processed
  uint4 hn{}, zn{};
  if (mbeg + rs < mend) {
    hn = *reinterpret_cast<const uint4 *>(H + (mbeg + rs) * kW + 8 * q);
    if (dZ) zn = *reinterpret_cast<const uint4 *>(Z + (mbeg + rs) * kW + 8 * q);
  }
  for (long m = mbeg + rs; m < mend; m += 8) {
    const uint4 hv = hn, zv = zn;
    if (m + 8 < mend) {
      hn = *reinterpret_cast<const uint4 *>(H + (m + 8) * kW + 8 * q);
      if (dZ) zn = *reinterpret_cast<const uint4 *>(Z + (m + 8) * kW + 8 * q);
    }
    const unsigned hw[4] = {hv.x, hv.y, hv.z, hv.w};
    float h[8];
    for (int c = 0; c < 8; ++c) h[c] = bf2f(u16((hw[c >> 1] >> (16 * (c & 1))) & 0xffffu));
    float y[kOut];
    for (int o = 0; o < kOut; ++o) {
      float t = 0.f;
      for (int c = 0; c < 8; ++c) t += h[c] * Ws[(8 * q + c) * kOut + o];
      y[o] = t;
    }
    for (int off = 16; off >= 1; off >>= 1)
      for (int o = 0; o < kOut; ++o) y[o] += __shfl_xor(y[o], off, 64);
    for (int o = 0; o < kOut; ++o) y[o] += bh[o];
    const bool valid = m < nvalid;
    if (Y && valid && q == 0)
      for (int o = 0; o < kOut; ++o) Y[m * kOut + o] = y[o];
    if (!dZ) continue;
    float dy[kOut];
    for (int o = 0; o < kOut; ++o) {
      const float r = valid ? y[o] - T[m * kOut + o] : 0.f;
      dy[o] = 2.f * r * inv_no;
      if (q == 0) ls += r * r;
    }
    const unsigned zw[4] = {zv.x, zv.y, zv.z, zv.w};
    unsigned pk[4] = {0, 0, 0, 0};
    for (int c = 0; c < 8; ++c) {
      float g = 0.f;
      for (int o = 0; o < kOut; ++o) g += dy[o] * Ws[(8 * q + c) * kOut + o];
      const float z = bf2f(u16((zw[c >> 1] >> (16 * (c & 1))) & 0xffffu));
      pk[c >> 1] |= unsigned(f2bf(g * gelu_grad(z))) << (16 * (c & 1));
      for (int o = 0; o < kOut; ++o) gw[c][o] += h[c] * dy[o];
    }
    *reinterpret_cast<uint4 *>(dZ + m * kW + 8 * q) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    if (q == 0)
      for (int o = 0; o < kOut; ++o) gb[o] += dy[o];
  }
  if (!dZ) return;
  // rows rs and rs^1 share a wavefront: add the two halves, then the 4 waves
  for (int c = 0; c < 8; ++c)
    for (int o = 0; o < kOut; ++o) gw[c][o] += __shfl_xor(gw[c][o], 32, 64);
  for (int o = 0; o < kOut; ++o) gb[o] += __shfl_xor(gb[o], 32, 64);
  ls += __shfl_xor(ls, 32, 64);
  if (lane < 32) {
    for (int c = 0; c < 8; ++c)
      for (int o = 0; o < kOut; ++o) red[wv][q][c * kOut + o] = gw[c][o];
    for (int o = 0; o < kOut; ++o) red[wv][q][8 * kOut + o] = gb[o];
    red[wv][q][8 * kOut + kOut] = ls;
  }
  __syncthreads();
  float *Gb = Gh + long(blockIdx.x) * kHeadPart;
  for (int e = tid; e < kW * kOut; e += 256) {
    const int k = e / kOut, o = e % kOut, qq = k / 8, c = k % 8;
    Gb[e] = red[0][qq][c * kOut + o] + red[1][qq][c * kOut + o] + red[2][qq][c * kOut + o] +
            red[3][qq][c * kOut + o];
  }
  if (tid < kOut)
    Gb[kW * kOut + tid] = red[0][0][8 * kOut + tid] + red[1][0][8 * kOut + tid] + red[2][0][8 * kOut + tid] +
                          red[3][0][8 * kOut + tid];
  if (tid == 0)
    lossp[blockIdx.x] = red[0][0][8 * kOut + kOut] + red[1][0][8 * kOut + kOut] + red[2][0][8 * kOut + kOut] +
                        red[3][0][8 * kOut + kOut];
}

// Output-layer gradient: column c of the nhb partial rows, summed by 16
// groups of rows (group g takes rows g, g+16, ...) and then the 16 group
// sums in a fixed tree; the order does not depend on scheduling.
__global__ __launch_bounds__(256) void k_head_reduce(int nhb, const float *__restrict__ Gh, float *__restrict__ gh) {
  __shared__ float red[16][16];
  const int t = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int c = blockIdx.x * 16 + t;
  float s = 0.f;
  if (c < kHeadPart)
    for (int b = g; b < nhb; b += 16) s += Gh[long(b) * kHeadPart + c];
  red[g][t] = s;
  __syncthreads();
  for (int h = 8; h >= 1; h >>= 1) {
    if (g < h) red[g][t] += red[g + h][t];
    __syncthreads();
  }
  if (g == 0 && c < kHeadPart) gh[c] = red[0][t];
}

// --------------------------------------------------------------- optimiser
struct LayerTab {
  long offW[kMaxLayers], offb[kMaxLayers];
  int in[kMaxLayers], out[kMaxLayers], inpad[kMaxLayers];
  u16 *Wt[kMaxLayers];  // [out][inpad] copies (forward operand)
  u16 *Wk[kMaxLayers];  // [in][out] copies (backward-data operand)
  int nl;
};

// mode 0: publish the bfloat16 copies only; 1: sum partials into gout;
// 2: sum partials, Adam update, publish.  Hidden-layer gradients come as
// nsplit slices of G, added in slice order; the output layer's (parameters
// from offHead on) already reduced in gh.  Workgroup 0 also reduces the nhb
// loss partials into the step's loss (lossh may be null).
__global__ __launch_bounds__(256) void k_adam(long P, int nsplit, int mode, const float *__restrict__ G,
                                              const float *__restrict__ gh, int nhb, long offHead,
                                              float *__restrict__ w, float *__restrict__ m1,
                                              float *__restrict__ m2, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2, LayerTab tab, float *__restrict__ gout,
                                              const float *__restrict__ lossp, double inv_no,
                                              double *__restrict__ lossh) {
  __shared__ double lred[256];
  const long p = long(blockIdx.x) * 256 + threadIdx.x;
  if (blockIdx.x == 0 && lossh && mode != 0) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nhb; k += 256) s += double(lossp[k]);
    lred[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
      if (int(threadIdx.x) < h) lred[threadIdx.x] += lred[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) *lossh = lred[0] * inv_no;
  }
  if (p >= P) return;
  if (mode != 0) {
    float g = 0.f;
    if (p >= offHead) {
      g = gh[p - offHead];
    } else {
      // the slices' partials in flight 16 at a time, summed in slice order
      int k = 0;
      for (; k + 16 <= nsplit; k += 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = G[long(k + u) * P + p];
#pragma unroll
        for (int u = 0; u < 16; ++u) g += v[u];
      }
      for (; k < nsplit; ++k) g += G[long(k) * P + p];
    }
    if (gout) gout[p] = g;
    if (mode == 2) {
      const float a = b1 * m1[p] + (1.f - b1) * g;
      const float b = b2 * m2[p] + (1.f - b2) * g * g;
      m1[p] = a;
      m2[p] = b;
      w[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
    }
  }
  if (mode == 1) return;
  for (int l = 0; l < tab.nl; ++l) {
    if (p >= tab.offW[l] && p < tab.offb[l]) {
      const long e = p - tab.offW[l];
      const int k = int(e / tab.out[l]), n = int(e % tab.out[l]);
      const u16 v = f2bf(w[p]);
      if (tab.Wt[l]) tab.Wt[l][long(n) * tab.inpad[l] + k] = v;
      if (tab.Wk[l]) tab.Wk[l][long(k) * tab.out[l] + n] = v;
    }
  }
}

}  // namespace

// ------------------------------------------------------------------ context
struct fvs_mlp {
  fvs_mlp_options opt{};
  int L = 0;   // hidden layers
  int M = 0;   // batch rows
  long P = 0;  // parameters
  LayerTab tab{};
  hipStream_t stream = nullptr;
  float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *G = nullptr, *Gh = nullptr, *gh = nullptr,
        *lossp = nullptr, *gtmp = nullptr;
  double *lossh = nullptr;
  u16 *X0 = nullptr, *dZa = nullptr, *dZb = nullptr;
  u16 *Z[kMaxDepth] = {}, *H[kMaxDepth] = {};
  float *T = nullptr, *Y = nullptr, *F32 = nullptr;
  float *raw = nullptr, *tgt = nullptr;  // training set
  long ndata = 0;
  float *iraw = nullptr, *itgt = nullptr;  // staging of caller rows (batch rows)
  long t = 0;                              // Adam step count
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
  std::vector<void *> allocs;
  // per-launch timing
  bool profiling = false;
  std::vector<hipEvent_t> evpool;
  std::vector<int> evclass;
  int evused = 0;

  template <class T_>
  int alloc(T_ *&p, size_t n) {
    void *q = nullptr;
    MLPC(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T_)));
    allocs.push_back(q);
    p = static_cast<T_ *>(q);
    return 0;
  }
  ~fvs_mlp() {
    for (void *q : allocs) (void)hipFree(q);
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

enum { MC_FEAT, MC_FWD, MC_HEAD, MC_WGRAD, MC_BWD, MC_ADAM, MC_N };

struct Prof {
  fvs_mlp &m;
  bool on;
  Prof(fvs_mlp &mm, int cls) : m(mm), on(mm.profiling && mm.evused + 2 <= int(mm.evpool.size())) {
    if (!on) return;
    (void)hipEventRecord(m.evpool[m.evused], m.stream);
    m.evclass.push_back(cls);
  }
  ~Prof() {
    if (!on) return;
    (void)hipEventRecord(m.evpool[m.evused + 1], m.stream);
    m.evused += 2;
  }
};

// largest divisor of M/32 that is at most kMaxSplit: row slices of the
// gradient reduction, each a multiple of 32 rows
int nsplit(int M) {
  const int u = M / kBK;
  for (int s = std::min(u, kMaxSplit); s >= 1; --s)
    if (u % s == 0) return s;
  return 1;
}

// forward (and with grad, backward) over the first M rows already in X0/T
int run_pass(fvs_mlp &m, int M, long nvalid, bool grad, float *Y) {
  const int S = nsplit(M), chunk = M / S;
  const int L = m.L;
  for (int l = 0; l < L; ++l) {
    Prof pf(m, MC_FWD);
    const u16 *A = (l == 0) ? m.X0 : m.H[l - 1];
    if (M % kBigRows == 0)
      hipLaunchKernelGGL(k_gemm_nt_big<0>, dim3(M / kBigRows), dim3(kBigThreads), 0, m.stream, m.tab.inpad[l], A,
                         m.tab.Wt[l], m.w + m.tab.offb[l], nullptr, m.Z[l], m.H[l]);
    else
      hipLaunchKernelGGL(k_gemm_nt<0>, dim3(M / kRows), dim3(256), 0, m.stream, m.tab.inpad[l], A, m.tab.Wt[l],
                         m.w + m.tab.offb[l], nullptr, m.Z[l], m.H[l]);
  }
  {
    Prof pf(m, MC_HEAD);
    const float inv_no = 1.0f / float(double(nvalid) * kOut);
    hipLaunchKernelGGL(k_head, dim3((M + kHeadRows - 1) / kHeadRows), dim3(256), 0, m.stream, long(M), nvalid, inv_no,
                       m.H[L - 1], m.Z[L - 1],
                       m.w + m.tab.offW[L], m.w + m.tab.offb[L], m.T, Y, grad ? m.dZa : nullptr, m.Gh, m.lossp);
  }
  if (grad) {
    u16 *cur = m.dZa, *oth = m.dZb;
    for (int l = L - 1; l >= 0; --l) {
      {
        Prof pf(m, MC_WGRAD);
        const u16 *A = (l == 0) ? m.X0 : m.H[l - 1];
        const int K = m.tab.inpad[l];
        if (chunk % (kTnGroups * kBK) == 0)
          hipLaunchKernelGGL(k_gemm_tn_g, dim3((K / kRows) * S), dim3(256 * kTnGroups), 0, m.stream, K,
                             m.tab.in[l], chunk, A, cur, m.G, m.P, m.tab.offW[l], m.tab.offb[l]);
        else
          hipLaunchKernelGGL(k_gemm_tn, dim3((K / kRows) * S), dim3(256), 0, m.stream, K, m.tab.in[l], chunk, A,
                             cur, m.G, m.P, m.tab.offW[l], m.tab.offb[l]);
      }
      if (l > 0) {
        Prof pf(m, MC_BWD);
        if (M % kBigRows == 0)
          hipLaunchKernelGGL(k_gemm_nt_big<1>, dim3(M / kBigRows), dim3(kBigThreads), 0, m.stream, kW, cur,
                             m.tab.Wk[l], nullptr, m.Z[l - 1], oth, nullptr);
        else
          hipLaunchKernelGGL(k_gemm_nt<1>, dim3(M / kRows), dim3(256), 0, m.stream, kW, cur, m.tab.Wk[l], nullptr,
                             m.Z[l - 1], oth, nullptr);
        std::swap(cur, oth);
      }
    }
  }
  MLPC(hipGetLastError());
  return 0;
}

int launch_adam(fvs_mlp &m, int mode, int S, int nhb, double inv_no, double *lossh, float *gout) {
  Prof pf(m, MC_ADAM);
  if (mode != 0)
    hipLaunchKernelGGL(k_head_reduce, dim3((kHeadPart + 15) / 16), dim3(256), 0, m.stream, nhb, m.Gh, m.gh);
  float bc1 = 1.f, bc2 = 1.f;
  if (mode == 2) {
    m.t += 1;
    bc1 = float(1.0 - std::pow(m.opt.beta1, double(m.t)));
    bc2 = float(1.0 - std::pow(m.opt.beta2, double(m.t)));
  }
  hipLaunchKernelGGL(k_adam, dim3(unsigned((m.P + 255) / 256)), dim3(256), 0, m.stream, m.P, S, mode, m.G, m.gh,
                     nhb, m.tab.offW[m.L], m.w, m.m1, m.m2, float(m.opt.lr), float(m.opt.beta1), float(m.opt.beta2),
                     float(m.opt.eps), bc1, bc2, m.tab, gout, m.lossp, inv_no, lossh);
  MLPC(hipGetLastError());
  return 0;
}

int stage_rows(fvs_mlp &m, int n, const float *raw, const float *target) {
  MLPC(hipMemcpyAsync(m.iraw, raw, size_t(n) * kRaw * sizeof(float), hipMemcpyHostToDevice, m.stream));
  if (target)
    MLPC(hipMemcpyAsync(m.itgt, target, size_t(n) * kOut * sizeof(float), hipMemcpyHostToDevice, m.stream));
  return 0;
}

void launch_features(fvs_mlp &m, int M, long nvalid, const float *raw, const float *tgt, long ndata, long step,
                     int sample, float *F32) {
  Prof pf(m, MC_FEAT);
  hipLaunchKernelGGL(k_features, dim3(unsigned((long(M) * 16 + 255) / 256)), dim3(256), 0, m.stream, M, nvalid, raw,
                     tgt, ndata, m.opt.seed, step, sample, m.X0, tgt ? m.T : nullptr, F32);
}

int round64(long n) { return int((n + kRows - 1) / kRows * kRows); }

}  // namespace

extern "C" {

int fvs_mlp_create(fvs_mlp **out, const fvs_mlp_options *opt) try {
  if (!out || !opt) return fail(FVS_EINVAL, "null argument");
  *out = nullptr;
  if (opt->depth < 1 || opt->depth > kMaxDepth) return fail(FVS_EINVAL, "depth must be 1..8");
  if (opt->batch < kRows || opt->batch % kRows) return fail(FVS_EINVAL, "batch must be a positive multiple of 64");
  if (!(opt->lr > 0) || !(opt->beta1 >= 0 && opt->beta1 < 1) || !(opt->beta2 >= 0 && opt->beta2 < 1) ||
      !(opt->eps > 0))
    return fail(FVS_EINVAL, "Adam parameters out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FVS_EHIP, "no HIP device");
  std::unique_ptr<fvs_mlp> holder(new fvs_mlp());  // freed on every early return or exception
  fvs_mlp *m = holder.get();
  m->opt = *opt;
  m->L = opt->depth;
  m->M = opt->batch;
  LayerTab &tb = m->tab;
  tb.nl = m->L + 1;
  long off = 0;
  for (int l = 0; l <= m->L; ++l) {
    tb.in[l] = (l == 0) ? kFeat : kW;
    tb.out[l] = (l == m->L) ? kOut : kW;
    tb.inpad[l] = (l == 0) ? kInPad : kW;
    tb.offW[l] = off;
    off += long(tb.in[l]) * tb.out[l];
    tb.offb[l] = off;
    off += tb.out[l];
  }
  m->P = off;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return fail(FVS_EHIP, "stream");
  if (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)
    return fail(FVS_EHIP, "event");
  const size_t M = size_t(m->M);
  int rc = 0;
  rc = rc ? rc : m->alloc(m->w, m->P);
  rc = rc ? rc : m->alloc(m->m1, m->P);
  rc = rc ? rc : m->alloc(m->m2, m->P);
  rc = rc ? rc : m->alloc(m->gtmp, m->P);
  rc = rc ? rc : m->alloc(m->G, size_t(kMaxSplit) * m->P);
  rc = rc ? rc : m->alloc(m->Gh, (M + kHeadRows - 1) / kHeadRows * kHeadPart);
  rc = rc ? rc : m->alloc(m->gh, kHeadPart);
  rc = rc ? rc : m->alloc(m->lossp, (M + kHeadRows - 1) / kHeadRows);
  rc = rc ? rc : m->alloc(m->lossh, 1);
  rc = rc ? rc : m->alloc(m->X0, M * kInPad);
  rc = rc ? rc : m->alloc(m->dZa, M * kW);
  rc = rc ? rc : m->alloc(m->dZb, M * kW);
  for (int l = 0; l < m->L && !rc; ++l) {
    rc = m->alloc(m->Z[l], M * kW);
    rc = rc ? rc : m->alloc(m->H[l], M * kW);
  }
  rc = rc ? rc : m->alloc(m->T, M * kOut);
  rc = rc ? rc : m->alloc(m->Y, M * kOut);
  rc = rc ? rc : m->alloc(m->F32, M * kFeat);
  rc = rc ? rc : m->alloc(m->iraw, M * kRaw);
  rc = rc ? rc : m->alloc(m->itgt, M * kOut);
  for (int l = 0; l < m->L && !rc; ++l) {
    rc = m->alloc(tb.Wt[l], size_t(kW) * tb.inpad[l]);
    if (!rc && l > 0) rc = m->alloc(tb.Wk[l], size_t(kW) * kW);
  }
  if (rc) return rc;
  tb.Wt[m->L] = tb.Wk[m->L] = nullptr;
  tb.Wk[0] = nullptr;
  if (hipMemsetAsync(tb.Wt[0], 0, size_t(kW) * kInPad * sizeof(u16), m->stream) != hipSuccess)
    return fail(FVS_EHIP, "memset");
  // Glorot-uniform weights from the seed, zero biases
  std::vector<float> w(m->P, 0.f);
  for (int l = 0; l <= m->L; ++l) {
    const double a = std::sqrt(6.0 / double(tb.in[l] + tb.out[l]));
    for (long e = 0; e < long(tb.in[l]) * tb.out[l]; ++e) {
      const unsigned long long r = mix64(opt->seed * 0x9E3779B97F4A7C15ull + (unsigned long long)(tb.offW[l] + e));
      const double u = double(r >> 11) * (1.0 / 9007199254740992.0);
      w[tb.offW[l] + e] = float((2.0 * u - 1.0) * a);
    }
  }
  rc = fvs_mlp_set_params(m, w.data());
  if (rc) return rc;
  *out = holder.release();
  return 0;
} catch (...) {
  return host_exception("fvs_mlp_create");
}

int fvs_mlp_destroy(fvs_mlp *m) try {
  delete m;
  return 0;
} catch (...) {
  return host_exception("fvs_mlp_destroy");
}

long long fvs_mlp_nparams(const fvs_mlp *m) { return m ? m->P : -1; }

int fvs_mlp_get_params(fvs_mlp *m, float *p) try {
  if (!m || !p) return fail(FVS_EINVAL, "null argument");
  MLPC(hipMemcpyAsync(p, m->w, size_t(m->P) * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLPC(hipStreamSynchronize(m->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_mlp_get_params");
}

int fvs_mlp_set_params(fvs_mlp *m, const float *p) try {
  if (!m || !p) return fail(FVS_EINVAL, "null argument");
  MLPC(hipMemcpyAsync(m->w, p, size_t(m->P) * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLPC(hipMemsetAsync(m->m1, 0, size_t(m->P) * sizeof(float), m->stream));
  MLPC(hipMemsetAsync(m->m2, 0, size_t(m->P) * sizeof(float), m->stream));
  m->t = 0;
  if (int rc = launch_adam(*m, 0, 0, 0, 0.0, nullptr, nullptr)) return rc;
  MLPC(hipStreamSynchronize(m->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_mlp_set_params");
}

int fvs_mlp_features(fvs_mlp *m, int n, const float *raw, float *feat) try {
  if (!m || (n > 0 && (!raw || !feat))) return fail(FVS_EINVAL, "null argument");
  if (n < 0 || n > m->M) return fail(FVS_EINVAL, "n must be 0..batch");
  if (n == 0) return 0;
  if (int rc = stage_rows(*m, n, raw, nullptr)) return rc;
  launch_features(*m, round64(n), n, m->iraw, nullptr, 0, 0, 0, m->F32);
  MLPC(hipGetLastError());
  MLPC(hipMemcpyAsync(feat, m->F32, size_t(n) * kFeat * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLPC(hipStreamSynchronize(m->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_mlp_features");
}

int fvs_mlp_set_data(fvs_mlp *m, long long n, const float *raw, const float *target) try {
  if (!m || !raw || !target) return fail(FVS_EINVAL, "null argument");
  if (n < 1) return fail(FVS_EINVAL, "empty training set");
  if (m->raw) {
    for (void *q : {static_cast<void *>(m->raw), static_cast<void *>(m->tgt)}) {
      (void)hipFree(q);
      m->allocs.erase(std::find(m->allocs.begin(), m->allocs.end(), q));
    }
    m->raw = m->tgt = nullptr;
  }
  if (int rc = m->alloc(m->raw, size_t(n) * kRaw)) return rc;
  if (int rc = m->alloc(m->tgt, size_t(n) * kOut)) return rc;
  MLPC(hipMemcpyAsync(m->raw, raw, size_t(n) * kRaw * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLPC(hipMemcpyAsync(m->tgt, target, size_t(n) * kOut * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLPC(hipStreamSynchronize(m->stream));
  m->ndata = long(n);
  return 0;
} catch (...) {
  return host_exception("fvs_mlp_set_data");
}

int fvs_mlp_train(fvs_mlp *m, int steps, double *loss) try {
  if (!m) return fail(FVS_EINVAL, "null argument");
  if (!m->raw) return fail(FVS_ESTATE, "no training set (fvs_mlp_set_data)");
  if (steps < 0) return fail(FVS_EINVAL, "steps < 0");
  if (steps == 0) return 0;
  double *dl = nullptr;
  MLPC(hipMalloc(&dl, size_t(steps) * sizeof(double)));
  const int S = nsplit(m->M);
  const double inv_no = 1.0 / (double(m->M) * kOut);
  int rc = 0;
  if (hipEventRecord(m->ev0, m->stream) != hipSuccess) rc = fail(FVS_EHIP, "event");
  for (int s = 0; s < steps && !rc; ++s) {
    launch_features(*m, m->M, m->M, m->raw, m->tgt, m->ndata, m->t + 1, 1, nullptr);
    rc = run_pass(*m, m->M, m->M, true, nullptr);
    if (!rc) rc = launch_adam(*m, 2, S, (m->M + kHeadRows - 1) / kHeadRows, inv_no, dl + s, nullptr);
  }
  if (!rc) {
    hipError_t e = hipEventRecord(m->ev1, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&m->last_ms, m->ev0, m->ev1);
    if (e != hipSuccess) rc = fail(FVS_EHIP, std::string("training: ") + hipGetErrorString(e));
  }
  if (!rc && loss && hipMemcpy(loss, dl, size_t(steps) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(FVS_EHIP, "copy of losses");
  (void)hipFree(dl);
  return rc;
} catch (...) {
  return host_exception("fvs_mlp_train");
}

static int grad_or_step(fvs_mlp *m, int n, const float *raw, const float *target, double *loss, float *grad,
                        bool update) {
  if (!m || !raw || !target) return fail(FVS_EINVAL, "null argument");
  if (n < 1 || n > m->M) return fail(FVS_EINVAL, "n must be 1..batch");
  const int M = round64(n);
  const int S = nsplit(M);
  if (int rc = stage_rows(*m, n, raw, target)) return rc;
  launch_features(*m, M, n, m->iraw, m->itgt, 0, 0, 0, nullptr);
  if (int rc = run_pass(*m, M, n, true, nullptr)) return rc;
  if (int rc = launch_adam(*m, update ? 2 : 1, S, (M + kHeadRows - 1) / kHeadRows, 1.0 / (double(n) * kOut), m->lossh,
                           update ? nullptr : m->gtmp))
    return rc;
  if (grad) MLPC(hipMemcpyAsync(grad, m->gtmp, size_t(m->P) * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  double l = 0.0;
  MLPC(hipMemcpyAsync(&l, m->lossh, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  MLPC(hipStreamSynchronize(m->stream));
  if (loss) *loss = l;
  return 0;
}

int fvs_mlp_grad(fvs_mlp *m, int n, const float *raw, const float *target, double *loss, float *grad) try {
  return grad_or_step(m, n, raw, target, loss, grad, false);
} catch (...) {
  return host_exception("fvs_mlp_grad");
}

int fvs_mlp_step(fvs_mlp *m, int n, const float *raw, const float *target, double *loss) try {
  return grad_or_step(m, n, raw, target, loss, nullptr, true);
} catch (...) {
  return host_exception("fvs_mlp_step");
}

int fvs_mlp_predict(fvs_mlp *m, long long n, const float *raw, float *out) try {
  if (!m || (n > 0 && (!raw || !out))) return fail(FVS_EINVAL, "null argument");
  if (n < 0) return fail(FVS_EINVAL, "n < 0");
  for (long long b = 0; b < n; b += m->M) {
    const int nb = int(std::min<long long>(m->M, n - b));
    const int M = round64(nb);
    if (int rc = stage_rows(*m, nb, raw + b * kRaw, nullptr)) return rc;
    launch_features(*m, M, nb, m->iraw, nullptr, 0, 0, 0, nullptr);
    if (int rc = run_pass(*m, M, nb, false, m->Y)) return rc;
    MLPC(hipMemcpyAsync(out + b * kOut, m->Y, size_t(nb) * kOut * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLPC(hipStreamSynchronize(m->stream));
  }
  return 0;
} catch (...) {
  return host_exception("fvs_mlp_predict");
}

int fvs_mlp_kernel_times(fvs_mlp *m, int steps, double *ms6, double *flop6, int *launches6) try {
  if (!m || !ms6 || !flop6 || !launches6) return fail(FVS_EINVAL, "null argument");
  if (!m->raw) return fail(FVS_ESTATE, "no training set (fvs_mlp_set_data)");
  if (steps < 1) return fail(FVS_EINVAL, "steps must be positive");
  const int per_step = 2 * (1 + m->L + 1 + m->L + (m->L - 1) + 1);
  const int need = per_step * steps;
  while (int(m->evpool.size()) < need) {
    hipEvent_t e;
    MLPC(hipEventCreate(&e));
    m->evpool.push_back(e);
  }
  m->evused = 0;
  m->evclass.clear();
  m->profiling = true;
  int rc = fvs_mlp_train(m, steps, nullptr);
  m->profiling = false;
  if (rc) return rc;
  double ms[MC_N] = {0}, cnt[MC_N] = {0};
  for (size_t k = 0; k < m->evclass.size(); ++k) {
    float t = 0.f;
    MLPC(hipEventElapsedTime(&t, m->evpool[2 * k], m->evpool[2 * k + 1]));
    ms[m->evclass[k]] += t;
    cnt[m->evclass[k]] += 1;
  }
  // matrix-core flops per launch (2 per multiply-add), averaged over layers
  const double M = m->M;
  double fl[MC_N] = {0};
  double fwd = 0, wg = 0;
  for (int l = 0; l < m->L; ++l) {
    fwd += 2.0 * M * m->tab.inpad[l] * kW;
    wg += 2.0 * M * m->tab.inpad[l] * kW;
  }
  fl[MC_FWD] = fwd / m->L;
  fl[MC_WGRAD] = wg / m->L;
  fl[MC_BWD] = 2.0 * M * kW * kW;
  fl[MC_HEAD] = 2.0 * M * kW * kOut * 3.0;
  for (int k = 0; k < MC_N; ++k) {
    ms6[k] = cnt[k] ? ms[k] / cnt[k] : 0.0;
    flop6[k] = fl[k];
    launches6[k] = int(cnt[k] / steps);
  }
  return 0;
} catch (...) {
  return host_exception("fvs_mlp_kernel_times");
}

double fvs_mlp_last_elapsed_ms(const fvs_mlp *m) { return m ? double(m->last_ms) : 0.0; }

}  // extern "C"
