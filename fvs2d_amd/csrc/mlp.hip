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

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// splitmix64 finaliser (Steele, Lea, Flood 2014): row sampling and weight init
__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------- features
// 16 threads per row, thread q writes input columns 8q..8q+7 (16 bytes).
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
  if (valid) {
    const float *r = raw + src * kRaw;
    const double x = r[0], y = r[1], d = r[2], mach = r[3], aoa = r[4], th = r[5];
    const double dx = x - 0.5;
    const double rr = sqrt(dx * dx + y * y);
    const double ang = atan2(y, dx);
    const double s = log1p(4.0 * rr) / log1p(200.0);
    const double kPi = 3.14159265358979323846;
    for (int c = 0; c < 8; ++c) {
      const int j = 8 * q + c;
      double v = 0.0;
      if (j < 32) {
        const double a = double(j / 2 + 1) * kPi * s;
        v = (j & 1) ? cos(a) : sin(a);
      } else if (j < 64) {
        const double a = double((j - 32) / 2 + 1) * ang;
        v = (j & 1) ? cos(a) : sin(a);
      } else if (j == 64) {
        v = log1p(100.0 * d) / log1p(5000.0);
      } else if (j == 65) {
        v = (mach - 0.675) / 0.175;
      } else if (j == 66) {
        v = (aoa - 1.0) / 3.0;
      } else if (j == 67) {
        v = (th - 0.125) / 0.025;
      }
      f[c] = float(v);
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
// acc[i][j] += A_lds[16i.., :] * B_lds[64w + 16j.., :]^T over one kBK slab.
// Fragment layout of v_mfma_f32_16x16x32_bf16 (lane l): A row l%16, k
// 8(l/16)..+7; B column l%16, same k; D rows 4(l/16)..+3, column l%16.
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
    for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
}

// C[M][256] = A[M][K] * B[256][K]^T (both row-major, K a multiple of 32).
// EPI 0 (forward): z = C + bias; Z = bf16(z), H = bf16(GELU(z)).
// EPI 1 (backward data): dZ = bf16(C * GELU'(Zp)).
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
  for (int k0 = 0; k0 < K; k0 += kBK) {
    {
      const int row = tid >> 2, c = (tid & 3) * 8;
      *reinterpret_cast<uint4 *>(As + row * kLds + c) =
          *reinterpret_cast<const uint4 *>(A + (r0 + row) * K + k0 + c);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ch = tid + 256 * q, row = ch >> 2, c = (ch & 3) * 8;
      *reinterpret_cast<uint4 *>(Bs + row * kLds + c) = *reinterpret_cast<const uint4 *>(B + long(row) * K + k0 + c);
    }
    __syncthreads();
    mma_slab(As, Bs, lane, w, acc);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = 64 * w + 16 * j + (lane & 15);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long row = r0 + 16 * i + 4 * (lane >> 4) + e;
        const long idx = row * kW + col;
        if (EPI == 0) {
          const float z = acc[i][j][e] + bias[col];
          O1[idx] = f2bf(z);
          O2[idx] = f2bf(gelu(z));
        } else {
          O1[idx] = f2bf(acc[i][j][e] * gelu_grad(bf2f(Zp[idx])));
        }
      }
    }
}

// Weight-gradient partials: for k-tile kt (64 rows of dW) and row slice s,
// dW[k][n] = sum over m in slice of A[m][k] * D[m][n], A[M][K], D[M][256].
// Both operands are transposed into LDS (m contiguous); a thread moves two
// rows' worth of 8 columns and writes 32-bit pairs.  Workgroups of k-tile 0
// also sum D's columns (the bias gradient).  Output rows k >= Kreal (input
// padding) are not written.
__global__ __launch_bounds__(256) void k_gemm_tn(int K, int Kreal, int chunk, const u16 *__restrict__ A,
                                                 const u16 *__restrict__ D, float *__restrict__ G, long P,
                                                 long offW, long offb) {
  __shared__ __attribute__((aligned(16))) u16 As[kRows * kLds];
  __shared__ __attribute__((aligned(16))) u16 Bs[kW * kLds];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nkt = K / kRows;
  const int kt = blockIdx.x % nkt, s = blockIdx.x / nkt;
  const long m0 = long(s) * chunk;
  const int kb = kt * kRows;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float db = 0.f;
  for (int mm = 0; mm < chunk; mm += kBK) {
    if (tid < 128) {
      const int mp = tid & 15, kg = tid >> 4;  // rows 2mp, 2mp+1; columns 8kg..8kg+7
      const long src = (m0 + mm + 2 * mp) * K + kb + 8 * kg;
      const uint4 v0 = *reinterpret_cast<const uint4 *>(A + src);
      const uint4 v1 = *reinterpret_cast<const uint4 *>(A + src + K);
      const unsigned a0[4] = {v0.x, v0.y, v0.z, v0.w}, a1[4] = {v1.x, v1.y, v1.z, v1.w};
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const unsigned lo = (a0[c >> 1] >> (16 * (c & 1))) & 0xffffu, hi = (a1[c >> 1] >> (16 * (c & 1))) & 0xffffu;
        *reinterpret_cast<unsigned *>(As + (8 * kg + c) * kLds + 2 * mp) = lo | (hi << 16);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int ch = tid + 256 * q, mp = ch & 15, ng = ch >> 4;  // ng 0..31
      const long src = (m0 + mm + 2 * mp) * kW + 8 * ng;
      const uint4 v0 = *reinterpret_cast<const uint4 *>(D + src);
      const uint4 v1 = *reinterpret_cast<const uint4 *>(D + src + kW);
      const unsigned a0[4] = {v0.x, v0.y, v0.z, v0.w}, a1[4] = {v1.x, v1.y, v1.z, v1.w};
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const unsigned lo = (a0[c >> 1] >> (16 * (c & 1))) & 0xffffu, hi = (a1[c >> 1] >> (16 * (c & 1))) & 0xffffu;
        *reinterpret_cast<unsigned *>(Bs + (8 * ng + c) * kLds + 2 * mp) = lo | (hi << 16);
      }
    }
    __syncthreads();
    if (kt == 0) {
      const u16 *row = Bs + tid * kLds;
      for (int c = 0; c < kBK; ++c) db += bf2f(row[c]);
    }
    mma_slab(As, Bs, lane, w, acc);
    __syncthreads();
  }
  float *Gs = G + long(s) * P;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = 64 * w + 16 * j + (lane & 15);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = kb + 16 * i + 4 * (lane >> 4) + e;
        if (k < Kreal) Gs[offW + long(k) * kW + n] = acc[i][j][e];
      }
    }
  if (kt == 0) Gs[offb + tid] = db;
}

// ------------------------------------------------------------ output layer
// Block s handles rows [s*chunk, (s+1)*chunk), 8 rows at a time, 32 lanes
// per row, 8 hidden columns per lane.  y = H Wh + bh in float32.  With dZ
// set: dy = 2 (y - t) / (nvalid * 4) for rows < nvalid (0 otherwise), the
// last hidden layer's dZ = bf16((dy Wh^T) * GELU'(Z)), and this slice's
// partial gradients of Wh, bh and the loss sum.
__global__ __launch_bounds__(256) void k_head(int chunk, long nvalid, float inv_no, const u16 *__restrict__ H,
                                              const u16 *__restrict__ Z, const float *__restrict__ Wh,
                                              const float *__restrict__ bh, const float *__restrict__ T,
                                              float *__restrict__ Y, u16 *__restrict__ dZ, float *__restrict__ G,
                                              long P, long offW, long offb, float *__restrict__ lossp) {
  __shared__ float Ws[kW * kOut];
  __shared__ float red[4][32][8 * kOut + kOut + 1];
  const int tid = threadIdx.x, q = tid & 31, rs = tid >> 5, lane = tid & 63, wv = tid >> 6;
  for (int e = tid; e < kW * kOut; e += 256) Ws[e] = Wh[e];
  __syncthreads();
  float gw[8][kOut], gb[kOut], ls = 0.f;
  for (int c = 0; c < 8; ++c)
    for (int o = 0; o < kOut; ++o) gw[c][o] = 0.f;
  for (int o = 0; o < kOut; ++o) gb[o] = 0.f;
  const long mbeg = long(blockIdx.x) * chunk;
  for (long m = mbeg + rs; m < mbeg + chunk; m += 8) {
    const uint4 hv = *reinterpret_cast<const uint4 *>(H + m * kW + 8 * q);
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
    const uint4 zv = *reinterpret_cast<const uint4 *>(Z + m * kW + 8 * q);
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
  float *Gs = G + long(blockIdx.x) * P;
  for (int e = tid; e < kW * kOut; e += 256) {
    const int k = e / kOut, o = e % kOut, qq = k / 8, c = k % 8;
    Gs[offW + e] = red[0][qq][c * kOut + o] + red[1][qq][c * kOut + o] + red[2][qq][c * kOut + o] +
                   red[3][qq][c * kOut + o];
  }
  if (tid < kOut)
    Gs[offb + tid] = red[0][0][8 * kOut + tid] + red[1][0][8 * kOut + tid] + red[2][0][8 * kOut + tid] +
                     red[3][0][8 * kOut + tid];
  if (tid == 0)
    lossp[blockIdx.x] = red[0][0][8 * kOut + kOut] + red[1][0][8 * kOut + kOut] + red[2][0][8 * kOut + kOut] +
                        red[3][0][8 * kOut + kOut];
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
// 2: sum partials, Adam update, publish.  Block 0 thread 0 also turns the
// loss partials into the step's loss (lossh may be null).
__global__ __launch_bounds__(256) void k_adam(long P, int nsplit, int mode, const float *__restrict__ G,
                                              float *__restrict__ w, float *__restrict__ m1,
                                              float *__restrict__ m2, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2, LayerTab tab, float *__restrict__ gout,
                                              const float *__restrict__ lossp, double inv_no,
                                              double *__restrict__ lossh) {
  const long p = long(blockIdx.x) * 256 + threadIdx.x;
  if (p == 0 && lossh && mode != 0) {
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += double(lossp[k]);
    *lossh = s * inv_no;
  }
  if (p >= P) return;
  if (mode != 0) {
    float g = 0.f;
    for (int k = 0; k < nsplit; ++k) g += G[long(k) * P + p];
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
  float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *G = nullptr, *lossp = nullptr, *gtmp = nullptr;
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
    hipLaunchKernelGGL(k_gemm_nt<0>, dim3(M / kRows), dim3(256), 0, m.stream, m.tab.inpad[l], A, m.tab.Wt[l],
                       m.w + m.tab.offb[l], nullptr, m.Z[l], m.H[l]);
  }
  {
    Prof pf(m, MC_HEAD);
    const float inv_no = 1.0f / float(double(nvalid) * kOut);
    hipLaunchKernelGGL(k_head, dim3(S), dim3(256), 0, m.stream, chunk, nvalid, inv_no, m.H[L - 1], m.Z[L - 1],
                       m.w + m.tab.offW[L], m.w + m.tab.offb[L], m.T, Y, grad ? m.dZa : nullptr, m.G, m.P,
                       m.tab.offW[L], m.tab.offb[L], m.lossp);
  }
  if (grad) {
    u16 *cur = m.dZa, *oth = m.dZb;
    for (int l = L - 1; l >= 0; --l) {
      {
        Prof pf(m, MC_WGRAD);
        const u16 *A = (l == 0) ? m.X0 : m.H[l - 1];
        const int K = m.tab.inpad[l];
        hipLaunchKernelGGL(k_gemm_tn, dim3((K / kRows) * S), dim3(256), 0, m.stream, K, m.tab.in[l], chunk, A, cur,
                           m.G, m.P, m.tab.offW[l], m.tab.offb[l]);
      }
      if (l > 0) {
        Prof pf(m, MC_BWD);
        hipLaunchKernelGGL(k_gemm_nt<1>, dim3(M / kRows), dim3(256), 0, m.stream, kW, cur, m.tab.Wk[l], nullptr,
                           m.Z[l - 1], oth, nullptr);
        std::swap(cur, oth);
      }
    }
  }
  MLPC(hipGetLastError());
  return 0;
}

int launch_adam(fvs_mlp &m, int mode, int S, double inv_no, double *lossh, float *gout) {
  Prof pf(m, MC_ADAM);
  float bc1 = 1.f, bc2 = 1.f;
  if (mode == 2) {
    m.t += 1;
    bc1 = float(1.0 - std::pow(m.opt.beta1, double(m.t)));
    bc2 = float(1.0 - std::pow(m.opt.beta2, double(m.t)));
  }
  hipLaunchKernelGGL(k_adam, dim3(unsigned((m.P + 255) / 256)), dim3(256), 0, m.stream, m.P, S, mode, m.G, m.w, m.m1,
                     m.m2, float(m.opt.lr), float(m.opt.beta1), float(m.opt.beta2), float(m.opt.eps), bc1, bc2, m.tab,
                     gout, m.lossp, inv_no, lossh);
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

int fvs_mlp_create(fvs_mlp **out, const fvs_mlp_options *opt) {
  if (!out || !opt) return fail(FVS_EINVAL, "null argument");
  *out = nullptr;
  if (opt->depth < 1 || opt->depth > kMaxDepth) return fail(FVS_EINVAL, "depth must be 1..8");
  if (opt->batch < kRows || opt->batch % kRows) return fail(FVS_EINVAL, "batch must be a positive multiple of 64");
  if (!(opt->lr > 0) || !(opt->beta1 >= 0 && opt->beta1 < 1) || !(opt->beta2 >= 0 && opt->beta2 < 1) ||
      !(opt->eps > 0))
    return fail(FVS_EINVAL, "Adam parameters out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FVS_EHIP, "no HIP device");
  auto *m = new fvs_mlp();
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
  auto bail = [&](int rc) {
    delete m;
    return rc;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(FVS_EHIP, "stream"));
  if (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)
    return bail(fail(FVS_EHIP, "event"));
  const size_t M = size_t(m->M);
  int rc = 0;
  rc = rc ? rc : m->alloc(m->w, m->P);
  rc = rc ? rc : m->alloc(m->m1, m->P);
  rc = rc ? rc : m->alloc(m->m2, m->P);
  rc = rc ? rc : m->alloc(m->gtmp, m->P);
  rc = rc ? rc : m->alloc(m->G, size_t(kMaxSplit) * m->P);
  rc = rc ? rc : m->alloc(m->lossp, kMaxSplit);
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
  if (rc) return bail(rc);
  tb.Wt[m->L] = tb.Wk[m->L] = nullptr;
  tb.Wk[0] = nullptr;
  if (hipMemsetAsync(tb.Wt[0], 0, size_t(kW) * kInPad * sizeof(u16), m->stream) != hipSuccess)
    return bail(fail(FVS_EHIP, "memset"));
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
  *out = m;
  rc = fvs_mlp_set_params(m, w.data());
  if (rc) {
    *out = nullptr;
    return bail(rc);
  }
  return 0;
}

int fvs_mlp_destroy(fvs_mlp *m) {
  delete m;
  return 0;
}

long long fvs_mlp_nparams(const fvs_mlp *m) { return m ? m->P : -1; }

int fvs_mlp_get_params(fvs_mlp *m, float *p) {
  if (!m || !p) return fail(FVS_EINVAL, "null argument");
  MLPC(hipMemcpyAsync(p, m->w, size_t(m->P) * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLPC(hipStreamSynchronize(m->stream));
  return 0;
}

int fvs_mlp_set_params(fvs_mlp *m, const float *p) {
  if (!m || !p) return fail(FVS_EINVAL, "null argument");
  MLPC(hipMemcpyAsync(m->w, p, size_t(m->P) * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLPC(hipMemsetAsync(m->m1, 0, size_t(m->P) * sizeof(float), m->stream));
  MLPC(hipMemsetAsync(m->m2, 0, size_t(m->P) * sizeof(float), m->stream));
  m->t = 0;
  if (int rc = launch_adam(*m, 0, 0, 0.0, nullptr, nullptr)) return rc;
  MLPC(hipStreamSynchronize(m->stream));
  return 0;
}

int fvs_mlp_features(fvs_mlp *m, int n, const float *raw, float *feat) {
  if (!m || (n > 0 && (!raw || !feat))) return fail(FVS_EINVAL, "null argument");
  if (n < 0 || n > m->M) return fail(FVS_EINVAL, "n must be 0..batch");
  if (n == 0) return 0;
  if (int rc = stage_rows(*m, n, raw, nullptr)) return rc;
  launch_features(*m, round64(n), n, m->iraw, nullptr, 0, 0, 0, m->F32);
  MLPC(hipGetLastError());
  MLPC(hipMemcpyAsync(feat, m->F32, size_t(n) * kFeat * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLPC(hipStreamSynchronize(m->stream));
  return 0;
}

int fvs_mlp_set_data(fvs_mlp *m, long long n, const float *raw, const float *target) {
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
}

int fvs_mlp_train(fvs_mlp *m, int steps, double *loss) {
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
    if (!rc) rc = launch_adam(*m, 2, S, inv_no, dl + s, nullptr);
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
  if (int rc = launch_adam(*m, update ? 2 : 1, S, 1.0 / (double(n) * kOut), m->lossh, update ? nullptr : m->gtmp))
    return rc;
  if (grad) MLPC(hipMemcpyAsync(grad, m->gtmp, size_t(m->P) * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  double l = 0.0;
  MLPC(hipMemcpyAsync(&l, m->lossh, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  MLPC(hipStreamSynchronize(m->stream));
  if (loss) *loss = l;
  return 0;
}

int fvs_mlp_grad(fvs_mlp *m, int n, const float *raw, const float *target, double *loss, float *grad) {
  return grad_or_step(m, n, raw, target, loss, grad, false);
}

int fvs_mlp_step(fvs_mlp *m, int n, const float *raw, const float *target, double *loss) {
  return grad_or_step(m, n, raw, target, loss, nullptr, true);
}

int fvs_mlp_predict(fvs_mlp *m, long long n, const float *raw, float *out) {
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
}

int fvs_mlp_kernel_times(fvs_mlp *m, int steps, double *ms6, double *flop6, int *launches6) {
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
}

double fvs_mlp_last_elapsed_ms(const fvs_mlp *m) { return m ? double(m->last_ms) : 0.0; }

}  // extern "C"
