// lgame_mlp.hip -- per-column MLP of the radiation surrogate: BF16 MFMA GEMMs with FP32
// accumulation, GELU, mean-squared-error backward pass and FP32 Adam master weights.
//
// Network: n_in -> hidden x n_hidden_layers (GELU) -> n_out.  Dimensions are padded to multiples
// of 64 in device storage (zero weights, zero inputs).  Weights are stored [out][in] (in
// contiguous) with a BF16 copy in that layout and a transposed one; activations [column][feature]
// and, for training, also [feature][column].  Inputs and targets are normalised with
// per-feature means and scales that the caller sets (lgame_mlp_set_norm).
//
// GEMM: every product is C = A B^T with both operands contiguous along k (the forward pass, the
// data gradient through the transposed weights, the weight gradient through the transposed
// activations and output gradients, which the epilogues write), computed by k_gemm_nt with
// gfx950's v_mfma_f32_16x16x32_bf16.  The weight gradient, whose k is the batch, is split along
// k and the partial sums are added in split order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lgame_gpu.h"
#include "lgame_mlp.h"

namespace lgame {

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));  // 16 bytes, kept in registers

enum Epi { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS_F32 = 2, EPI_DGELU = 3 };

// C(m,n) = sum_k A[m*lda + k] * B[n*ldb + k] (both operands k-contiguous), over the K values
// [z*K, (z+1)*K) of split z of a split-K launch.  Epilogues:
//   EPI_F32       C[z*split_stride + m*ldc + n] = acc                       (weight gradient partials)
//   EPI_BIAS_F32  C[m*ldc + n] = acc + bias[n]                              (output layer)
//   EPI_BIAS_GELU z = acc + bias[n]; Z[m*ldz + n] = z (if Z), A[m*ldz + n] = gelu(z), and the
//                 transposed AT[n*ldt + m] = gelu(z) (if AT)                (hidden layers)
//   EPI_DGELU     d = acc * gelu'(Zin[m*ldz + n]); D[m*ldz + n] = d, DT[n*ldt + m] = d (data gradient)
struct NtArgs {
  const uint16_t *A;
  long long lda;
  const uint16_t *B;
  long long ldb;
  int M, N, K;
  float *C;
  long long ldc, split_stride;
  const float *bias;
  uint16_t *Z, *Aout, *AT;
  const uint16_t *Zin;
  uint16_t *D, *DT;
  long long ldz, ldt;
};

constexpr int MAX_SPLITS = 16;  // split-K limit of the weight gradients
constexpr int GK = 64;          // k values per split of the weight gradients: a multiple of the stage depth

// Block tile BM x BN with 4 wavefronts in a 2 x 2 arrangement; a wavefront computes (BM/2) x
// (BN/2) as (BM/32) x (BN/32) v_mfma_f32_16x16x32_bf16 tiles.  K goes in stages of BK values
// staged in LDS (rows of BK + 8 values: the 16-B operand reads of a wavefront hit distinct
// banks); the next stage is loaded from global memory into registers with 16-B loads while the
// current one is multiplied.  Deep stages (BK = 128) keep the number of exposed load latencies
// per block small.
template <int EPI, int BM, int BN, int BK>
__global__ void __launch_bounds__(256) k_gemm_nt(const NtArgs g) {
  constexpr int LD = BK + 8;
  constexpr int TMW = BM / 32, TNW = BN / 32;  // MFMA tiles per wavefront along m and n
  constexpr int CPR = BK / 8;                  // 16-B chunks per row of a stage
  constexpr int CA = BM * CPR / 256, CB = BN * CPR / 256;  // chunks per thread
  constexpr int RS = 256 / CPR;                // rows between a thread's chunks
  __shared__ __attribute__((aligned(16))) uint16_t As[BM * LD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LD];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // Tile of this block.  The 1-D grid is dealt round-robin to the 8 XCDs; XCD k takes a
  // contiguous range of tiles in (split, m, n) order, so the blocks that share an A tile (the n
  // tiles of one m tile) run on one XCD, whose L2 then reads that tile from memory once.
  int bx, by, bz;
  {
    const int nb = gridDim.x, b = blockIdx.x, nxb = g.N / BN, nyb = g.M / BM;
    const int k = b % 8, s = b / 8, q = nb / 8, r = nb % 8;
    const int tt = k * q + (k < r ? k : r) + s;
    bx = tt % nxb;
    by = (tt / nxb) % nyb;
    bz = tt / (nxb * nyb);
  }
  const int m0 = by * BM, n0 = bx * BN;
  const long long kz = (long long)bz * g.K;
  const int wm = (wave >> 1) * (BM / 2), wn = (wave & 1) * (BN / 2);
  float4_t acc[TMW][TNW];
#pragma unroll
  for (int a = 0; a < TMW; ++a)
#pragma unroll
    for (int b = 0; b < TNW; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
  // this thread's chunks of a stage: row crow + RS*q, k values ckc..ckc+7
  const int crow = t / CPR, ckc = (t % CPR) * 8;
  const uint16_t *pa = g.A + (long long)(m0 + crow) * g.lda + kz + ckc;
  const uint16_t *pb = g.B + (long long)(n0 + crow) * g.ldb + kz + ckc;
  const long long sa = RS * g.lda, sb = RS * g.ldb;
  u32x4_t ra[CA], rb[CB];
#pragma unroll
  for (int q = 0; q < CA; ++q) ra[q] = *reinterpret_cast<const u32x4_t *>(pa + q * sa);
#pragma unroll
  for (int q = 0; q < CB; ++q) rb[q] = *reinterpret_cast<const u32x4_t *>(pb + q * sb);
  for (int kt = 0; kt < g.K; kt += BK) {
    __syncthreads();  // the previous stage has been read
#pragma unroll
    for (int q = 0; q < CA; ++q) *reinterpret_cast<u32x4_t *>(&As[(crow + RS * q) * LD + ckc]) = ra[q];
#pragma unroll
    for (int q = 0; q < CB; ++q) *reinterpret_cast<u32x4_t *>(&Bs[(crow + RS * q) * LD + ckc]) = rb[q];
    __syncthreads();
    if (kt + BK < g.K) {  // the next stage, in flight while this one is multiplied
      pa += BK;
      pb += BK;
#pragma unroll
      for (int q = 0; q < CA; ++q) ra[q] = *reinterpret_cast<const u32x4_t *>(pa + q * sa);
#pragma unroll
      for (int q = 0; q < CB; ++q) rb[q] = *reinterpret_cast<const u32x4_t *>(pb + q * sb);
    }
    // lane layout of v_mfma_f32_16x16x32_bf16: row (A) / column (B) lane % 16, k = 8*(lane/16) + 0..7
    const int r = lane & 15, kq = 8 * (lane >> 4);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      bf16x8_t av[TMW], bv[TNW];
#pragma unroll
      for (int a = 0; a < TMW; ++a) av[a] = *reinterpret_cast<const bf16x8_t *>(&As[(wm + 16 * a + r) * LD + ks + kq]);
#pragma unroll
      for (int b = 0; b < TNW; ++b) bv[b] = *reinterpret_cast<const bf16x8_t *>(&Bs[(wn + 16 * b + r) * LD + ks + kq]);
#pragma unroll
      for (int a = 0; a < TMW; ++a)
#pragma unroll
        for (int b = 0; b < TNW; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
  }
  // result layout: lane holds rows 4*(lane/16) + j (j = 0..3) of column lane % 16
#pragma unroll
  for (int a = 0; a < TMW; ++a)
#pragma unroll
    for (int b = 0; b < TNW; ++b) {
      const int mb = m0 + wm + 16 * a + 4 * (lane >> 4);
      const int n = n0 + wn + 16 * b + (lane & 15);
      uint16_t tr[4];  // the lane's four values for the transposed copy (consecutive m)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long m = mb + j;
        const float v = acc[a][b][j];
        if (EPI == EPI_F32) {
          g.C[(long long)bz * g.split_stride + m * g.ldc + n] = v;
        } else if (EPI == EPI_BIAS_F32) {
          g.C[m * g.ldc + n] = v + g.bias[n];
        } else if (EPI == EPI_BIAS_GELU) {
          const float z = v + g.bias[n];
          if (g.Z) g.Z[m * g.ldz + n] = f2bf(z);
          const uint16_t h = f2bf(gelu(z));
          g.Aout[m * g.ldz + n] = h;
          tr[j] = h;
        } else {
          const float z = bf2f(g.Zin[m * g.ldz + n]);
          const uint16_t h = f2bf(v * gelu_grad(z));
          g.D[m * g.ldz + n] = h;
          tr[j] = h;
        }
      }
      uint16_t *T = EPI == EPI_BIAS_GELU ? g.AT : EPI == EPI_DGELU ? g.DT : nullptr;
      if (T) {
        uint2 pk;
        pk.x = (uint32_t)tr[0] | ((uint32_t)tr[1] << 16);
        pk.y = (uint32_t)tr[2] | ((uint32_t)tr[3] << 16);
        *reinterpret_cast<uint2 *>(T + (long long)n * g.ldt + mb) = pk;
      }
    }
}

// sum of the split-K partials in split order: out[i] = sum_z part[z*n + i]
__global__ void k_sum_splits(const float *part, int splits, long long n, float *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int z = 0; z < splits; ++z) s += part[(long long)z * n + i];
  out[i] = s;
}

// row sums of a [rows][ld] BF16 matrix over its first `len` columns (one block per row, fixed
// reduction tree): the bias gradient from the transposed output gradient
__global__ void __launch_bounds__(256) k_rowsum(const uint16_t *T, long long ld, int len, float *out) {
  __shared__ float sh[256];
  const uint16_t *row = T + (long long)blockIdx.x * ld;
  float s = 0.0f;
  // 8 values (16 B) per load; len and ld are multiples of 64
  for (int i = 8 * threadIdx.x; i < len; i += 8 * 256) {
    const uint4 v = *reinterpret_cast<const uint4 *>(row + i);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) s += bf2f((uint16_t)(w[q] & 0xffffu)) + bf2f((uint16_t)(w[q] >> 16));
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// normalised BF16 inputs, padded: X[c][i] = (x - mean)/scale for i < n_in, 0 above
__global__ void k_norm_in(const float *x, long long ncol, int n_in, int ldi, const float *mean, const float *scale, uint16_t *X) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncol * ldi) return;
  const long long c = idx / ldi;
  const int i = (int)(idx % ldi);
  X[idx] = i < n_in ? f2bf((x[c * n_in + i] - mean[i]) / scale[i]) : (uint16_t)0;
}

// physical outputs from the network's normalised ones
__global__ void k_denorm_out(const float *Yn, long long ncol, int n_out, int ldo, const float *mean, const float *scale, float *y) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncol * n_out) return;
  const long long c = idx / n_out;
  const int o = (int)(idx % n_out);
  y[idx] = Yn[c * ldo + o] * scale[o] + mean[o];
}

// normalised FP32 targets, padded
__global__ void k_norm_out(const float *y, long long ncol, int n_out, int ldo, const float *mean, const float *scale, float *Yn) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncol * ldo) return;
  const long long c = idx / ldo;
  const int o = (int)(idx % ldo);
  Yn[idx] = o < n_out ? (y[c * n_out + o] - mean[o]) / scale[o] : 0.0f;
}

// mean-squared-error gradient of a batch: dY = 2 (yhat - y) / (B * n_out) on the valid outputs,
// also transposed; per-block partial sums of the squared error
__global__ void __launch_bounds__(256) k_mse_grad(const float *yhat, const float *y, const int *rows, int B, int n_out, int ldo,
                                                  uint16_t *dY, uint16_t *dYT, float *partial) {
  __shared__ float sh[256];
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float se = 0.0f;
  if (idx < (long long)B * ldo) {
    const int b = (int)(idx / ldo), o = (int)(idx % ldo);
    float d = 0.0f;
    if (o < n_out) {
      const long long r = rows ? (long long)rows[b] : b;
      const float e = yhat[(long long)b * ldo + o] - y[r * ldo + o];
      se = e * e;
      d = 2.0f * e / (float)((long long)B * n_out);
    }
    dY[idx] = f2bf(d);
    dYT[(long long)o * B + b] = f2bf(d);  // transposed copy [ldo][B] for the weight gradient
  }
  sh[threadIdx.x] = se;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// acc += sum of the per-block squared-error partials of one batch (one block, fixed order)
__global__ void __launch_bounds__(256) k_loss_add(const float *partial, int n, float *acc) {
  __shared__ float sh[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *acc += sh[0];
}

// dst[c*ldd + r] = src[r*ld + c] for r < rows, c < cols (BF16), through 32x32 LDS tiles
__global__ void __launch_bounds__(256) k_transpose16(const uint16_t *src, long long ld, int rows, int cols, uint16_t *dst, long long ldd) {
  __shared__ uint16_t tile[32][34];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + tx;
    if (r < rows && c < cols) tile[j][tx] = src[(long long)r * ld + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j, r = r0 + tx;
    if (r < rows && c < cols) dst[(long long)c * ldd + r] = tile[tx][j];
  }
}

// A batch of rows of the normalised data set gathered into X [B][ld] and, through 32x32 LDS tiles,
// its transpose XT [ld][ldt] (B and ld multiples of 32)
__global__ void __launch_bounds__(256) k_gather_t(const uint16_t *DX, const int *rows, int ld, uint16_t *X, uint16_t *XT,
                                                  long long ldt) {
  __shared__ uint16_t tile[32][34];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + tx;
    const uint16_t v = DX[(long long)rows[r] * ld + c];
    X[(long long)r * ld + c] = v;
    tile[j][tx] = v;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) XT[(long long)(c0 + j) * ldt + r0 + tx] = tile[tx][j];
}



struct AdamArgs {
  float lr, b1, b2, eps, c1, c2;
};

// Adam (Kingma and Ba) on element i of an FP32 master-weight block, with the bias corrections
// c1 = 1 - b1^t and c2 = 1 - b2^t; returns the new weight
__device__ __forceinline__ float adam_elem(float *w, float *m, float *v, long long i, float gi, const AdamArgs &a) {
  const float mi = a.b1 * m[i] + (1.0f - a.b1) * gi;
  const float vi = a.b2 * v[i] + (1.0f - a.b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
  w[i] = wi;
  return wi;
}

// The weight gradient summed over its split-K partials in split order (as k_sum_splits), then Adam
// on the padded [rows][ld] weight block (entries outside the valid rows/columns stay zero) and its
// BF16 copies [rows][ld] and [ld][ldt]: one pass instead of two.
__global__ void k_adam_splits(const float *part, int splits, long long n, float *w, float *m, float *v, uint16_t *wbf,
                              uint16_t *wbfT, long long ldt, int ld, int valid_cols, int valid_rows, const AdamArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int col = (int)(i % ld);
  const long long row = i / ld;
  if (col >= valid_cols || row >= valid_rows) return;
  float gi = 0.0f;
  for (int z = 0; z < splits; ++z) gi += part[(long long)z * n + i];
  const float wi = adam_elem(w, m, v, i, gi, a);
  wbf[i] = f2bf(wi);
  wbfT[(long long)col * ldt + row] = f2bf(wi);
}

// The bias gradient as the row sum of the transposed output gradient (k_rowsum, one block per
// output, the same reduction tree), then Adam on the bias.
__global__ void __launch_bounds__(256) k_rowsum_adam(const uint16_t *T, long long ld, int len, int valid, float *b, float *mb,
                                                      float *vb, const AdamArgs a) {
  __shared__ float sh[256];
  const uint16_t *row = T + (long long)blockIdx.x * ld;
  float s = 0.0f;
  for (int i = 8 * threadIdx.x; i < len; i += 8 * 256) {
    const uint4 q4 = *reinterpret_cast<const uint4 *>(row + i);
    const uint32_t wq[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) s += bf2f((uint16_t)(wq[q] & 0xffffu)) + bf2f((uint16_t)(wq[q] >> 16));
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && (int)blockIdx.x < valid) (void)adam_elem(b, mb, vb, blockIdx.x, sh[0], a);
}

// BF16 copies of a [rows][ld] FP32 weight block: as it is and transposed ([ld][rows])
__global__ void k_to_bf16(const float *w, uint16_t *wbf, uint16_t *wbfT, int rows, int ld) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * ld) return;
  const long long row = i / ld, col = i % ld;
  wbf[i] = f2bf(w[i]);
  wbfT[col * rows + row] = f2bf(w[i]);
}

}  // namespace lgame

using namespace lgame;

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
struct MlpLayer {
  int in, out;      // logical sizes
  int in_p, out_p;  // padded to 64
  float *w = nullptr, *b = nullptr;          // FP32 master [out_p][in_p], [out_p]
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;  // Adam moments
  float *gw = nullptr, *gb = nullptr;        // gradients
  uint16_t *wbf = nullptr, *wbfT = nullptr;  // BF16 copies [out_p][in_p] and [in_p][out_p]
};

struct lgame_mlp {
  int n_in, n_out, hidden, nh;
  std::vector<MlpLayer> L;
  float *x_mean = nullptr, *x_scale = nullptr, *y_mean = nullptr, *y_scale = nullptr;
  bool has_norm = false;  // lgame_mlp_set_norm was called (the outputs' scales define the envelope)
  int device = 0;         // the GPU the network lives on (current at lgame_mlp_create)
  lgame::MlpWork *host_work = nullptr;  // inference buffers of lgame_mlp_forward
  // activations for up to `cap` rows: per layer Z and A (BF16), gradient buffers
  long long cap = 0;
  std::vector<uint16_t *> Z, A, AT;  // AT: activations transposed, [feature][row]
  uint16_t *X = nullptr, *XT = nullptr, *dA = nullptr, *dB = nullptr, *dYb = nullptr, *dTa = nullptr, *dTb = nullptr;
  float *gpart = nullptr;  // split-K partials of a weight gradient
  float *Yf = nullptr, *Yt = nullptr, *partial = nullptr, *xin = nullptr, *yout = nullptr;
  int *rows = nullptr;
  // data set (normalised, device)
  uint16_t *DX = nullptr;
  float *DY = nullptr;
  long long ndata = 0;
  long long adam_t = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

namespace lgame {
// Buffers of one inference pass of up to `rows` rows: the normalised input (BF16, padded width),
// every hidden layer's activations (BF16) and the outputs (FP32, padded width).
struct InferBufs {
  long long rows = 0;
  uint16_t *X = nullptr;
  uint16_t *A[8] = {};
  float *Yf = nullptr;
};
// Inference buffers of one user of the network, allocated by mlp_work_create.
struct MlpWork {
  InferBufs b;
};
}  // namespace lgame

namespace {

int mfail(lgame_mlp *m, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (m) m->err = buf;
  return code;
}

#define MCHK(m, call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return mfail(m, LGAME_EHIP, "%s: %s", #call, hipGetErrorString(e_));  \
  } while (0)

inline int pad64(int n) { return (n + 63) / 64 * 64; }
inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

void free_training_buffers(lgame_mlp *m) {
  for (auto p : m->Z) (void)hipFree(p);
  for (auto p : m->A) (void)hipFree(p);
  for (auto p : m->AT) (void)hipFree(p);
  m->Z.clear();
  m->A.clear();
  m->AT.clear();
  for (void **p : {(void **)&m->X, (void **)&m->XT, (void **)&m->dA, (void **)&m->dB, (void **)&m->dYb, (void **)&m->dTa,
                   (void **)&m->dTb, (void **)&m->gpart, (void **)&m->Yf, (void **)&m->Yt, (void **)&m->partial,
                   (void **)&m->xin, (void **)&m->yout, (void **)&m->rows}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  m->cap = 0;
}

// Activation and gradient buffers of training for `rows` rows.  On a failed allocation every
// buffer is freed and the capacity is 0, so no later call uses a stale or freed buffer.
int ensure_cap(lgame_mlp *m, long long rows) {
  if (rows <= m->cap) return 0;
  free_training_buffers(m);
  const int hp = pad64(m->hidden), ip = pad64(m->n_in), op = pad64(m->n_out);
  const int wmax = hp > ip ? hp : ip;
  bool ok = true;
  auto alloc = [&](void **p, size_t bytes) {
    if (ok && hipMalloc(p, bytes) != hipSuccess) {
      *p = nullptr;
      ok = false;
    }
  };
  for (int l = 0; l < m->nh && ok; ++l) {
    uint16_t *z = nullptr, *a = nullptr, *at = nullptr;
    alloc((void **)&z, (size_t)rows * hp * 2);
    alloc((void **)&a, (size_t)rows * hp * 2);
    alloc((void **)&at, (size_t)rows * hp * 2);
    m->Z.push_back(z);
    m->A.push_back(a);
    m->AT.push_back(at);
  }
  alloc((void **)&m->X, (size_t)rows * ip * 2);
  alloc((void **)&m->XT, (size_t)rows * ip * 2);
  alloc((void **)&m->dA, (size_t)rows * wmax * 2);
  alloc((void **)&m->dB, (size_t)rows * wmax * 2);
  alloc((void **)&m->dTa, (size_t)rows * wmax * 2);
  alloc((void **)&m->dTb, (size_t)rows * wmax * 2);
  alloc((void **)&m->gpart, (size_t)MAX_SPLITS * wmax * wmax * 4);
  alloc((void **)&m->dYb, (size_t)rows * op * 2);
  alloc((void **)&m->Yf, (size_t)rows * op * 4);
  alloc((void **)&m->Yt, (size_t)rows * op * 4);
  alloc((void **)&m->partial, (size_t)nblk(rows * op) * 4);
  alloc((void **)&m->xin, (size_t)rows * m->n_in * 4);
  alloc((void **)&m->yout, (size_t)rows * m->n_out * 4);
  alloc((void **)&m->rows, (size_t)rows * 4);
  if (ok) ok = hipMemset(m->X, 0, (size_t)rows * ip * 2) == hipSuccess;
  if (!ok) {
    free_training_buffers(m);
    return mfail(m, LGAME_ENOMEM, "activation buffers for %lld rows", rows);
  }
  m->cap = rows;
  return 0;
}

template <int EPI, int BK>
void gemm_launch(hipStream_t st, const NtArgs &g, dim3 grid, bool m128, bool n128) {
  if (m128 && n128)
    hipLaunchKernelGGL((k_gemm_nt<EPI, 128, 128, BK>), grid, dim3(256), 0, st, g);
  else if (m128)
    hipLaunchKernelGGL((k_gemm_nt<EPI, 128, 64, BK>), grid, dim3(256), 0, st, g);
  else if (n128)
    hipLaunchKernelGGL((k_gemm_nt<EPI, 64, 128, BK>), grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL((k_gemm_nt<EPI, 64, 64, BK>), grid, dim3(256), 0, st, g);
}

// C = A B^T with both operands k-contiguous; M, N multiples of 64, K (per split) of 64.  The
// largest tiles that still give at least 4 blocks per CU (1024 blocks): a stage's global loads
// are in flight for one stage of multiplication only, so the other blocks of a CU hide them.
template <int EPI>
void gemm_nt(hipStream_t st, const NtArgs &g, int splits = 1) {
  const long long want = 1024;
  auto nblocks = [&](int bm, int bn) { return (long long)(g.M / bm) * (g.N / bn) * splits; };
  bool m128 = g.M % 128 == 0, n128 = g.N % 128 == 0;
  if (m128 && n128 && nblocks(128, 128) < want) {  // drop to a 64-wide side first
    if (nblocks(128, 64) >= want || nblocks(64, 128) >= want) {
      if (nblocks(128, 64) >= want) n128 = false; else m128 = false;
    } else {
      m128 = n128 = false;
    }
  } else if ((m128 != n128) && nblocks(m128 ? 128 : 64, n128 ? 128 : 64) < want) {
    m128 = n128 = false;
  }
  const dim3 grid((unsigned)((g.N / (n128 ? 128 : 64)) * (g.M / (m128 ? 128 : 64)) * splits));  // tiles, see k_gemm_nt
  if (g.K % 128 == 0)
    gemm_launch<EPI, 128>(st, g, grid, m128, n128);
  else
    gemm_launch<EPI, 64>(st, g, grid, m128, n128);
}

// forward pass of B normalised rows X (padded) -> Yf (FP32, padded); for training it keeps the
// pre-activations Z and the activations in both layouts (A [row][feature], AT [feature][row])
void forward(lgame_mlp *m, const uint16_t *X, int B, bool train) {
  const uint16_t *in = X;
  int in_p = pad64(m->n_in);
  for (int l = 0; l <= m->nh; ++l) {
    const MlpLayer &Ly = m->L[l];
    NtArgs g{};
    g.A = in;
    g.lda = in_p;
    g.B = Ly.wbf;
    g.ldb = Ly.in_p;
    g.M = B;
    g.N = Ly.out_p;
    g.K = Ly.in_p;
    g.bias = Ly.b;
    if (l < m->nh) {
      g.Z = train ? m->Z[l] : nullptr;
      g.Aout = m->A[l];
      g.AT = train ? m->AT[l] : nullptr;
      g.ldz = Ly.out_p;
      g.ldt = B;
      gemm_nt<EPI_BIAS_GELU>(m->stream, g);
      in = m->A[l];
      in_p = Ly.out_p;
    } else {
      g.C = m->Yf;
      g.ldc = Ly.out_p;
      gemm_nt<EPI_BIAS_F32>(m->stream, g);
    }
  }
}

// the forward pass of inference on a workspace's buffers, enqueued on `st`
void forward_infer(const lgame_mlp *m, const InferBufs *w, int B, hipStream_t st) {
  const uint16_t *in = w->X;
  int in_p = pad64(m->n_in);
  for (int l = 0; l <= m->nh; ++l) {
    const MlpLayer &Ly = m->L[l];
    NtArgs g{};
    g.A = in;
    g.lda = in_p;
    g.B = Ly.wbf;
    g.ldb = Ly.in_p;
    g.M = B;
    g.N = Ly.out_p;
    g.K = Ly.in_p;
    g.bias = Ly.b;
    if (l < m->nh) {
      g.Aout = w->A[l];
      g.ldz = Ly.out_p;
      g.ldt = B;
      gemm_nt<EPI_BIAS_GELU>(st, g);
      in = w->A[l];
      in_p = Ly.out_p;
    } else {
      g.C = w->Yf;
      g.ldc = Ly.out_p;
      gemm_nt<EPI_BIAS_F32>(st, g);
    }
  }
}

// backward pass from the output gradient (dYb [B][out_p] and dTa = its transpose) with the
// activations of the last training forward pass; XT is the transposed input
// With `adam` the weight and bias gradients are not stored: each layer's parameters take their
// Adam update as soon as its gradients are formed and its old weights are no longer needed (after
// the data gradient through them).
void backward(lgame_mlp *m, const uint16_t *XT, int B, const AdamArgs *adam = nullptr) {
  const uint16_t *dcur = m->dYb, *dcurT = m->dTa;
  // split-K of the weight gradients (K = B): enough blocks to fill the chip, >= 256 k values each
  int splits = 1;
  while (splits < MAX_SPLITS && B / (2 * splits) >= 256 && (B / (2 * splits)) % GK == 0) splits *= 2;
  for (int l = m->nh; l >= 0; --l) {
    MlpLayer &Ly = m->L[l];
    const uint16_t *inT = l == 0 ? XT : m->AT[l - 1];
    // weight gradient: gW[o][i] = sum_b d[b][o] * in[b][i] = sum_b dT[o][b] * inT[i][b]
    NtArgs gw{};
    gw.A = dcurT;
    gw.lda = B;
    gw.B = inT;
    gw.ldb = B;
    gw.M = Ly.out_p;
    gw.N = Ly.in_p;
    gw.K = B / splits;
    gw.C = m->gpart;
    gw.ldc = Ly.in_p;
    gw.split_stride = (long long)Ly.out_p * Ly.in_p;
    gemm_nt<EPI_F32>(m->stream, gw, splits);
    const long long nw = gw.split_stride;
    auto update_weights = [&]() {
      hipLaunchKernelGGL(k_adam_splits, dim3(nblk(nw)), dim3(256), 0, m->stream, (const float *)m->gpart, splits, nw, Ly.w, Ly.mw,
                         Ly.vw, Ly.wbf, Ly.wbfT, (long long)Ly.out_p, Ly.in_p, Ly.in, Ly.out, *adam);
    };
    if (adam) {
      hipLaunchKernelGGL(k_rowsum_adam, dim3(Ly.out_p), dim3(256), 0, m->stream, dcurT, (long long)B, B, Ly.out, Ly.b, Ly.mb, Ly.vb,
                         *adam);
    } else {
      hipLaunchKernelGGL(k_sum_splits, dim3(nblk(nw)), dim3(256), 0, m->stream, m->gpart, splits, nw, Ly.gw);
      hipLaunchKernelGGL(k_rowsum, dim3(Ly.out_p), dim3(256), 0, m->stream, dcurT, (long long)B, B, Ly.gb);
    }
    if (l == 0) {
      if (adam) update_weights();
      break;
    }
    // data gradient through the GELU of layer l-1: d_prev[b][i] = (sum_o d[b][o] W[o][i]) gelu'(Z)
    uint16_t *dnext = (dcur == m->dA) ? m->dB : m->dA;
    uint16_t *dnextT = (dcurT == m->dTa) ? m->dTb : m->dTa;
    NtArgs gd{};
    gd.A = dcur;
    gd.lda = Ly.out_p;
    gd.B = Ly.wbfT;
    gd.ldb = Ly.out_p;
    gd.M = B;
    gd.N = Ly.in_p;
    gd.K = Ly.out_p;
    gd.Zin = m->Z[l - 1];
    gd.D = dnext;
    gd.DT = dnextT;
    gd.ldz = Ly.in_p;
    gd.ldt = B;
    gemm_nt<EPI_DGELU>(m->stream, gd);
    if (adam) update_weights();  // the data gradient has read this layer's old weights
    dcur = dnext;
    dcurT = dnextT;
  }
}

float loss_sum(lgame_mlp *m, int B) {
  const int op = pad64(m->n_out);
  std::vector<float> p(nblk((long long)B * op));
  if (hipMemcpyAsync(p.data(), m->partial, p.size() * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess) return NAN;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return NAN;
  double s = 0.0;
  for (float v : p) s += v;
  return (float)s;
}

}  // namespace

namespace lgame {

// Device-side inference used by the radiation surrogate: x (ncol x n_in, FP32, device) ->
// y (ncol x n_out, FP32, device), in chunks of the activation capacity, on `stream`.
}  // namespace lgame

// ---- the scratch of lgame_mlp_forward_noalloc ------------------------------------------------
// One inference pass of R rows needs the FP32 input and output rows, the normalised BF16 input,
// every hidden layer's BF16 activations and the padded FP32 outputs; R is the largest multiple of
// 64 (at most 4096) that fits.  Every GPU the library runs on holds one copy of this array until
// the process ends; calls that overlap on one GPU share it and race.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[MAX_DEVICES] = {};  // the array's address on each device

// The device address of `scratch` on the current device, looked up once per device with that
// device current (the array's name is only a host-side stand-in).
int scratch_address(unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return LGAME_EHIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LGAME_EHIP;
    g_scratch_dev[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

__global__ void k_zero_u16(uint16_t *p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

namespace lgame {

void mlp_work_destroy(MlpWork *w) {
  if (!w) return;
  for (auto p : w->b.A)
    if (p) (void)hipFree(p);
  if (w->b.X) (void)hipFree(w->b.X);
  if (w->b.Yf) (void)hipFree(w->b.Yf);
  delete w;
}

int mlp_work_create(const lgame_mlp *m, long long rows, MlpWork **out) {
  *out = nullptr;
  rows = (rows + 63) / 64 * 64;
  if (rows < 64) rows = 64;
  MlpWork *w = new MlpWork();
  w->b.rows = rows;
  const int hp = pad64(m->hidden), ip = pad64(m->n_in), op = pad64(m->n_out);
  bool ok = hipMalloc(&w->b.X, (size_t)rows * ip * 2) == hipSuccess;
  // the padding columns of the input stay zero (k_norm_in writes the first n_in of each row)
  ok = ok && hipMemset(w->b.X, 0, (size_t)rows * ip * 2) == hipSuccess;
  for (int l = 0; l < m->nh && ok; ++l) ok = hipMalloc(&w->b.A[l], (size_t)rows * hp * 2) == hipSuccess;
  ok = ok && hipMalloc(&w->b.Yf, (size_t)rows * op * 4) == hipSuccess;
  if (!ok) {
    mlp_work_destroy(w);
    return LGAME_ENOMEM;
  }
  *out = w;
  return 0;
}

// x (ncol x n_in, FP32, device) -> y (ncol x n_out, FP32, device) on `stream`, in passes of the
// workspace's rows.  Allocates nothing; the network's weights are only read, so users with
// their own workspaces may run it at the same time on different streams.
static int infer_passes(const lgame_mlp *m, const InferBufs *w, const float *x, long long ncol, float *y, hipStream_t stream) {
  const long long chunk = w->rows;
  const int ip = pad64(m->n_in), op = pad64(m->n_out);
  for (long long c0 = 0; c0 < ncol; c0 += chunk) {
    const long long nc = ncol - c0 < chunk ? ncol - c0 : chunk;
    const int B = (int)((nc + 63) / 64 * 64);
    // rows beyond nc hold earlier inputs (or zeros): finite, and their outputs are not copied
    hipLaunchKernelGGL(k_norm_in, dim3(nblk(nc * ip)), dim3(256), 0, stream, x + c0 * m->n_in, nc, m->n_in, ip, m->x_mean,
                       m->x_scale, w->X);
    forward_infer(m, w, B, stream);
    hipLaunchKernelGGL(k_denorm_out, dim3(nblk(nc * m->n_out)), dim3(256), 0, stream, w->Yf, nc, m->n_out, op, m->y_mean,
                       m->y_scale, y + c0 * m->n_out);
  }
  return hipGetLastError() == hipSuccess ? 0 : LGAME_EHIP;
}

int mlp_infer_device(const lgame_mlp *m, MlpWork *w, const float *x, long long ncol, float *y, hipStream_t stream) {
  return infer_passes(m, &w->b, x, ncol, y, stream);
}

bool mlp_has_norm(const lgame_mlp *m) { return m->has_norm; }

int mlp_dims(const lgame_mlp *m, int *n_in, int *n_out) {
  *n_in = m->n_in;
  *n_out = m->n_out;
  return 0;
}

void mlp_out_norm(const lgame_mlp *m, const float **mean, const float **scale) {
  *mean = m->y_mean;
  *scale = m->y_scale;
}

}  // namespace lgame

extern "C" {

int lgame_mlp_create(int n_in, int hidden, int n_hidden_layers, int n_out, unsigned long long seed, lgame_mlp **out) {
  if (!out) return LGAME_EINVAL;
  *out = nullptr;
  if (n_in < 1 || hidden < 1 || n_hidden_layers < 1 || n_out < 1 || n_hidden_layers > 8) return LGAME_EINVAL;
  lgame_mlp *m = new lgame_mlp();
  if (hipGetDevice(&m->device) != hipSuccess) {
    delete m;
    return LGAME_EHIP;
  }
  m->n_in = n_in;
  m->n_out = n_out;
  m->hidden = hidden;
  m->nh = n_hidden_layers;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return LGAME_EHIP;
  }
  // He-style uniform initialisation from a 64-bit LCG (reproducible across platforms)
  unsigned long long st = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto urand = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l <= m->nh; ++l) {
    MlpLayer Ly;
    Ly.in = l == 0 ? n_in : hidden;
    Ly.out = l == m->nh ? n_out : hidden;
    Ly.in_p = pad64(Ly.in);
    Ly.out_p = pad64(Ly.out);
    const size_t nw = (size_t)Ly.in_p * Ly.out_p, nb = Ly.out_p;
    std::vector<float> w(nw, 0.0f), b(nb, 0.0f);
    const double lim = std::sqrt(6.0 / Ly.in) * (l == m->nh ? 0.5 : 1.0);
    for (int o = 0; o < Ly.out; ++o)
      for (int i = 0; i < Ly.in; ++i) w[(size_t)o * Ly.in_p + i] = (float)((2.0 * urand() - 1.0) * lim);
    bool ok = true;
    for (float **p : {&Ly.w, &Ly.mw, &Ly.vw, &Ly.gw}) ok = ok && hipMalloc(p, nw * 4) == hipSuccess;
    for (float **p : {&Ly.b, &Ly.mb, &Ly.vb, &Ly.gb}) ok = ok && hipMalloc(p, nb * 4) == hipSuccess;
    ok = ok && hipMalloc(&Ly.wbf, nw * 2) == hipSuccess;
    ok = ok && hipMalloc(&Ly.wbfT, nw * 2) == hipSuccess;
    ok = ok && hipMemcpy(Ly.w, w.data(), nw * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(Ly.b, b.data(), nb * 4, hipMemcpyHostToDevice) == hipSuccess;
    for (float *p : {Ly.mw, Ly.vw, Ly.gw}) ok = ok && hipMemset(p, 0, nw * 4) == hipSuccess;
    for (float *p : {Ly.mb, Ly.vb, Ly.gb}) ok = ok && hipMemset(p, 0, nb * 4) == hipSuccess;
    m->L.push_back(Ly);
    if (!ok) {
      lgame_mlp_destroy(m);
      return LGAME_ENOMEM;
    }
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(nw)), dim3(256), 0, m->stream, Ly.w, Ly.wbf, Ly.wbfT, Ly.out_p, Ly.in_p);
  }
  bool ok = true;
  for (float **p : {&m->x_mean, &m->x_scale}) ok = ok && hipMalloc(p, (size_t)n_in * 4) == hipSuccess;
  for (float **p : {&m->y_mean, &m->y_scale}) ok = ok && hipMalloc(p, (size_t)n_out * 4) == hipSuccess;
  if (!ok) {
    lgame_mlp_destroy(m);
    return LGAME_ENOMEM;
  }
  std::vector<float> zeros((size_t)(n_in > n_out ? n_in : n_out), 0.0f), ones(zeros.size(), 1.0f);
  (void)hipMemcpy(m->x_mean, zeros.data(), (size_t)n_in * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->x_scale, ones.data(), (size_t)n_in * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->y_mean, zeros.data(), (size_t)n_out * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->y_scale, ones.data(), (size_t)n_out * 4, hipMemcpyHostToDevice);
  if (hipStreamSynchronize(m->stream) != hipSuccess) {
    lgame_mlp_destroy(m);
    return LGAME_EHIP;
  }
  *out = m;
  return LGAME_OK;
}

long long lgame_mlp_n_params(const lgame_mlp *m) {
  if (!m) return -1;
  long long n = 0;
  for (auto &Ly : m->L) n += (long long)Ly.in * Ly.out + Ly.out;
  return n;
}

// parameters in the order W_1 [out][in], b_1, W_2, b_2, ... (unpadded, FP32)
static int params_io(lgame_mlp *m, float *p, long long n, bool set) {
  if (!m || !p) return LGAME_EINVAL;
  if (n != lgame_mlp_n_params(m)) return mfail(m, LGAME_EINVAL, "expected %lld parameters, got %lld", lgame_mlp_n_params(m), n);
  long long off = 0;
  for (auto &Ly : m->L) {
    std::vector<float> w((size_t)Ly.in_p * Ly.out_p, 0.0f), b(Ly.out_p, 0.0f);
    if (!set) {
      MCHK(m, hipMemcpy(w.data(), Ly.w, w.size() * 4, hipMemcpyDeviceToHost));
      MCHK(m, hipMemcpy(b.data(), Ly.b, b.size() * 4, hipMemcpyDeviceToHost));
    }
    for (int o = 0; o < Ly.out; ++o)
      for (int i = 0; i < Ly.in; ++i) {
        float &x = w[(size_t)o * Ly.in_p + i];
        if (set) x = p[off]; else p[off] = x;
        ++off;
      }
    for (int o = 0; o < Ly.out; ++o) {
      if (set) b[o] = p[off]; else p[off] = b[o];
      ++off;
    }
    if (set) {
      MCHK(m, hipMemcpy(Ly.w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
      MCHK(m, hipMemcpy(Ly.b, b.data(), b.size() * 4, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_to_bf16, dim3(nblk((long long)w.size())), dim3(256), 0, m->stream, Ly.w, Ly.wbf, Ly.wbfT, Ly.out_p,
                         Ly.in_p);
      MCHK(m, hipStreamSynchronize(m->stream));
    }
  }
  if (set) m->adam_t = 0;
  return LGAME_OK;
}

int lgame_mlp_set_params(lgame_mlp *m, const float *p, long long n) { return params_io(m, const_cast<float *>(p), n, true); }
int lgame_mlp_get_params(lgame_mlp *m, float *p, long long n) { return params_io(m, p, n, false); }

int lgame_mlp_set_norm(lgame_mlp *m, const float *x_mean, const float *x_scale, const float *y_mean, const float *y_scale) {
  if (!m || !x_mean || !x_scale || !y_mean || !y_scale) return LGAME_EINVAL;
  for (int i = 0; i < m->n_in; ++i)
    if (!(x_scale[i] > 0.0f)) return mfail(m, LGAME_EINVAL, "x_scale[%d] must be positive", i);
  for (int i = 0; i < m->n_out; ++i)
    if (!(y_scale[i] > 0.0f)) return mfail(m, LGAME_EINVAL, "y_scale[%d] must be positive", i);
  MCHK(m, hipMemcpy(m->x_mean, x_mean, (size_t)m->n_in * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->x_scale, x_scale, (size_t)m->n_in * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->y_mean, y_mean, (size_t)m->n_out * 4, hipMemcpyHostToDevice));
  MCHK(m, hipMemcpy(m->y_scale, y_scale, (size_t)m->n_out * 4, hipMemcpyHostToDevice));
  m->has_norm = true;
  return LGAME_OK;
}

int lgame_mlp_forward(lgame_mlp *m, const float *x, long long n, float *y) {
  if (!m || !x || !y || n < 0) return LGAME_EINVAL;
  if (n == 0) return LGAME_OK;
  const long long chunk = 65536;
  if (ensure_cap(m, chunk)) return LGAME_ENOMEM;
  if (!m->host_work && mlp_work_create(m, chunk, &m->host_work)) return mfail(m, LGAME_ENOMEM, "inference buffers");
  for (long long c0 = 0; c0 < n; c0 += chunk) {
    const long long nc = n - c0 < chunk ? n - c0 : chunk;
    MCHK(m, hipMemcpyAsync(m->xin, x + c0 * m->n_in, (size_t)nc * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
    int rc = mlp_infer_device(m, m->host_work, m->xin, nc, m->yout, m->stream);
    if (rc) return mfail(m, rc, "inference failed");
    MCHK(m, hipMemcpyAsync(y + c0 * m->n_out, m->yout, (size_t)nc * m->n_out * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
  }
  return LGAME_OK;
}

// lgame_mlp_forward without any allocation: the passes run in the static scratch array above.
int lgame_mlp_forward_noalloc(lgame_mlp *m, const float *x, long long n, float *y) {
  if (!m || !x || !y || n < 0) return LGAME_EINVAL;
  if (n == 0) return LGAME_OK;
  int dev = -1;
  MCHK(m, hipGetDevice(&dev));
  if (dev != m->device) return mfail(m, LGAME_EINVAL, "call with the network's device (%d) current, not %d", m->device, dev);
  unsigned char *base = nullptr;
  if (scratch_address(&base)) return mfail(m, LGAME_EHIP, "device address of the inference scratch");
  const int ip = pad64(m->n_in), hp = pad64(m->hidden), op = pad64(m->n_out);
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  auto need = [&](long long R) {
    return al((size_t)R * m->n_in * 4) + al((size_t)R * m->n_out * 4) + al((size_t)R * ip * 2) +
           (size_t)m->nh * al((size_t)R * hp * 2) + al((size_t)R * op * 4);
  };
  long long R = 4096;
  while (R >= 64 && need(R) > NOALLOC_SCRATCH_BYTES) R -= 64;
  if (R < 64)
    return mfail(m, LGAME_EINVAL, "the network needs %zu bytes of scratch for 64 rows; lgame_mlp_forward_noalloc has %zu",
                 need(64), NOALLOC_SCRATCH_BYTES);
  unsigned char *q = base;
  auto carve = [&](size_t bytes) {
    unsigned char *r = q;
    q += al(bytes);
    return r;
  };
  float *xin = (float *)carve((size_t)R * m->n_in * 4);
  float *yout = (float *)carve((size_t)R * m->n_out * 4);
  InferBufs b;
  b.rows = R;
  b.X = (uint16_t *)carve((size_t)R * ip * 2);
  for (int l = 0; l < m->nh; ++l) b.A[l] = (uint16_t *)carve((size_t)R * hp * 2);
  b.Yf = (float *)carve((size_t)R * op * 4);
  // the scratch keeps the last call's contents: the input's padding columns must start at zero
  hipLaunchKernelGGL(k_zero_u16, dim3(nblk((long long)R * ip)), dim3(256), 0, m->stream, b.X, (long long)R * ip);
  for (long long c0 = 0; c0 < n; c0 += R) {
    const long long nc = n - c0 < R ? n - c0 : R;
    MCHK(m, hipMemcpyAsync(xin, x + c0 * m->n_in, (size_t)nc * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
    int rc = infer_passes(m, &b, xin, nc, yout, m->stream);
    if (rc) return mfail(m, rc, "inference failed");
    MCHK(m, hipMemcpyAsync(y + c0 * m->n_out, yout, (size_t)nc * m->n_out * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
  }
  return LGAME_OK;
}

// Gradient of the mean-squared error (normalised outputs) over n rows, parameters in the order of
// lgame_mlp_get_params; *loss receives the mean squared error.  n must be a multiple of 64.
int lgame_mlp_gradient(lgame_mlp *m, const float *x, const float *y, long long n, float *grad, float *loss) {
  if (!m || !x || !y || !grad || !loss || n <= 0 || n % 64) return LGAME_EINVAL;
  if (ensure_cap(m, n)) return mfail(m, LGAME_ENOMEM, "activation buffers");
  const int ip = pad64(m->n_in), op = pad64(m->n_out);
  float *dy;
  MCHK(m, hipMalloc(&dy, (size_t)n * m->n_out * 4));
  MCHK(m, hipMemcpyAsync(m->xin, x, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(dy, y, (size_t)n * m->n_out * 4, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_norm_in, dim3(nblk(n * ip)), dim3(256), 0, m->stream, m->xin, n, m->n_in, ip, m->x_mean, m->x_scale, m->X);
  hipLaunchKernelGGL(k_norm_out, dim3(nblk(n * op)), dim3(256), 0, m->stream, dy, n, m->n_out, op, m->y_mean, m->y_scale, m->Yt);
  hipLaunchKernelGGL(k_transpose16, dim3(ip / 32, (unsigned)((n + 31) / 32)), dim3(256), 0, m->stream, m->X, (long long)ip, (int)n, ip,
                     m->XT, (long long)n);
  forward(m, m->X, (int)n, true);
  hipLaunchKernelGGL(k_mse_grad, dim3(nblk(n * op)), dim3(256), 0, m->stream, m->Yf, m->Yt, (const int *)nullptr, (int)n, m->n_out,
                     op, m->dYb, m->dTa, m->partial);
  backward(m, m->XT, (int)n);
  *loss = loss_sum(m, (int)n) / (float)(n * m->n_out);
  (void)hipFree(dy);
  long long off = 0;
  for (auto &Ly : m->L) {
    std::vector<float> w((size_t)Ly.in_p * Ly.out_p), b(Ly.out_p);
    MCHK(m, hipMemcpy(w.data(), Ly.gw, w.size() * 4, hipMemcpyDeviceToHost));
    MCHK(m, hipMemcpy(b.data(), Ly.gb, b.size() * 4, hipMemcpyDeviceToHost));
    for (int o = 0; o < Ly.out; ++o)
      for (int i = 0; i < Ly.in; ++i) grad[off++] = w[(size_t)o * Ly.in_p + i];
    for (int o = 0; o < Ly.out; ++o) grad[off++] = b[o];
  }
  return LGAME_OK;
}

int lgame_mlp_load_data(lgame_mlp *m, const float *x, const float *y, long long n) {
  if (!m || !x || !y || n < 64) return LGAME_EINVAL;
  const int ip = pad64(m->n_in), op = pad64(m->n_out);
  if (m->DX) (void)hipFree(m->DX);
  if (m->DY) (void)hipFree(m->DY);
  m->DX = nullptr;
  m->DY = nullptr;
  MCHK(m, hipMalloc(&m->DX, (size_t)n * ip * 2));
  MCHK(m, hipMalloc(&m->DY, (size_t)n * op * 4));
  const long long chunk = 65536;
  float *tx, *ty;
  MCHK(m, hipMalloc(&tx, (size_t)chunk * m->n_in * 4));
  MCHK(m, hipMalloc(&ty, (size_t)chunk * m->n_out * 4));
  for (long long c0 = 0; c0 < n; c0 += chunk) {
    const long long nc = n - c0 < chunk ? n - c0 : chunk;
    MCHK(m, hipMemcpyAsync(tx, x + c0 * m->n_in, (size_t)nc * m->n_in * 4, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipMemcpyAsync(ty, y + c0 * m->n_out, (size_t)nc * m->n_out * 4, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_norm_in, dim3(nblk(nc * ip)), dim3(256), 0, m->stream, tx, nc, m->n_in, ip, m->x_mean, m->x_scale,
                       m->DX + c0 * ip);
    hipLaunchKernelGGL(k_norm_out, dim3(nblk(nc * op)), dim3(256), 0, m->stream, ty, nc, m->n_out, op, m->y_mean, m->y_scale,
                       m->DY + c0 * op);
    MCHK(m, hipStreamSynchronize(m->stream));
  }
  (void)hipFree(tx);
  (void)hipFree(ty);
  m->ndata = n;
  return LGAME_OK;
}

// Adam training on the loaded data set: `epochs` passes in shuffled batches of `batch` rows (a
// multiple of 64); loss[e] receives the mean squared error of epoch e over all its batches (each
// taken before its update).  *seconds (if non-null) the device time of the run.
int lgame_mlp_train(lgame_mlp *m, int epochs, int batch, float lr, unsigned long long seed, float *loss, double *seconds) {
  if (!m || epochs < 1 || batch < 64 || batch % 64 || !loss) return LGAME_EINVAL;
  if (m->ndata < batch) return mfail(m, LGAME_ESTATE, "load at least one batch of data first");
  if (ensure_cap(m, batch)) return mfail(m, LGAME_ENOMEM, "activation buffers");
  const int ip = pad64(m->n_in), op = pad64(m->n_out);
  const long long nb = m->ndata / batch;
  const int nparts = (int)nblk((long long)batch * op);
  unsigned long long st = seed * 0x9E3779B97F4A7C15ull + 1;
  // The permutation of the next epoch is drawn on the host while the device trains the current
  // one: two pinned host buffers and two device copies.  The squared error is summed on the
  // device, batch by batch in order, so nothing waits for the device within the run.
  int *hperm[2] = {nullptr, nullptr}, *drows[2] = {nullptr, nullptr};
  float *dloss = nullptr;
  hipEvent_t e0, e1, copied[2];
  MCHK(m, hipEventCreate(&e0));
  MCHK(m, hipEventCreate(&e1));
  for (int q = 0; q < 2; ++q) {
    MCHK(m, hipHostMalloc(&hperm[q], (size_t)m->ndata * 4, hipHostMallocDefault));
    MCHK(m, hipMalloc(&drows[q], (size_t)m->ndata * 4));
    MCHK(m, hipEventCreate(&copied[q]));
  }
  MCHK(m, hipMalloc(&dloss, (size_t)epochs * 4));
  MCHK(m, hipMemsetAsync(dloss, 0, (size_t)epochs * 4, m->stream));
  auto shuffle = [&](int *perm) {
    for (long long i = 0; i < m->ndata; ++i) perm[i] = (int)i;
    for (long long i = m->ndata - 1; i > 0; --i) {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      std::swap(perm[i], perm[st % (unsigned long long)(i + 1)]);
    }
  };
  shuffle(hperm[0]);
  MCHK(m, hipMemcpyAsync(drows[0], hperm[0], (size_t)m->ndata * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipEventRecord(copied[0], m->stream));
  MCHK(m, hipEventRecord(e0, m->stream));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    const int cur = ep % 2;
    for (long long b = 0; b < nb; ++b) {
      const int *rows = drows[cur] + b * batch;
      hipLaunchKernelGGL(k_gather_t, dim3(ip / 32, batch / 32), dim3(256), 0, m->stream, m->DX, rows, ip, m->X, m->XT,
                         (long long)batch);
      forward(m, m->X, batch, true);
      hipLaunchKernelGGL(k_mse_grad, dim3(nparts), dim3(256), 0, m->stream, m->Yf, m->DY, rows, batch, m->n_out, op, m->dYb, m->dTa,
                         m->partial);
      hipLaunchKernelGGL(k_loss_add, dim3(1), dim3(256), 0, m->stream, (const float *)m->partial, nparts, dloss + ep);
      ++m->adam_t;
      const AdamArgs adam{lr, b1, b2, eps, 1.0f - std::pow(b1, (float)m->adam_t), 1.0f - std::pow(b2, (float)m->adam_t)};
      backward(m, m->XT, batch, &adam);  // with each layer's Adam update
    }
    if (ep + 1 < epochs) {  // the next epoch's rows, drawn while this epoch runs
      const int nxt = (ep + 1) % 2;
      MCHK(m, hipEventSynchronize(copied[nxt]));  // its buffer's previous copy has completed
      shuffle(hperm[nxt]);
      MCHK(m, hipMemcpyAsync(drows[nxt], hperm[nxt], (size_t)m->ndata * 4, hipMemcpyHostToDevice, m->stream));
      MCHK(m, hipEventRecord(copied[nxt], m->stream));
    }
  }
  MCHK(m, hipEventRecord(e1, m->stream));
  MCHK(m, hipEventSynchronize(e1));
  float ms = 0.0f;
  MCHK(m, hipEventElapsedTime(&ms, e0, e1));
  if (seconds) *seconds = ms * 1e-3;
  std::vector<float> se(epochs);
  MCHK(m, hipMemcpy(se.data(), dloss, (size_t)epochs * 4, hipMemcpyDeviceToHost));
  for (int ep = 0; ep < epochs; ++ep) loss[ep] = (float)((double)se[ep] / ((double)nb * batch * m->n_out));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int q = 0; q < 2; ++q) {
    (void)hipEventDestroy(copied[q]);
    (void)hipHostFree(hperm[q]);
    (void)hipFree(drows[q]);
  }
  (void)hipFree(dloss);
  MCHK(m, hipGetLastError());
  return LGAME_OK;
}

const char *lgame_mlp_last_error(const lgame_mlp *m) { return m ? m->err.c_str() : ""; }

void lgame_mlp_destroy(lgame_mlp *m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  mlp_work_destroy(m->host_work);
  for (auto &Ly : m->L)
    for (void *p : {(void *)Ly.w, (void *)Ly.b, (void *)Ly.mw, (void *)Ly.vw, (void *)Ly.mb, (void *)Ly.vb, (void *)Ly.gw,
                    (void *)Ly.gb, (void *)Ly.wbf, (void *)Ly.wbfT})
      if (p) (void)hipFree(p);
  for (auto p : m->Z) (void)hipFree(p);
  for (auto p : m->A) (void)hipFree(p);
  for (auto p : m->AT) (void)hipFree(p);
  for (void *p : {(void *)m->X, (void *)m->XT, (void *)m->dA, (void *)m->dB, (void *)m->dYb, (void *)m->dTa, (void *)m->dTb,
                  (void *)m->gpart, (void *)m->Yf, (void *)m->Yt, (void *)m->partial, (void *)m->xin, (void *)m->yout,
                  (void *)m->rows, (void *)m->DX, (void *)m->DY, (void *)m->x_mean, (void *)m->x_scale, (void *)m->y_mean,
                  (void *)m->y_scale})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

}  // extern "C"
