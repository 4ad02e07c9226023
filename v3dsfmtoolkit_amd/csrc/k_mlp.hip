// k_mlp.hip -- the learned proposal network: embedding MLP 128 -> 256 (GELU)
// -> 64, L2-normalised, with BF16 weights; forward, backward, symmetric
// InfoNCE loss and Adam, all on the GPU.
//
// Input x = u / 64 (u the uint8 descriptor).  Forward, per row:
//   a1 = W1r x + b1, h = gelu(a1) = a1 (1 + erf(a1 / sqrt 2)) / 2,
//   z = W2r h + b2, e = z / sqrt(|z|^2 + 1e-12),
// with W1r, W2r the BF16 roundings (nearest even) of the FP32 master weights
// (training updates the masters, the forward pass always uses the rounded
// weights, so inference runs exactly the trained network).  Loss over a
// batch of B matched pairs (l_i, r_i), s_ij = e(l_i) . e(r_j) / tau:
//   L = 1/(2B) sum_i (lse_j s_ij - s_ii) + 1/(2B) sum_j (lse_i s_ij - s_jj).
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace v3dm {

constexpr int kGT = 64;  // GEMM tile
constexpr int kGK = 16;  // k per LDS stage

// C = alpha op(A) op(B) (+ C when accumulate) (+ bias[n]); row-major.
// op(A) is M x K: A[m*lda + k], or A[k*lda + m] when TA.  op(B) is K x N:
// B[k*ldb + n], or B[n*ldb + k] when TB.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_f32(int M, int N, int K, const float *__restrict__ A, int lda,
                                                  const float *__restrict__ B, int ldb, float *__restrict__ C,
                                                  int ldc, const float *__restrict__ bias, float alpha,
                                                  int accumulate)
{
   __shared__ float As[kGK][kGT + 4];
   __shared__ float Bs[kGK][kGT + 4];
   const int tid = threadIdx.x;
   const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
   const int tx = tid & 15, ty = tid >> 4;
   float acc[4][4];
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
   for (int k0 = 0; k0 < K; k0 += kGK)
   {
#pragma unroll
      for (int e = tid; e < kGT * kGK; e += 256)
      {
         int mm, kk;
         if (TA) { kk = e / kGT; mm = e % kGT; }
         else { mm = e / kGK; kk = e % kGK; }
         const int m = m0 + mm, k = k0 + kk;
         float va = 0.0f;
         if (m < M && k < K) va = TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
         As[kk][mm] = va;
         int nn, kb;
         if (TB) { nn = e / kGK; kb = e % kGK; }
         else { kb = e / kGT; nn = e % kGT; }
         const int n = n0 + nn, kq = k0 + kb;
         float vb = 0.0f;
         if (n < N && kq < K) vb = TB ? B[(size_t)n * ldb + kq] : B[(size_t)kq * ldb + n];
         Bs[kb][nn] = vb;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < kGK; ++kk)
      {
         float a[4], b[4];
#pragma unroll
         for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
         for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
         for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
      }
      __syncthreads();
   }
#pragma unroll
   for (int i = 0; i < 4; ++i)
   {
      const int m = m0 + ty * 4 + i;
      if (m >= M) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
         const int n = n0 + tx * 4 + j;
         if (n >= N) continue;
         float v = alpha * acc[i][j];
         if (bias) v += bias[n];
         float *c = C + (size_t)m * ldc + n;
         *c = accumulate ? *c + v : v;
      }
   }
}

__device__ inline float gelu(float a) { return 0.5f * a * (1.0f + erff(a * 0.70710678118654752f)); }

__device__ inline float gelu_grad(float a)
{
   return 0.5f * (1.0f + erff(a * 0.70710678118654752f)) + a * 0.39894228040143268f * expf(-0.5f * a * a);
}

__global__ void k_gelu_fwd(int64_t n, const float *__restrict__ a, float *__restrict__ h)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) h[i] = gelu(a[i]);
}

__global__ void k_gelu_bwd(int64_t n, const float *__restrict__ a, const float *__restrict__ dh,
                           float *__restrict__ da)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) da[i] = dh[i] * gelu_grad(a[i]);
}

__device__ inline float wave_sum(float v)
{
#pragma unroll
   for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
   return v;
}

__device__ inline float wave_max(float v)
{
#pragma unroll
   for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
   return v;
}

// One wave per 64-wide row: e = z / sqrt(|z|^2 + 1e-12); nrm = that root.
__global__ __launch_bounds__(256) void k_l2n_fwd(int rows, const float *__restrict__ z, float *__restrict__ e,
                                                 float *__restrict__ nrm, uint16_t *__restrict__ ebf)
{
   const int row = blockIdx.x * 4 + (threadIdx.x >> 6), d = threadIdx.x & 63;
   if (row >= rows) return;
   const float v = z[(size_t)row * 64 + d];
   const float n = sqrtf(wave_sum(v * v) + 1e-12f);
   const float o = v / n;
   if (e) e[(size_t)row * 64 + d] = o;
   if (nrm && d == 0) nrm[row] = n;
   if (ebf)
   {
      const uint32_t u = __float_as_uint(o);
      ebf[(size_t)row * 64 + d] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
   }
}

// dz = (de - e (e . de)) / nrm
__global__ __launch_bounds__(256) void k_l2n_bwd(int rows, const float *__restrict__ e, const float *__restrict__ nrm,
                                                 const float *__restrict__ de, float *__restrict__ dz)
{
   const int row = blockIdx.x * 4 + (threadIdx.x >> 6), d = threadIdx.x & 63;
   if (row >= rows) return;
   const size_t i = (size_t)row * 64 + d;
   const float dot = wave_sum(e[i] * de[i]);
   dz[i] = (de[i] - e[i] * dot) / nrm[row];
}

// Log-sum-exp of every row (byCol = 0) or column (byCol = 1) of the B x B
// logits, one wave each.
__global__ __launch_bounds__(256) void k_lse(int B, const float *__restrict__ S, int byCol, float *__restrict__ out)
{
   const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
   if (r >= B) return;
   float m = -INFINITY;
   for (int j = lane; j < B; j += 64) m = fmaxf(m, byCol ? S[(size_t)j * B + r] : S[(size_t)r * B + j]);
   m = wave_max(m);
   float s = 0.0f;
   for (int j = lane; j < B; j += 64) s += expf((byCol ? S[(size_t)j * B + r] : S[(size_t)r * B + j]) - m);
   s = wave_sum(s);
   if (lane == 0) out[r] = m + logf(s);
}

// G_ij = dL/ds_ij = (exp(s_ij - lseR_i) + exp(s_ij - lseC_j) - 2 [i == j]) / (2B)
__global__ void k_infonce_grad(int B, const float *__restrict__ S, const float *__restrict__ lseR,
                               const float *__restrict__ lseC, float *__restrict__ G)
{
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= (int64_t)B * B) return;
   const int i = (int)(t / B), j = (int)(t % B);
   const float s = S[t];
   float g = expf(s - lseR[i]) + expf(s - lseC[j]);
   if (i == j) g -= 2.0f;
   G[t] = g / (2.0f * (float)B);
}

// loss = sum_i (lseR_i + lseC_i - 2 s_ii) / (2B); one workgroup.
__global__ __launch_bounds__(256) void k_infonce_loss(int B, const float *__restrict__ S, const float *__restrict__ lseR,
                                                      const float *__restrict__ lseC, float *__restrict__ loss)
{
   __shared__ float part[4];
   float acc = 0.0f;
   for (int i = threadIdx.x; i < B; i += 256) acc += lseR[i] + lseC[i] - 2.0f * S[(size_t)i * B + i];
   acc = wave_sum(acc);
   if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
   __syncthreads();
   if (threadIdx.x == 0) *loss = (part[0] + part[1] + part[2] + part[3]) / (2.0f * (float)B);
}

// out[c] = sum_r X[r * cols + c]
__global__ void k_colsum(int rows, int cols, const float *__restrict__ X, float *__restrict__ out)
{
   const int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= cols) return;
   float s = 0.0f;
   for (int r = 0; r < rows; ++r) s += X[(size_t)r * cols + c];
   out[c] = s;
}

__device__ inline float bf16_round(float f)
{
   const uint32_t u = __float_as_uint(f);
   return __uint_as_float(((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16) << 16);
}

// Adam on FP32 masters; Wr receives the forward copy (BF16-rounded weights
// when round != 0, the masters themselves for biases).
__global__ void k_adam(int n, float *__restrict__ W, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float *__restrict__ Wr, float lr, float b1, float b2, float eps,
                       float c1, float c2, int round)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const float gi = g[i];
   const float mi = b1 * m[i] + (1.0f - b1) * gi;
   const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
   m[i] = mi;
   v[i] = vi;
   const float w = W[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
   W[i] = w;
   Wr[i] = round ? bf16_round(w) : w;
}

__global__ void k_round_copy(int n, const float *__restrict__ W, float *__restrict__ Wr, int round)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) Wr[i] = round ? bf16_round(W[i]) : W[i];
}

// x rows from uint8 descriptors: x[i] = u[idx ? idx[i] : i] / 64.
__global__ void k_u8_rows_to_x(int n, const uint8_t *__restrict__ u, const int64_t *__restrict__ idx,
                               float *__restrict__ x)
{
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= (int64_t)n * 128) return;
   const int64_t row = t >> 7, d = t & 127;
   const int64_t src = idx ? idx[row] : row;
   x[t] = (float)u[src * 128 + d] * (1.0f / 64.0f);
}

// Batch sampling with replacement: idx[i] = hash(seed, step, i) mod n.
__global__ void k_sample(int B, uint64_t seed, uint64_t step, int64_t n, int64_t *__restrict__ idx)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= B) return;
   uint64_t z = seed + 0x9E3779B97F4A7C15ull * (step * 0x100000001B3ull + (uint64_t)i + 1);
   z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
   z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
   z ^= z >> 31;
   idx[i] = (int64_t)(z % (uint64_t)n);
}

// ---- fused inference (MFMA) ---------------------------------------------------
//
// One wave embeds 32 rows, computing the transposed layers so that every lane
// owns one data row as an MFMA column:
//   layer 1, a1^T = W1r x^T on v_mfma_f32_32x32x16_bf16: A = W1r (hidden
//   units as rows, BF16), B = x^T (x = u / 64 is exact in BF16, so every
//   product is exact; FP32 accumulation).  Accumulator register v of output
//   block n in lane l holds unit 32n + 8(v/4) + 4(l/32) + v%4 of row l%32.
//   layer 2, z^T = W2r h^T on v_mfma_f32_32x32x2_f32 (FP32 throughout): in
//   step t = 16n + v the lane half h supplies k = h, paired with unit
//   U = 32n + 8(v/4) + 4h + v%4 -- exactly the h value that lane already
//   holds in register (n, v) -- and A supplies W2r[o][U].  The sum over k
//   does not depend on which unit a step assigns to which k slot.
// The L2 norm of a row combines the two lane halves (outputs 8a + b and
// 8a + 4 + b).  Same arithmetic as the training forward pass up to the order
// of the FP32 sums.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v4i_ __attribute__((ext_vector_type(4)));

// W1bf[u][k] = W1r[u][k] as BF16 bits (exact: already rounded);
// W2T[u][o] = W2r[o][u].
__global__ void k_embed_weights(const float *__restrict__ Pr, uint16_t *__restrict__ W1bf, float *__restrict__ W2T)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < 256 * 128) W1bf[i] = (uint16_t)(__float_as_uint(Pr[i]) >> 16);
   const int w2 = 256 * 128 + 256;
   if (i < 64 * 256)
   {
      const int o = i / 256, u = i % 256;
      W2T[u * 64 + o] = Pr[w2 + i];
   }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_embed_mfma(int n, const uint8_t *__restrict__ U8,
                                                    const uint16_t *__restrict__ W1bf, const float *__restrict__ b1,
                                                    const float *__restrict__ W2T, const float *__restrict__ b2,
                                                    float *__restrict__ outF, uint16_t *__restrict__ outBF)
{
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int half = lane >> 5, lj = lane & 31;
   const int row0 = (blockIdx.x * 4 + wave) * 32;
   if (row0 >= n) return;
   const int row = row0 + lj;
   const bool live = row < n;
   // B operand of layer 1: x^T fragments, k in [16s + 8h, 16s + 8h + 8)
   v4i_ X[8];
#pragma unroll
   for (int s = 0; s < 8; ++s)
   {
      uint2 u = make_uint2(0, 0);
      if (live) u = *reinterpret_cast<const uint2 *>(U8 + (size_t)row * 128 + 16 * s + 8 * half);
      uint32_t w[4];
#pragma unroll
      for (int e = 0; e < 4; ++e)
      {
         const uint32_t lo = ((e < 2 ? u.x : u.y) >> (16 * (e & 1))) & 0xFFu;
         const uint32_t hi = ((e < 2 ? u.x : u.y) >> (16 * (e & 1) + 8)) & 0xFFu;
         w[e] = (__float_as_uint((float)lo * 0.015625f) >> 16) | (__float_as_uint((float)hi * 0.015625f) & 0xFFFF0000u);
      }
      X[s] = v4i_{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
   }
   // per hidden block nb: layer-1 MFMAs, GELU, then that block's layer-2
   // steps, so only one block of hidden values is live at a time
   const v16f zero = {};
   v16f z[2] = {zero, zero};
#pragma unroll 1
   for (int nb = 0; nb < 8; ++nb)
   {
      float w2[2][16];
#pragma unroll
      for (int v = 0; v < 16; ++v)
      {
         const int u = 32 * nb + 8 * (v >> 2) + 4 * half + (v & 3);
#pragma unroll
         for (int ob = 0; ob < 2; ++ob) w2[ob][v] = W2T[u * 64 + 32 * ob + lj];
      }
      const uint16_t *wrow = W1bf + (size_t)(32 * nb + lj) * 128 + 8 * half;
      v4i_ W[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) W[s] = *reinterpret_cast<const v4i_ *>(wrow + 16 * s);
      v16f a = zero;
#pragma unroll
      for (int s = 0; s < 8; ++s)
         a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, W[s]), __builtin_bit_cast(v8bf, X[s]), a,
                                                    0, 0, 0);
      // register v is unit 32 nb + 8(v/4) + 4 half + v%4
#pragma unroll
      for (int v = 0; v < 16; ++v)
      {
         const int u = 32 * nb + 8 * (v >> 2) + 4 * half + (v & 3);
         const float hv = gelu(a[v] + b1[u]);
#pragma unroll
         for (int ob = 0; ob < 2; ++ob) z[ob] = __builtin_amdgcn_mfma_f32_32x32x2f32(w2[ob][v], hv, z[ob], 0, 0, 0);
      }
   }
   // bias, norm over the 64 outputs of the row (this half holds 32 of them)
   float ss = 0.0f;
#pragma unroll
   for (int ob = 0; ob < 2; ++ob)
#pragma unroll
      for (int v = 0; v < 16; ++v)
      {
         const int o = 32 * ob + 8 * (v >> 2) + 4 * half + (v & 3);
         z[ob][v] += b2[o];
         ss = fmaf(z[ob][v], z[ob][v], ss);
      }
   ss += __shfl_xor(ss, 32, 64);
   const float nrm = sqrtf(ss + 1e-12f);
   if (!live) return;
#pragma unroll
   for (int ob = 0; ob < 2; ++ob)
#pragma unroll
      for (int a = 0; a < 4; ++a)
      {
         const int o = 32 * ob + 8 * a + 4 * half;
         float e[4];
#pragma unroll
         for (int b = 0; b < 4; ++b) e[b] = z[ob][4 * a + b] / nrm;
         if (outF) *reinterpret_cast<float4 *>(outF + (size_t)row * 64 + o) = make_float4(e[0], e[1], e[2], e[3]);
         if (outBF)
         {
            uint32_t h2[2];
#pragma unroll
            for (int b = 0; b < 2; ++b)
            {
               const uint32_t x0 = __float_as_uint(e[2 * b]), x1 = __float_as_uint(e[2 * b + 1]);
               h2[b] = ((x0 + 0x7FFFu + ((x0 >> 16) & 1u)) >> 16) | (((x1 + 0x7FFFu + ((x1 >> 16) & 1u)) >> 16) << 16);
            }
            *reinterpret_cast<uint2 *>(outBF + (size_t)row * 64 + o) = make_uint2(h2[0], h2[1]);
         }
      }
}

// ---- launchers --------------------------------------------------------------

hipError_t launch_embed_weights(const float *Pr, uint16_t *W1bf, float *W2T, hipStream_t st)
{
   hipLaunchKernelGGL(k_embed_weights, dim3(128), dim3(256), 0, st, Pr, W1bf, W2T);
   return hipGetLastError();
}

hipError_t launch_embed_mfma(int n, const uint8_t *U8, const uint16_t *W1bf, const float *b1, const float *W2T,
                             const float *b2, float *outF, uint16_t *outBF, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_embed_mfma, dim3((n + 127) / 128), dim3(256), 0, st, n, U8, W1bf, b1, W2T, b2, outF, outBF);
   return hipGetLastError();
}

hipError_t launch_gemm_f32(bool ta, bool tb, int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                           float *C, int ldc, const float *bias, float alpha, bool accumulate, hipStream_t st)
{
   if (M <= 0 || N <= 0) return hipSuccess;
   dim3 grid((N + kGT - 1) / kGT, (M + kGT - 1) / kGT);
   const int acc = accumulate ? 1 : 0;
   if (!ta && !tb)
      hipLaunchKernelGGL((k_gemm_f32<false, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias,
                         alpha, acc);
   else if (!ta && tb)
      hipLaunchKernelGGL((k_gemm_f32<false, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias,
                         alpha, acc);
   else if (ta && !tb)
      hipLaunchKernelGGL((k_gemm_f32<true, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias,
                         alpha, acc);
   else
      hipLaunchKernelGGL((k_gemm_f32<true, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias,
                         alpha, acc);
   return hipGetLastError();
}

static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_gelu_fwd(int64_t n, const float *a, float *h, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_gelu_fwd, g1(n), dim3(256), 0, st, n, a, h);
   return hipGetLastError();
}

hipError_t launch_gelu_bwd(int64_t n, const float *a, const float *dh, float *da, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_gelu_bwd, g1(n), dim3(256), 0, st, n, a, dh, da);
   return hipGetLastError();
}

hipError_t launch_l2n_fwd(int rows, const float *z, float *e, float *nrm, uint16_t *ebf, hipStream_t st)
{
   if (rows <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_l2n_fwd, dim3((rows + 3) / 4), dim3(256), 0, st, rows, z, e, nrm, ebf);
   return hipGetLastError();
}

hipError_t launch_l2n_bwd(int rows, const float *e, const float *nrm, const float *de, float *dz, hipStream_t st)
{
   if (rows <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_l2n_bwd, dim3((rows + 3) / 4), dim3(256), 0, st, rows, e, nrm, de, dz);
   return hipGetLastError();
}

hipError_t launch_infonce(int B, const float *S, float *lseR, float *lseC, float *G, float *loss, hipStream_t st)
{
   if (B <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_lse, dim3((B + 3) / 4), dim3(256), 0, st, B, S, 0, lseR);
   hipLaunchKernelGGL(k_lse, dim3((B + 3) / 4), dim3(256), 0, st, B, S, 1, lseC);
   hipLaunchKernelGGL(k_infonce_grad, g1((int64_t)B * B), dim3(256), 0, st, B, S, lseR, lseC, G);
   hipLaunchKernelGGL(k_infonce_loss, dim3(1), dim3(256), 0, st, B, S, lseR, lseC, loss);
   return hipGetLastError();
}

hipError_t launch_colsum(int rows, int cols, const float *X, float *out, hipStream_t st)
{
   if (cols <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_colsum, g1(cols), dim3(256), 0, st, rows, cols, X, out);
   return hipGetLastError();
}

hipError_t launch_adam(int n, float *W, const float *g, float *m, float *v, float *Wr, float lr, float b1, float b2,
                       float eps, float c1, float c2, bool round, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_adam, g1(n), dim3(256), 0, st, n, W, g, m, v, Wr, lr, b1, b2, eps, c1, c2, round ? 1 : 0);
   return hipGetLastError();
}

hipError_t launch_round_copy(int n, const float *W, float *Wr, bool round, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_round_copy, g1(n), dim3(256), 0, st, n, W, Wr, round ? 1 : 0);
   return hipGetLastError();
}

hipError_t launch_u8_rows_to_x(int n, const uint8_t *u, const int64_t *idx, float *x, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_u8_rows_to_x, g1((int64_t)n * 128), dim3(256), 0, st, n, u, idx, x);
   return hipGetLastError();
}

hipError_t launch_sample(int B, uint64_t seed, uint64_t step, int64_t n, int64_t *idx, hipStream_t st)
{
   if (B <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_sample, g1(B), dim3(256), 0, st, B, seed, step, n, idx);
   return hipGetLastError();
}

} // namespace v3dm
