// v3dm_mlp.cpp -- the learned proposal path behind the C ABI: the embedding
// network (parameters, GPU training with Adam, batch loss and gradients,
// embedding inference) and v3dm_learned_find, which matches the current pair
// of a matcher context by retrieval + exact rescoring + exhaustive fallback.
//
// No counterpart in the reference (V3D/src/CUDA/v3d_cudamatching.cpp matches
// every pair exhaustively); the decisions it feeds are the reference rule
// (v3d_cudamatching.cu:117-194, v3d_cudamatching.cpp:121-126) evaluated on
// exact scores.
#include <chrono>
#include <cmath>
#include <random>

#include <vector>

#include "v3dm_ctx.h"

namespace v3dm {
hipError_t launch_gemm_f32(bool ta, bool tb, int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                           float *C, int ldc, const float *bias, float alpha, bool accumulate, hipStream_t st);
hipError_t launch_gelu_fwd(int64_t n, const float *a, float *h, hipStream_t st);
hipError_t launch_gelu_bwd(int64_t n, const float *a, const float *dh, float *da, hipStream_t st);
hipError_t launch_l2n_fwd(int rows, const float *z, float *e, float *nrm, uint16_t *ebf, hipStream_t st);
hipError_t launch_l2n_bwd(int rows, const float *e, const float *nrm, const float *de, float *dz, hipStream_t st);
hipError_t launch_infonce(int B, const float *S, float *lseR, float *lseC, float *G, float *loss, hipStream_t st);
hipError_t launch_colsum(int rows, int cols, const float *X, float *out, hipStream_t st);
hipError_t launch_adam(int n, float *W, const float *g, float *m, float *v, float *Wr, float lr, float b1, float b2,
                       float eps, float c1, float c2, bool round, hipStream_t st);
hipError_t launch_round_copy(int n, const float *W, float *Wr, bool round, hipStream_t st);
hipError_t launch_u8_rows_to_x(int n, const uint8_t *u, const int64_t *idx, float *x, hipStream_t st);
hipError_t launch_sample(int B, uint64_t seed, uint64_t step, int64_t n, int64_t *idx, hipStream_t st);
hipError_t launch_embed_weights(const float *Pr, uint16_t *W1bf, float *W2T, hipStream_t st);
hipError_t launch_embed_mfma(int n, const uint8_t *U8, const uint16_t *W1bf, const float *b1, const float *W2T,
                             const float *b2, float *outF, uint16_t *outBF, hipStream_t st);
int emb_topk_chunks(int nR, int K);
size_t emb_topk_scratch_bytes(int nQ, int nChunks);
hipError_t launch_emb_topk(int nQ, int nR, int K, const uint16_t *EQ, const uint16_t *ER, int nChunks, void *scratch,
                           int *cand, float *candS, int *ovf, hipStream_t st);
hipError_t launch_cand_insert(int nL, int nR, int K, const int *cand, const int8_t *L, const int8_t *R,
                              uint64_t *top1L, uint64_t *top2L, uint64_t *top1R, uint64_t *top2R,
                              const uint32_t *epoch, hipStream_t st);
hipError_t launch_flag(int n, int isRight, const uint64_t *top1, const uint64_t *top2, const uint32_t *epoch,
                       float margin, const int *force, int *list, int *count, hipStream_t st);
hipError_t launch_gather_rows(int n, int nPad, const int *nDev, const int *idx, const int8_t *src,
                              const int32_t *srcSum, int8_t *dst, int32_t *dstSum, hipStream_t st);
hipError_t launch_fallback_pairs(const int *count, int nL, int nR, int tileM, int tileN, PairDesc *out,
                                 hipStream_t st);
} // namespace v3dm

namespace {

constexpr int kIn = 128, kHid = 256, kOut = 64;
constexpr int kOffW1 = 0, kOffB1 = kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kOut * kHid;
constexpr int kParams = kOffB2 + kOut; // 49472
constexpr int kMaxCandK = 32; // candidates per query: 16 or 32

} // namespace

struct v3dm_mlp
{
   int device = 0;
   hipStream_t stream = nullptr;
   std::string err;
   float *dP = nullptr;  // FP32 masters: W1 | b1 | W2 | b2
   float *dPr = nullptr; // forward copy: BF16-rounded W1, W2; biases as is
   float *dG = nullptr;  // gradients, same layout
   float *dM = nullptr, *dV = nullptr;
   int64_t step = 0;
   uint8_t *dTL = nullptr, *dTR = nullptr;
   int64_t nTrain = 0;
   int batchCap = 0;
   uint8_t *dBL = nullptr, *dBR = nullptr;
   int64_t *dIdx = nullptr;
   float *dX = nullptr, *dA1 = nullptr, *dH = nullptr, *dZ = nullptr, *dE = nullptr, *dN = nullptr;
   float *dS = nullptr, *dGs = nullptr, *dLseR = nullptr, *dLseC = nullptr, *dLoss = nullptr;
   float *dDE = nullptr, *dDZ = nullptr, *dDH = nullptr, *dDA1 = nullptr;
   float *dLossHist = nullptr;
   int lossHistCap = 0;
   float *iE = nullptr;
   uint16_t *dW1bf = nullptr; // inference copies of the forward weights
   float *dW2T = nullptr;
   uint8_t *iU8 = nullptr;
   int iU8Cap = 0;
};

namespace v3dm_host {

// Buffers of the learned path on a matcher context.
struct LearnedState
{
   int capL = 0, capR = 0;
   uint16_t *embL = nullptr, *embR = nullptr;
   void *scratch = nullptr; // retrieval sweeps: group maxima, thresholds, candidate lists
   size_t scratchBytes = 0;
   int *cand = nullptr;
   int *ovf = nullptr;      // per left: candidate list overflowed
   int *flagL = nullptr, *flagR = nullptr, *dCount = nullptr, *hCount = nullptr;
   int8_t *gI8 = nullptr;
   int32_t *gSum = nullptr;
   int gCap = 0;
   PairDesc *dPairFb = nullptr; // the two fallback scans' pair descriptors
   int lastNL = 0;
   int lastK = 0;
};

void freeLearned(v3dm_ctx *c)
{
   LearnedState *s = c->learned;
   if (!s) return;
   dfree(s->embL);
   dfree(s->embR);
   dfree(s->scratch);
   dfree(s->cand);
   dfree(s->ovf);
   dfree(s->flagL);
   dfree(s->flagR);
   dfree(s->dCount);
   if (s->hCount) (void)hipHostFree(s->hCount);
   dfree(s->gI8);
   dfree(s->gSum);
   dfree(s->dPairFb);
   delete s;
   c->learned = nullptr;
}

} // namespace v3dm_host

namespace {

int mfail(v3dm_mlp *m, int code, const char *fmt, ...)
{
   char buf[512];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof(buf), fmt, ap);
   va_end(ap);
   if (m) m->err = buf;
   else g_createError = buf;
   return code;
}

#define MLPCHK(m, call)                                                                                  \
   do                                                                                                    \
   {                                                                                                     \
      hipError_t e_ = (call);                                                                            \
      if (e_ != hipSuccess)                                                                              \
         return mfail((m), e_ == hipErrorOutOfMemory ? V3DM_ERR_NOMEM : V3DM_ERR_DEVICE, "%s: %s (%s:%d)", \
                      #call, hipGetErrorString(e_), __FILE__, __LINE__);                                 \
   } while (0)

void freeMlp(v3dm_mlp *m)
{
   for (float **p : {&m->dP, &m->dPr, &m->dG, &m->dM, &m->dV, &m->dX, &m->dA1, &m->dH, &m->dZ, &m->dE, &m->dN,
                     &m->dS, &m->dGs, &m->dLseR, &m->dLseC, &m->dLoss, &m->dDE, &m->dDZ, &m->dDH, &m->dDA1,
                     &m->dLossHist, &m->iE, &m->dW2T})
      dfree(*p);
   dfree(m->dW1bf);
   dfree(m->dTL);
   dfree(m->dTR);
   dfree(m->dBL);
   dfree(m->dBR);
   dfree(m->dIdx);
   dfree(m->iU8);
}

int ensureBatch(v3dm_mlp *m, int B)
{
   if (B <= m->batchCap) return V3DM_OK;
   for (float **p : {&m->dX, &m->dA1, &m->dH, &m->dZ, &m->dE, &m->dN, &m->dS, &m->dGs, &m->dLseR, &m->dLseC,
                     &m->dLoss, &m->dDE, &m->dDZ, &m->dDH, &m->dDA1})
      dfree(*p);
   dfree(m->dBL);
   dfree(m->dBR);
   dfree(m->dIdx);
   m->batchCap = 0;
   const size_t R = 2 * (size_t)B;
   MLPCHK(m, hipMalloc((void **)&m->dBL, (size_t)B * kIn));
   MLPCHK(m, hipMalloc((void **)&m->dBR, (size_t)B * kIn));
   MLPCHK(m, hipMalloc((void **)&m->dIdx, sizeof(int64_t) * B));
   MLPCHK(m, hipMalloc((void **)&m->dX, sizeof(float) * R * kIn));
   MLPCHK(m, hipMalloc((void **)&m->dA1, sizeof(float) * R * kHid));
   MLPCHK(m, hipMalloc((void **)&m->dH, sizeof(float) * R * kHid));
   MLPCHK(m, hipMalloc((void **)&m->dZ, sizeof(float) * R * kOut));
   MLPCHK(m, hipMalloc((void **)&m->dE, sizeof(float) * R * kOut));
   MLPCHK(m, hipMalloc((void **)&m->dN, sizeof(float) * R));
   MLPCHK(m, hipMalloc((void **)&m->dS, sizeof(float) * (size_t)B * B));
   MLPCHK(m, hipMalloc((void **)&m->dGs, sizeof(float) * (size_t)B * B));
   MLPCHK(m, hipMalloc((void **)&m->dLseR, sizeof(float) * B));
   MLPCHK(m, hipMalloc((void **)&m->dLseC, sizeof(float) * B));
   MLPCHK(m, hipMalloc((void **)&m->dLoss, sizeof(float)));
   MLPCHK(m, hipMalloc((void **)&m->dDE, sizeof(float) * R * kOut));
   MLPCHK(m, hipMalloc((void **)&m->dDZ, sizeof(float) * R * kOut));
   MLPCHK(m, hipMalloc((void **)&m->dDH, sizeof(float) * R * kHid));
   MLPCHK(m, hipMalloc((void **)&m->dDA1, sizeof(float) * R * kHid));
   m->batchCap = B;
   return V3DM_OK;
}

// Forward copy of the parameters (rounded weights) from the masters.
int refreshForward(v3dm_mlp *m)
{
   MLPCHK(m, launch_round_copy(kOffB1, m->dP, m->dPr, true, m->stream));
   MLPCHK(m, launch_round_copy(kHid, m->dP + kOffB1, m->dPr + kOffB1, false, m->stream));
   MLPCHK(m, launch_round_copy(kOut * kHid, m->dP + kOffW2, m->dPr + kOffW2, true, m->stream));
   MLPCHK(m, launch_round_copy(kOut, m->dP + kOffB2, m->dPr + kOffB2, false, m->stream));
   return V3DM_OK;
}

// One batch through forward, loss and backward (gradients into dG); the
// batch's rows are already in dX (rows 0..B-1 left, B..2B-1 right).
int forwardBackward(v3dm_mlp *m, int B, float tau, float *lossDst)
{
   hipStream_t st = m->stream;
   const int R = 2 * B;
   const float *W1 = m->dPr + kOffW1, *b1 = m->dPr + kOffB1, *W2 = m->dPr + kOffW2, *b2 = m->dPr + kOffB2;
   MLPCHK(m, launch_gemm_f32(false, true, R, kHid, kIn, m->dX, kIn, W1, kIn, m->dA1, kHid, b1, 1.0f, false, st));
   MLPCHK(m, launch_gelu_fwd((int64_t)R * kHid, m->dA1, m->dH, st));
   MLPCHK(m, launch_gemm_f32(false, true, R, kOut, kHid, m->dH, kHid, W2, kHid, m->dZ, kOut, b2, 1.0f, false, st));
   MLPCHK(m, launch_l2n_fwd(R, m->dZ, m->dE, m->dN, nullptr, st));
   const float *EL = m->dE, *ER = m->dE + (size_t)B * kOut;
   MLPCHK(m, launch_gemm_f32(false, true, B, B, kOut, EL, kOut, ER, kOut, m->dS, B, nullptr, 1.0f / tau, false, st));
   MLPCHK(m, launch_infonce(B, m->dS, m->dLseR, m->dLseC, m->dGs, lossDst, st));
   MLPCHK(m, launch_gemm_f32(false, false, B, kOut, B, m->dGs, B, ER, kOut, m->dDE, kOut, nullptr, 1.0f / tau, false,
                             st));
   MLPCHK(m, launch_gemm_f32(true, false, B, kOut, B, m->dGs, B, EL, kOut, m->dDE + (size_t)B * kOut, kOut, nullptr,
                             1.0f / tau, false, st));
   MLPCHK(m, launch_l2n_bwd(R, m->dE, m->dN, m->dDE, m->dDZ, st));
   MLPCHK(m, launch_gemm_f32(true, false, kOut, kHid, R, m->dDZ, kOut, m->dH, kHid, m->dG + kOffW2, kHid, nullptr, 1.0f,
                             false, st));
   MLPCHK(m, launch_colsum(R, kOut, m->dDZ, m->dG + kOffB2, st));
   MLPCHK(m, launch_gemm_f32(false, false, R, kHid, kOut, m->dDZ, kOut, W2, kHid, m->dDH, kHid, nullptr, 1.0f, false,
                             st));
   MLPCHK(m, launch_gelu_bwd((int64_t)R * kHid, m->dA1, m->dDH, m->dDA1, st));
   MLPCHK(m, launch_gemm_f32(true, false, kHid, kIn, R, m->dDA1, kHid, m->dX, kIn, m->dG + kOffW1, kIn, nullptr, 1.0f,
                             false, st));
   MLPCHK(m, launch_colsum(R, kHid, m->dDA1, m->dG + kOffB1, st));
   return V3DM_OK;
}

int adamStep(v3dm_mlp *m, float lr)
{
   const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
   m->step += 1;
   const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
   const int seg[4][2] = {{kOffW1, kOffB1 - kOffW1}, {kOffB1, kHid}, {kOffW2, kOut * kHid}, {kOffB2, kOut}};
   for (int s = 0; s < 4; ++s)
   {
      const int o = seg[s][0], n = seg[s][1];
      MLPCHK(m, launch_adam(n, m->dP + o, m->dG + o, m->dM + o, m->dV + o, m->dPr + o, lr, b1, b2, eps, c1, c2,
                            s == 0 || s == 2, m->stream));
   }
   return V3DM_OK;
}

// Embeddings of n uint8 rows (device), on stream st, by the fused MFMA
// kernel: FP32 rows to outF and/or BF16 rows to outBF (rows n..padTo-1 of
// outBF are zeroed).
int embedRows(v3dm_mlp *m, hipStream_t st, int n, const uint8_t *dU8, float *outF, uint16_t *outBF, int padTo)
{
   if (!m->dW1bf)
   {
      MLPCHK(m, hipMalloc((void **)&m->dW1bf, sizeof(uint16_t) * kHid * kIn));
      MLPCHK(m, hipMalloc((void **)&m->dW2T, sizeof(float) * kHid * kOut));
   }
   MLPCHK(m, launch_embed_weights(m->dPr, m->dW1bf, m->dW2T, st));
   MLPCHK(m, launch_embed_mfma(n, dU8, m->dW1bf, m->dPr + kOffB1, m->dW2T, m->dPr + kOffB2, outF, outBF, st));
   if (outBF && padTo > n)
      MLPCHK(m, hipMemsetAsync(outBF + (size_t)n * kOut, 0, sizeof(uint16_t) * (size_t)(padTo - n) * kOut, st));
   return V3DM_OK;
}

uint64_t splitmix(uint64_t &s)
{
   uint64_t z = (s += 0x9E3779B97F4A7C15ull);
   z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
   z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
   return z ^ (z >> 31);
}

double normal(uint64_t &s)
{
   const double u1 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
   const double u2 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
   return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

} // namespace

extern "C" {

int v3dm_mlp_param_count(void) { return kParams; }

int v3dm_mlp_create(int device, uint64_t seed, v3dm_mlp **out)
{
   if (!out) return mfail(nullptr, V3DM_ERR_ARG, "mlp_create: out is NULL");
   *out = nullptr;
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
      return mfail(nullptr, V3DM_ERR_DEVICE, "mlp_create: no HIP device available");
   if (device < 0 || device >= n) return mfail(nullptr, V3DM_ERR_ARG, "mlp_create: device %d out of range", device);
   if (hipSetDevice(device) != hipSuccess) return mfail(nullptr, V3DM_ERR_DEVICE, "mlp_create: hipSetDevice failed");
   v3dm_mlp *m = new v3dm_mlp();
   m->device = device;
   int rc = V3DM_OK;
   auto chk = [&](hipError_t e) {
      if (e != hipSuccess && rc == V3DM_OK) rc = mfail(nullptr, V3DM_ERR_DEVICE, "mlp_create: %s", hipGetErrorString(e));
   };
   chk(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
   for (float **p : {&m->dP, &m->dPr, &m->dG, &m->dM, &m->dV})
      chk(hipMalloc((void **)p, sizeof(float) * kParams));
   if (rc == V3DM_OK)
   {
      // He initialisation of W1, 1/sqrt(fan-in) for W2, zero biases
      std::vector<float> p(kParams, 0.0f);
      uint64_t s = seed;
      for (int i = 0; i < kHid * kIn; ++i) p[kOffW1 + i] = (float)(normal(s) * std::sqrt(2.0 / kIn));
      for (int i = 0; i < kOut * kHid; ++i) p[kOffW2 + i] = (float)(normal(s) * std::sqrt(1.0 / kHid));
      chk(hipMemcpy(m->dP, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
      chk(hipMemset(m->dM, 0, sizeof(float) * kParams));
      chk(hipMemset(m->dV, 0, sizeof(float) * kParams));
   }
   if (rc == V3DM_OK)
   {
      rc = refreshForward(m);
      if (rc) g_createError = m->err;
   }
   if (rc == V3DM_OK && hipStreamSynchronize(m->stream) != hipSuccess)
      rc = mfail(nullptr, V3DM_ERR_DEVICE, "mlp_create: initialisation failed");
   if (rc != V3DM_OK)
   {
      freeMlp(m);
      if (m->stream) (void)hipStreamDestroy(m->stream);
      delete m;
      return rc;
   }
   *out = m;
   return V3DM_OK;
}

void v3dm_mlp_destroy(v3dm_mlp *m)
{
   if (!m) return;
   (void)hipSetDevice(m->device);
   (void)hipStreamSynchronize(m->stream);
   freeMlp(m);
   (void)hipStreamDestroy(m->stream);
   delete m;
}

const char *v3dm_mlp_last_error(const v3dm_mlp *m) { return m ? m->err.c_str() : g_createError.c_str(); }

int v3dm_mlp_get_params(v3dm_mlp *m, float *params)
{
   if (!m || !params) return V3DM_ERR_ARG;
   MLPCHK(m, hipSetDevice(m->device));
   MLPCHK(m, hipStreamSynchronize(m->stream));
   MLPCHK(m, hipMemcpy(params, m->dP, sizeof(float) * kParams, hipMemcpyDeviceToHost));
   return V3DM_OK;
}

// ---- persistence -------------------------------------------------------------
// File: "V3DMMLP" + NUL, then little-endian uint32 version (1), input, hidden
// and output widths, parameter count, FNV-1a checksum of the payload; the
// payload is what inference uses, in the parameter layout: W1 as BF16,
// b1 as FP32, W2 as BF16, b2 as FP32 (99584 bytes).  Loading sets the FP32
// masters to exactly those values, so a loaded network embeds bit for bit
// like the saved one (training resumes from BF16-valued weights).

namespace {
constexpr char kMlpMagic[8] = {'V', '3', 'D', 'M', 'M', 'L', 'P', 0};
constexpr uint32_t kMlpVersion = 1;

uint16_t bf16_bits(float f) // round to nearest even, as the forward copy is rounded on the device
{
   uint32_t u;
   memcpy(&u, &f, 4);
   return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
float bf16_value(uint16_t b)
{
   const uint32_t u = (uint32_t)b << 16;
   float f;
   memcpy(&f, &u, 4);
   return f;
}
uint32_t fnv1a(const std::vector<uint8_t> &v)
{
   uint32_t h = 2166136261u;
   for (uint8_t x : v) h = (h ^ x) * 16777619u;
   return h;
}
bool isWeight(int i) { return i < kOffB1 || (i >= kOffW2 && i < kOffB2); }
} // namespace

int v3dm_mlp_save(v3dm_mlp *m, const char *path)
{
   if (!m || !path) return V3DM_ERR_ARG;
   std::vector<float> p(kParams);
   int rc = v3dm_mlp_get_params(m, p.data());
   if (rc) return rc;
   std::vector<uint8_t> payload;
   payload.reserve(2 * (kOffB1 + kOut * kHid) + 4 * (kHid + kOut));
   for (int i = 0; i < kParams; ++i)
   {
      if (isWeight(i))
      {
         const uint16_t b = bf16_bits(p[i]);
         payload.push_back((uint8_t)(b & 0xFF));
         payload.push_back((uint8_t)(b >> 8));
      }
      else
      {
         uint32_t u;
         memcpy(&u, &p[i], 4);
         for (int k = 0; k < 4; ++k) payload.push_back((uint8_t)(u >> (8 * k)));
      }
   }
   const uint32_t head[6] = {kMlpVersion, (uint32_t)kIn, (uint32_t)kHid, (uint32_t)kOut, (uint32_t)kParams,
                             fnv1a(payload)};
   FILE *f = fopen(path, "wb");
   if (!f) return mfail(m, V3DM_ERR_ARG, "mlp_save: cannot open %s for writing", path);
   bool ok = fwrite(kMlpMagic, 1, 8, f) == 8;
   for (uint32_t h : head)
   {
      const uint8_t le[4] = {(uint8_t)h, (uint8_t)(h >> 8), (uint8_t)(h >> 16), (uint8_t)(h >> 24)};
      ok = ok && fwrite(le, 1, 4, f) == 4;
   }
   ok = ok && fwrite(payload.data(), 1, payload.size(), f) == payload.size();
   ok = (fclose(f) == 0) && ok;
   if (!ok) return mfail(m, V3DM_ERR_ARG, "mlp_save: write to %s failed", path);
   return V3DM_OK;
}

int v3dm_mlp_load(v3dm_mlp *m, const char *path)
{
   if (!m || !path) return V3DM_ERR_ARG;
   FILE *f = fopen(path, "rb");
   if (!f) return mfail(m, V3DM_ERR_ARG, "mlp_load: cannot open %s", path);
   std::vector<uint8_t> all;
   uint8_t buf[65536];
   size_t got;
   while ((got = fread(buf, 1, sizeof(buf), f)) > 0) all.insert(all.end(), buf, buf + got);
   fclose(f);
   const size_t payloadBytes = 2 * (size_t)(kOffB1 + kOut * kHid) + 4 * (size_t)(kHid + kOut);
   if (all.size() != 8 + 24 + payloadBytes || memcmp(all.data(), kMlpMagic, 8) != 0)
      return mfail(m, V3DM_ERR_ARG, "mlp_load: %s is not a saved network (%zu bytes)", path, all.size());
   uint32_t head[6];
   for (int k = 0; k < 6; ++k)
      head[k] = (uint32_t)all[8 + 4 * k] | ((uint32_t)all[9 + 4 * k] << 8) | ((uint32_t)all[10 + 4 * k] << 16) |
                ((uint32_t)all[11 + 4 * k] << 24);
   if (head[0] != kMlpVersion || head[1] != (uint32_t)kIn || head[2] != (uint32_t)kHid || head[3] != (uint32_t)kOut ||
       head[4] != (uint32_t)kParams)
      return mfail(m, V3DM_ERR_UNSUPPORTED, "mlp_load: %s: version %u, shape %u-%u-%u, %u parameters", path, head[0],
                   head[1], head[2], head[3], head[4]);
   const std::vector<uint8_t> payload(all.begin() + 32, all.end());
   if (fnv1a(payload) != head[5]) return mfail(m, V3DM_ERR_ARG, "mlp_load: %s: checksum mismatch", path);
   std::vector<float> p(kParams);
   size_t o = 0;
   for (int i = 0; i < kParams; ++i)
   {
      if (isWeight(i))
      {
         p[i] = bf16_value((uint16_t)(payload[o] | (payload[o + 1] << 8)));
         o += 2;
      }
      else
      {
         const uint32_t u = (uint32_t)payload[o] | ((uint32_t)payload[o + 1] << 8) | ((uint32_t)payload[o + 2] << 16) |
                            ((uint32_t)payload[o + 3] << 24);
         memcpy(&p[i], &u, 4);
         o += 4;
      }
   }
   return v3dm_mlp_set_params(m, p.data());
}

int v3dm_mlp_set_params(v3dm_mlp *m, const float *params)
{
   if (!m || !params) return V3DM_ERR_ARG;
   MLPCHK(m, hipSetDevice(m->device));
   MLPCHK(m, hipStreamSynchronize(m->stream));
   MLPCHK(m, hipMemcpy(m->dP, params, sizeof(float) * kParams, hipMemcpyHostToDevice));
   MLPCHK(m, hipMemset(m->dM, 0, sizeof(float) * kParams));
   MLPCHK(m, hipMemset(m->dV, 0, sizeof(float) * kParams));
   m->step = 0;
   int rc = refreshForward(m);
   if (rc) return rc;
   MLPCHK(m, hipStreamSynchronize(m->stream));
   return V3DM_OK;
}

int v3dm_mlp_loss_grad(v3dm_mlp *m, int batch, const uint8_t *left, const uint8_t *right, float temperature,
                       float *loss, float *grad)
{
   if (!m) return V3DM_ERR_ARG;
   if (batch < 2 || !left || !right || !(temperature > 0))
      return mfail(m, V3DM_ERR_ARG, "mlp_loss_grad: need batch >= 2, descriptors and temperature > 0");
   MLPCHK(m, hipSetDevice(m->device));
   int rc = ensureBatch(m, batch);
   if (rc) return rc;
   MLPCHK(m, hipMemcpyAsync(m->dBL, left, (size_t)batch * kIn, hipMemcpyHostToDevice, m->stream));
   MLPCHK(m, hipMemcpyAsync(m->dBR, right, (size_t)batch * kIn, hipMemcpyHostToDevice, m->stream));
   MLPCHK(m, launch_u8_rows_to_x(batch, m->dBL, nullptr, m->dX, m->stream));
   MLPCHK(m, launch_u8_rows_to_x(batch, m->dBR, nullptr, m->dX + (size_t)batch * kIn, m->stream));
   rc = forwardBackward(m, batch, temperature, m->dLoss);
   if (rc) return rc;
   if (loss) MLPCHK(m, hipMemcpyAsync(loss, m->dLoss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
   if (grad) MLPCHK(m, hipMemcpyAsync(grad, m->dG, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
   MLPCHK(m, hipStreamSynchronize(m->stream));
   return V3DM_OK;
}

int v3dm_mlp_set_training_pairs(v3dm_mlp *m, int64_t n, const uint8_t *left, const uint8_t *right)
{
   if (!m) return V3DM_ERR_ARG;
   if (n < 1 || !left || !right) return mfail(m, V3DM_ERR_ARG, "mlp_set_training_pairs: empty pair set");
   MLPCHK(m, hipSetDevice(m->device));
   MLPCHK(m, hipStreamSynchronize(m->stream));
   dfree(m->dTL);
   dfree(m->dTR);
   m->nTrain = 0;
   MLPCHK(m, hipMalloc((void **)&m->dTL, (size_t)n * kIn));
   MLPCHK(m, hipMalloc((void **)&m->dTR, (size_t)n * kIn));
   MLPCHK(m, hipMemcpy(m->dTL, left, (size_t)n * kIn, hipMemcpyHostToDevice));
   MLPCHK(m, hipMemcpy(m->dTR, right, (size_t)n * kIn, hipMemcpyHostToDevice));
   m->nTrain = n;
   return V3DM_OK;
}

int v3dm_mlp_train(v3dm_mlp *m, int steps, int batch, float lr, float temperature, uint64_t seed, float *losses)
{
   if (!m) return V3DM_ERR_ARG;
   if (m->nTrain < 1) return mfail(m, V3DM_ERR_STATE, "mlp_train: no training pairs set");
   if (steps < 0 || batch < 2 || !(temperature > 0) || !(lr >= 0))
      return mfail(m, V3DM_ERR_ARG, "mlp_train: bad arguments");
   MLPCHK(m, hipSetDevice(m->device));
   int rc = ensureBatch(m, batch);
   if (rc) return rc;
   if (steps > m->lossHistCap)
   {
      dfree(m->dLossHist);
      m->lossHistCap = 0;
      MLPCHK(m, hipMalloc((void **)&m->dLossHist, sizeof(float) * std::max(steps, 1)));
      m->lossHistCap = std::max(steps, 1);
   }
   for (int t = 0; t < steps; ++t)
   {
      MLPCHK(m, launch_sample(batch, seed, (uint64_t)m->step, m->nTrain, m->dIdx, m->stream));
      MLPCHK(m, launch_u8_rows_to_x(batch, m->dTL, m->dIdx, m->dX, m->stream));
      MLPCHK(m, launch_u8_rows_to_x(batch, m->dTR, m->dIdx, m->dX + (size_t)batch * kIn, m->stream));
      rc = forwardBackward(m, batch, temperature, m->dLossHist + t);
      if (rc) return rc;
      rc = adamStep(m, lr);
      if (rc) return rc;
   }
   if (losses && steps > 0)
      MLPCHK(m, hipMemcpyAsync(losses, m->dLossHist, sizeof(float) * steps, hipMemcpyDeviceToHost, m->stream));
   MLPCHK(m, hipStreamSynchronize(m->stream));
   return V3DM_OK;
}

int v3dm_mlp_embed(v3dm_mlp *m, int n, const uint8_t *desc, float *out)
{
   if (!m) return V3DM_ERR_ARG;
   if (n < 0 || (n > 0 && (!desc || !out))) return mfail(m, V3DM_ERR_ARG, "mlp_embed: bad arguments");
   if (n == 0) return V3DM_OK;
   MLPCHK(m, hipSetDevice(m->device));
   if (n > m->iU8Cap)
   {
      dfree(m->iU8);
      dfree(m->iE);
      m->iU8Cap = 0;
      MLPCHK(m, hipMalloc((void **)&m->iU8, (size_t)n * kIn));
      MLPCHK(m, hipMalloc((void **)&m->iE, sizeof(float) * (size_t)n * kOut));
      m->iU8Cap = n;
   }
   MLPCHK(m, hipMemcpyAsync(m->iU8, desc, (size_t)n * kIn, hipMemcpyHostToDevice, m->stream));
   int rc = embedRows(m, m->stream, n, m->iU8, m->iE, nullptr, 0);
   if (rc) return rc;
   MLPCHK(m, hipMemcpyAsync(out, m->iE, sizeof(float) * (size_t)n * kOut, hipMemcpyDeviceToHost, m->stream));
   MLPCHK(m, hipStreamSynchronize(m->stream));
   return V3DM_OK;
}

int v3dm_learned_find(v3dm_ctx *c, v3dm_mlp *m, int topK, float margin, float minScore, float minRatio,
                      int32_t *pairs, int capacity, int *count, int64_t *stats)
{
   if (!c) return V3DM_ERR_ARG;
   if (!m) return fail(c, V3DM_ERR_ARG, "learned_find: network is NULL");
   if (topK != 16 && topK != 32) return fail(c, V3DM_ERR_ARG, "learned_find: topK must be 16 or 32");
   if (capacity < 0) return fail(c, V3DM_ERR_ARG, "learned_find: negative capacity");
   if (!c->allocated) return fail(c, V3DM_ERR_STATE, "learned_find: allocate() has not been called");
   if (m->device != c->device) return fail(c, V3DM_ERR_ARG, "learned_find: network and matcher on different devices");
   if (!c->L.isU8 || !c->R.isU8)
      return fail(c, V3DM_ERR_UNSUPPORTED, "learned_find: both sides must be uploaded as uint8 descriptors");
   HIPCHK(c, hipSetDevice(c->device));
   HIPCHK(c, hipStreamSynchronize(m->stream)); // weights of finished training steps
   if (c->pendingAsync)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      c->pendingAsync = false;
   }
   c->masks.clear();
   int rc = uploadPair(c, 0);
   if (rc) return rc;
   rc = useKeyFormat(c, 0);
   if (rc) return rc;
   const int nL = c->L.n, nR = c->R.n;
   if (!c->learned) c->learned = new LearnedState();
   LearnedState *s = c->learned;
   const int capL = std::max(roundUp(std::max(c->L.nMax, 1), kRowPad), kRowPad);
   const int capR = std::max(roundUp(std::max(c->R.nMax, 1), kRowPad), kRowPad);
   if (s->capL < capL || s->capR < capR)
   {
      dfree(s->embL);
      dfree(s->embR);
      dfree(s->cand);
      dfree(s->ovf);
      dfree(s->flagL);
      dfree(s->flagR);
      dfree(s->gI8);
      dfree(s->gSum);
      s->capL = s->capR = 0;
      HIPCHK(c, hipMalloc((void **)&s->embL, sizeof(uint16_t) * (size_t)capL * kOut));
      HIPCHK(c, hipMalloc((void **)&s->embR, sizeof(uint16_t) * (size_t)capR * kOut));
      HIPCHK(c, hipMalloc((void **)&s->cand, sizeof(int) * (size_t)capL * kMaxCandK));
      HIPCHK(c, hipMalloc((void **)&s->ovf, sizeof(int) * (size_t)capL));
      HIPCHK(c, hipMalloc((void **)&s->flagL, sizeof(int) * (size_t)capL));
      HIPCHK(c, hipMalloc((void **)&s->flagR, sizeof(int) * (size_t)capR));
      const int g = std::max(capL, capR);
      HIPCHK(c, hipMalloc((void **)&s->gI8, (size_t)g * kDim));
      HIPCHK(c, hipMalloc((void **)&s->gSum, sizeof(int32_t) * (size_t)g));
      s->capL = capL;
      s->capR = capR;
   }
   if (!s->dCount)
   {
      HIPCHK(c, hipMalloc((void **)&s->dCount, 2 * sizeof(int)));
      HIPCHK(c, hipHostMalloc((void **)&s->hCount, 2 * sizeof(int), hipHostMallocDefault));
      HIPCHK(c, hipMalloc((void **)&s->dPairFb, 2 * sizeof(PairDesc)));
   }
   const int nChunks = emb_topk_chunks(nR, topK);
   const size_t needScratch = emb_topk_scratch_bytes(nL, nChunks);
   if (needScratch > s->scratchBytes)
   {
      dfree(s->scratch);
      s->scratchBytes = 0;
      HIPCHK(c, hipMalloc(&s->scratch, needScratch));
      s->scratchBytes = needScratch;
   }
   const PairDesc &pd = c->hPairUploaded;
   int64_t nFlag[2] = {0, 0};
   // this job's kernels read the epoch from device memory (its finalize
   // advances it there); the host's count is the current one (single-pair
   // int8 and float jobs pass it by value and leave device memory alone)
   HIPCHK(c, launch_fill_i32(1, (int32_t)c->hostEpoch, reinterpret_cast<int32_t *>(c->dEpoch), c->stream));
   if (nL > 0 && nR > 0)
   {
      hipStream_t st = c->stream;
      // 1. embeddings, 2. top-16 retrieval, 3. exact rescoring of the candidates
      rc = embedRows(m, st, nL, c->L.dRawU8, nullptr, s->embL, roundUp(nL, 32));
      if (rc) return fail(c, rc, "learned_find: %s", m->err.c_str());
      rc = embedRows(m, st, nR, c->R.dRawU8, nullptr, s->embR, roundUp(nR, 32));
      if (rc) return fail(c, rc, "learned_find: %s", m->err.c_str());
      HIPCHK(c, launch_emb_topk(nL, nR, topK, s->embL, s->embR, nChunks, s->scratch, s->cand, nullptr, s->ovf, st));
      HIPCHK(c, launch_cand_insert(nL, nR, topK, s->cand, c->L.dI8, c->R.dI8, c->L.dTop1, c->L.dTop2, c->R.dTop1,
                                   c->R.dTop2, c->dEpoch, st));
      // 4. margin flags
      HIPCHK(c, hipMemsetAsync(s->dCount, 0, 2 * sizeof(int), st));
      HIPCHK(c, launch_flag(nL, 0, c->L.dTop1, c->L.dTop2, c->dEpoch, margin, s->ovf, s->flagL, s->dCount, st));
      HIPCHK(c, launch_flag(nR, 1, c->R.dTop1, c->R.dTop2, c->dEpoch, margin, nullptr, s->flagR, s->dCount + 1,
                            st));
      // 5. exhaustive int8 scans of the flagged lefts (all rights) and flagged
      // rights (all lefts), with no host round trip: the flag counts stay on
      // the device, the gather and the scans are launched for the largest
      // count (every row flagged) and read the real one -- the scans from
      // pair descriptors a kernel writes from the counts
      HIPCHK(c, launch_fallback_pairs(s->dCount, nL, nR, scan_variant_tile_m(0), scan_variant_tile_n(0), s->dPairFb,
                                      st));
      for (int side = 0; side < 2; ++side)
      {
         const int nMax = side == 0 ? nL : nR;
         const int nPad = std::max(roundUp(nMax, kRowPad), kRowPad);
         const Side &src = side == 0 ? c->L : c->R;
         HIPCHK(c, launch_gather_rows(nMax, nPad, s->dCount + side, side == 0 ? s->flagL : s->flagR, src.dI8, src.dSum,
                                      s->gI8, s->gSum, st));
         PairDesc worst{};
         worst.nL = nL;
         worst.nR = nR;
         worst.tilesM = (roundUp(nL, kRowPad) + scan_variant_tile_m(0) - 1) / scan_variant_tile_m(0);
         worst.tilesN = (roundUp(nR, kRowPad) + scan_variant_tile_n(0) - 1) / scan_variant_tile_n(0);
         ScanArgs a{};
         a.pair0 = worst; // not read: pairsOnDevice
         a.L = side == 0 ? s->gI8 : c->L.dI8;
         a.Lsum = side == 0 ? s->gSum : c->L.dSum;
         a.R = side == 0 ? c->R.dI8 : s->gI8;
         a.Rsum = side == 0 ? c->R.dSum : s->gSum;
         a.pairs = s->dPairFb + side;
         a.nPairs = 1;
         a.pairsOnDevice = 1;
         a.nTiles = worst.tilesM * worst.tilesN;
         a.thr = survivorThreshold(minScore, minRatio);
         a.top1L = c->L.dTop1;
         a.top2L = c->L.dTop2;
         a.top1R = c->R.dTop1;
         a.top2R = c->R.dTop2;
         a.epoch = c->dEpoch;
         a.surv = c->dSurv;
         a.survCount = c->dSurvCount;
         a.survCap = scan_variant_surv_cap(0);
         a.waves = scan_variant_waves(0);
         a.tileM = scan_variant_tile_m(0);
         a.tileN = scan_variant_tile_n(0);
         a.tileCounter = c->dEpoch + 2;
         a.persistent = 0;
         a.mapL = side == 0 ? s->flagL : nullptr;
         a.mapR = side == 0 ? nullptr : s->flagR;
         a.inlineInsert = a.nTiles <= scan_variant_resident(0, c->device) ? 1 : 0;
         HIPCHK(c, launch_i8_scan(0, a.nTiles, a, st));
         if (!a.inlineInsert)
            HIPCHK(c, launch_i8_insert(a, insert_form_for((int64_t)a.nTiles * a.waves, c->device, c->insertForm), st));
      }
      // the counts for the statistics come back with the match table
      HIPCHK(c, hipMemcpyAsync(s->hCount, s->dCount, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
   }
   // 6. the exact matcher's ratio test and mutual check on the scored pairs
   FinalizeArgs f{};
   f.pairs = c->dPair;
   f.nPairs = 1;
   f.nBlocks = std::max(1, (pd.nL + 255) / 256);
   f.top1L = c->L.dTop1;
   f.top2L = c->L.dTop2;
   f.top1R = c->R.dTop1;
   f.top2R = c->R.dTop2;
   f.epoch = c->dEpoch;
   f.done = c->dEpoch + 1;
   f.tileCounter = c->dEpoch + 2;
   f.minScore = minScore;
   f.minRatio = minRatio;
   f.matchOfLeft = c->dMatch;
   f.pair0 = pd;
   f.hasPair0 = 1;
   HIPCHK(c, launch_i8_finalize(f, c->stream));
   c->hostEpoch += 1;
   c->lastKernel = V3DM_KERNEL_I8;
   c->lastNL = nL;
   s->lastNL = nL;
   s->lastK = topK;
   c->pendingAsync = true;
   rc = fetch(c, pairs, capacity, count); // waits for the job
   if (nL > 0 && nR > 0)
   {
      nFlag[0] = s->hCount[0];
      nFlag[1] = s->hCount[1];
   }
   if (stats)
   {
      stats[0] = nL;
      stats[1] = nR;
      stats[2] = (int64_t)nL * topK;
      stats[3] = nFlag[0];
      stats[4] = nFlag[1];
      stats[5] = nFlag[0] * nR + (int64_t)nL * nFlag[1];
   }
   return rc;
}

int v3dm_learned_candidates(v3dm_ctx *c, int n, int32_t *cand, int *topK)
{
   if (!c) return V3DM_ERR_ARG;
   if (!c->learned || n < 0 || n > c->learned->lastNL || (n > 0 && !cand))
      return fail(c, V3DM_ERR_ARG, "learned_candidates: no learned job or n out of range");
   HIPCHK(c, hipSetDevice(c->device));
   int rc = syncStream(c);
   if (rc) return rc;
   const int K = c->learned->lastK;
   if (topK) *topK = K;
   if (n > 0) HIPCHK(c, hipMemcpy(cand, c->learned->cand, sizeof(int32_t) * (size_t)n * K, hipMemcpyDeviceToHost));
   return V3DM_OK;
}

} // extern "C"
