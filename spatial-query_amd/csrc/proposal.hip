// proposal.hip -- the learned subtree proposal: a small MLP that predicts, per
// ray, which subtrees of the BVH cut at a fixed depth hold the nearest hit.
// Its top-4 subtrees are walked first to tighten the limit; the full exact walk
// then confirms, so results stay bit-exact (the walk's result does not depend
// on visiting order, see traverse.h).
//
// Model: x = (origin normalised to the scene box in [-1,1]^3, direction);
// 64 random Fourier features sin/cos(2 pi x B), B ~ N(0, rff_scale^2) -> 128; three 256-wide ReLU
// layers; logits over the K subtrees.  Weights BF16 (FP32 master copies for
// Adam), accumulation FP32 on the MFMA units (v_mfma_f32_16x16x32_bf16).
//
// Every contraction is a "TN" GEMM  C[M x N] = A[M x K] * B[N x K]^T  with
// both operands contiguous along K, which is the natural operand order of the
// MFMA instruction (8 consecutive K values per lane).
#include <cmath>
#include <cstring>

#include "engine_internal.h"
#include "pool.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int BM = 128, BN = 128, BK = 32, LDP = BK + 8;   // LDS rows padded by 16 B
constexpr int N_RFF = 64, F_IN = 2 * N_RFF, HID = 256;

// Epilogues of the GEMM.
enum { EPI_F32 = 0,          // C (fp32)
       EPI_BIAS_RELU_BF16,   // relu(C + bias) (bf16)
       EPI_BIAS_F32,         // C + bias (fp32)
       EPI_MASK_BF16,        // C * (mask > 0) (bf16): backward through a ReLU
       EPI_PARTIAL };        // C (fp32) into partial buffer `split` (split-K)

__device__ __forceinline__ float bf2f(bf16 x) { return (float)x; }
__device__ __forceinline__ bf16 f2bf(float x) { return (bf16)x; }

// One 128x128 block tile of C = A * B^T over K range [k0, k1).
// 256 threads = 4 waves in a 2x2 arrangement; each wave owns 64x64 = 4x4
// MFMA tiles of 16x16.  MFMA 16x16x32 operand layout: lane l supplies row
// (l % 16) and K values [8 (l / 16), 8 (l / 16) + 8) of its 16x32 A slice, and
// column (l % 16), same K values of its 32x16 B slice; it receives rows
// 4 (l / 16) + r, r < 4, of column (l % 16) of the 16x16 result.
__device__ __forceinline__ void tile_mma(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb, int m0, int n0,
                                         int k0, int k1, bf16 (*As)[LDP], bf16 (*Bs)[LDP], f32x4 acc[4][4]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int lr = tid >> 2, lc = (tid & 3) * 8;   // loader: 64 rows x 4 chunks of 8 per pass
  for (int k = k0; k < k1; k += BK) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int r = lr + 64 * p;
      *reinterpret_cast<bf16x8*>(&As[r][lc]) = *reinterpret_cast<const bf16x8*>(A + (size_t)(m0 + r) * lda + k + lc);
      *reinterpret_cast<bf16x8*>(&Bs[r][lc]) = *reinterpret_cast<const bf16x8*>(B + (size_t)(n0 + r) * ldb + k + lc);
    }
    __syncthreads();
    const int kb = (lane >> 4) * 8, li = lane & 15;
    bf16x8 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + li][kb]);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * j + li][kb]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
}

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_tn(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb,
                                                 int K, int k_chunk, float* Cf, bf16* Cb, int ldc, size_t split_stride,
                                                 const float* __restrict__ bias, const bf16* __restrict__ mask, int ldm) {
  __shared__ bf16 As[BM][LDP];
  __shared__ bf16 Bs[BN][LDP];
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int k0 = blockIdx.z * k_chunk, k1 = min(K, k0 + k_chunk);
  f32x4 acc[4][4];
  tile_mma(A, lda, B, ldb, m0, n0, k0, k1, As, Bs, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int col = n0 + wn + 16 * j + (lane & 15);
        float v = acc[i][j][r];
        const size_t o = (size_t)row * ldc + col;
        if constexpr (EPI == EPI_F32) Cf[o] = v;
        else if constexpr (EPI == EPI_PARTIAL) Cf[blockIdx.z * split_stride + o] = v;
        else if constexpr (EPI == EPI_BIAS_F32) Cf[o] = v + bias[col];
        else if constexpr (EPI == EPI_BIAS_RELU_BF16) Cb[o] = f2bf(fmaxf(v + bias[col], 0.0f));
        else Cb[o] = f2bf(bf2f(mask[(size_t)row * ldm + col]) > 0.0f ? v : 0.0f);
      }
}

// Sum of split-K partials (fixed order: deterministic).
__global__ void k_sum_splits(const float* __restrict__ part, int splits, size_t n, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int p = 0; p < splits; ++p) s += part[p * n + i];
  out[i] = s;
}

// Network input of a ray: origin mapped from the scene box to [-1, 1]^3, and
// the direction.
__device__ __forceinline__ void ray_features(const float* r, const float* __restrict__ box, float* x) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x[k] = (r[k] - box[k]) / (box[3 + k] - box[k]) * 2.0f - 1.0f;
    x[3 + k] = r[3 + k];
  }
}

// Random Fourier features F[m] = (sin 2 pi x.B, cos 2 pi x.B) (bf16, [rows][128]);
// rows >= n are zero.
__global__ void k_features(const float* __restrict__ rays, uint64_t n, uint64_t rows, const float* __restrict__ Bm,
                           const float* __restrict__ box, bf16* F) {
  const uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= rows) return;
  float x[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (m < n) ray_features(rays + m * 6, box, x);
  bf16* f = F + m * F_IN;
  for (int j = 0; j < N_RFF; ++j) {
    float z = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) z += x[k] * Bm[k * N_RFF + j];
    const float a = 6.28318530717958647692f * z;
    f[j] = f2bf(m < n ? sinf(a) : 0.0f);
    f[N_RFF + j] = f2bf(m < n ? cosf(a) : 0.0f);
  }
}

// Softmax cross-entropy over the first n_cls logits of each row; label < 0
// rows are ignored.  dZ = (softmax - onehot) * inv_count as bf16 (padded
// columns 0); loss[m] = -log p[label] (0 when ignored).  One wave per row.
__global__ __launch_bounds__(256) void k_softmax_ce(const float* __restrict__ Z, int ldz, int n_cls, const int32_t* __restrict__ label,
                                                    uint64_t rows, const uint32_t* __restrict__ count, bf16* dZ, float* loss) {
  const uint64_t m = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (m >= rows) return;
  const float* z = Z + m * ldz;
  const int y = label[m] < n_cls ? label[m] : -1;   // out-of-range labels are ignored
  float mx = -INFINITY;
  for (int c = lane; c < n_cls; c += 64) mx = fmaxf(mx, z[c]);
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  float se = 0.0f;
  for (int c = lane; c < n_cls; c += 64) se += expf(z[c] - mx);
  for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o);
  const float inv = 1.0f / se;
  const float inv_count = *count ? 1.0f / (float)*count : 0.0f;
  bf16* g = dZ + m * ldz;
  for (int c = lane; c < ldz; c += 64) {
    float v = 0.0f;
    if (y >= 0 && c < n_cls) v = (expf(z[c] - mx) * inv - (c == y ? 1.0f : 0.0f)) * inv_count;
    g[c] = f2bf(v);
  }
  if (lane == 0) loss[m] = y >= 0 ? -(z[y] - mx - logf(se)) : 0.0f;
}

// out[c][r] = in[r][c] for a rows x cols bf16 matrix (both multiples of 64).
__global__ void k_transpose(const bf16* __restrict__ in, int rows, int cols, bf16* out) {
  __shared__ bf16 t[64][65];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  for (int i = threadIdx.y; i < 64; i += blockDim.y) t[i][threadIdx.x] = in[(size_t)(r0 + i) * cols + c0 + threadIdx.x];
  __syncthreads();
  for (int i = threadIdx.y; i < 64; i += blockDim.y) out[(size_t)(c0 + i) * rows + r0 + threadIdx.x] = t[threadIdx.x][i];
}

// Column sums of a rows x cols bf16 matrix (bias gradients): block (x, y)
// sums rows [y * chunk, (y + 1) * chunk) of 64 columns, 4 row lanes, into
// part[y][col]; k_sum_splits adds the chunks (fixed order: deterministic).
__global__ __launch_bounds__(256) void k_colsum(const bf16* __restrict__ g, uint64_t rows, int cols, uint64_t chunk, float* part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
  const uint64_t r0 = blockIdx.y * chunk, r1 = r0 + chunk < rows ? r0 + chunk : rows;
  float s = 0.0f;
  for (uint64_t r = r0 + q; r < r1; r += 4) s += bf2f(g[r * cols + c]);
  red[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q == 0) part[(size_t)blockIdx.y * cols + c] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// Number of labelled rows (0 <= label < n_cls) into *count (cleared by the caller).
__global__ void k_count_labels(const int32_t* __restrict__ label, uint64_t n, int n_cls, uint32_t* count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool lab = i < n && label[i] >= 0 && label[i] < n_cls;
  const uint64_t b = __ballot(lab);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (uint32_t)__popcll(b));
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Adam on fp32 masters (bias-corrected), then the bf16 working copies:
// w_nk[n][k] (forward operand) and, when w_kn != null, w_kn[k][n] (backward).
__global__ void k_adam(float* w, float* mo, float* ve, const float* __restrict__ g, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, bf16* w_nk, bf16* w_kn, int rows, int cols) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float m = b1 * mo[i] + (1.0f - b1) * gi;
  const float v = b2 * ve[i] + (1.0f - b2) * gi * gi;
  mo[i] = m; ve[i] = v;
  const float wi = w[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  w[i] = wi;
  if (w_nk) {
    w_nk[i] = f2bf(wi);
    if (w_kn) { const int r = (int)(i / cols), c = (int)(i % cols); w_kn[(size_t)c * rows + r] = f2bf(wi); }
  }
}

// Floats as unsigned integers in the same order (LDS atomic max of a float).
__device__ __forceinline__ uint32_t ordered_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// (v, c) ranks before (w, d): larger logit first, then the smaller class id.
__device__ __forceinline__ bool beats(float v, int c, float w, int d) { return v > w || (v == w && c < d); }

// Insert (v, c) into a best-first list of four, all in registers (a chain of
// selects: no dynamic indexing).
__device__ __forceinline__ void top4_insert(float (&tv)[4], int (&ti)[4], float v, int c) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool b = beats(v, c, tv[k], ti[k]);
    const float ov = tv[k];
    const int oc = ti[k];
    tv[k] = b ? v : ov;
    ti[k] = b ? c : oc;
    v = b ? ov : v;
    c = b ? oc : c;
  }
}

// Insert (v, c) into a best-first list of four whose entries all have smaller
// class ids than c and whose fourth v beats (v > tv[3]): v goes before every
// entry it exceeds and after every entry it ties (the smaller id ranks first),
// the entries below it shift down one place, branch-free.
__device__ __forceinline__ void top4_insert_after(float (&tv)[4], int (&ti)[4], float v, int c) {
  const bool b0 = v > tv[0], b1 = v > tv[1], b2 = v > tv[2];
  const float v3 = b2 ? tv[2] : v, v2 = b1 ? tv[1] : (b2 ? v : tv[2]);
  const float v1 = b0 ? tv[0] : (b1 ? v : tv[1]), v0 = b0 ? v : tv[0];
  const int c3 = b2 ? ti[2] : c, c2 = b1 ? ti[1] : (b2 ? c : ti[2]);
  const int c1 = b0 ? ti[0] : (b1 ? c : ti[1]), c0 = b0 ? c : ti[0];
  tv[0] = v0; tv[1] = v1; tv[2] = v2; tv[3] = v3;
  ti[0] = c0; ti[1] = c1; ti[2] = c2; ti[3] = c3;
}

// Fused inference: features -> 3 ReLU layers -> logits -> top-4, for IR = 128
// rays per block of 512 threads (8 waves).  Activations stay in LDS (one
// [128][HID + 8] bf16 buffer, overwritten layer by layer); weights stream from
// L2 in fragment order (k_hidden_fragments, k_w4_fragments: each wave's
// operand fragment is 1 KB of consecutive memory, eight whole cache lines per
// load).  Hidden layers: wave w owns rays 64 (w >> 2) .. +64 and outputs
// 64 (w & 3) .. +64 (4 x 4 MFMA tiles of 16 x 16).  The last layer is computed
// transposed (classes as MFMA rows, rays as columns) on 32 x 32 MFMA tiles:
// per 128-class tile, wave w owns classes 32 (w >> 1) .. +32 and rays
// 64 (w & 1) .. +64, so each lane holds 16 classes of each of its 2 rays and
// keeps every ray's best four in registers, skipping candidates below the
// ray's shared threshold; the 8 partial lists of a ray (4 waves x 2 lane
// halves) are merged in LDS.
constexpr int IR = 128, ALD = HID + 8, IT = 512;

// A hidden layer's W [HID][KIN] (bf16) in the order dense_relu reads it: for
// K step s (K 32 s ..), output group g (outputs 64 g ..), sub-tile j (outputs
// + 16 j ..) and lane l, the 8 values lane l feeds to one 16 x 16 x 32 MFMA
// (output 64 g + 16 j + l % 16, K 32 s + 8 (l / 16) ..) are contiguous: one
// fragment is 1 KB of consecutive memory.
__global__ void k_hidden_fragments(const bf16* __restrict__ w, uint32_t kin, bf16* out) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (uint32_t)HID * kin / 8) return;
  const uint32_t l = f % 64, j = (f / 64) % 4, g = (f / 256) % 4, s = f / 1024;
  *reinterpret_cast<bf16x8*>(out + (size_t)f * 8) =
      *reinterpret_cast<const bf16x8*>(w + (size_t)(64 * g + 16 * j + l % 16) * kin + 32 * s + 8 * (l / 16));
}

template <int KIN>
__device__ __forceinline__ void dense_relu(bf16 (*act)[ALD], const bf16* __restrict__ Wf, const float* __restrict__ bias) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = (wave >> 2) * 64, c0 = (wave & 3) * 64;
  const int li = lane & 15, kb = (lane >> 4) * 8;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // the W fragments (fragment order: ((s * 4 + output group) * 4 + j) * 64 +
  // lane) of the next k slice load while this slice's MFMAs run
  const bf16* wf = Wf + ((size_t)(wave & 3) * 4 * 64 + lane) * 8;
  bf16x8 b_next[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) b_next[j] = *reinterpret_cast<const bf16x8*>(wf + j * 64 * 8);
#pragma unroll 1
  for (int k = 0; k < KIN; k += 32) {
    bf16x8 a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = b_next[j];
    if (k + 32 < KIN)
#pragma unroll
      for (int j = 0; j < 4; ++j) b_next[j] = *reinterpret_cast<const bf16x8*>(wf + ((size_t)(k / 32 + 1) * 16 + j) * 64 * 8);
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&act[r0 + 16 * i + li][k + kb]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  __syncthreads();   // every wave has read the layer input
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = c0 + 16 * j + li;
    const float bj = bias[col];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) act[r0 + 16 * i + 4 * (lane >> 4) + r][col] = f2bf(fmaxf(acc[i][j][r] + bj, 0.0f));
  }
  __syncthreads();
}

struct MlpWeights {
  const bf16* w[4];     // [out_pad][in] bf16
  const bf16* w4f;      // the last layer's weights in fragment order (k_w4_fragments)
  const bf16* wf[3];    // the hidden layers' weights in fragment order (k_hidden_fragments)
  const float* b[4];    // [out_pad]
  const float* rff;     // [6][N_RFF]
  const float* box;     // scene min[3], max[3]
};

// The last layer's W4 [n_pad][HID] (bf16) in the order k_infer_top4 reads it:
// for tile T (classes 128 T ..), step s (K 32 s ..), class group c (classes
// + 32 c ..), slice sl (K + 16 sl ..) and lane l, the 8 values lane l feeds
// to one 32 x 32 x 16 MFMA (class 128 T + 32 c + l % 32, K 32 s + 16 sl +
// 8 (l / 32) ..) are contiguous, so a wave reads each fragment as 1 KB of
// consecutive memory: eight whole cache lines per load instruction instead of
// 32-byte pieces of 32 rows.
__global__ void k_w4_fragments(const bf16* __restrict__ w4, uint32_t n_pad, bf16* out) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;   // one 8-value fragment piece per thread
  const uint64_t pieces = (uint64_t)n_pad * HID / 8;
  if (f >= pieces) return;
  const uint32_t l = (uint32_t)(f % 64);
  uint64_t r = f / 64;
  const uint32_t sl = (uint32_t)(r % 2); r /= 2;
  const uint32_t c = (uint32_t)(r % 4); r /= 4;
  const uint32_t s = (uint32_t)(r % (HID / 32)); r /= HID / 32;
  const uint64_t T = r;
  const uint64_t row = 128 * T + 32 * c + (l % 32);
  const uint32_t k = 32 * s + 16 * sl + 8 * (l / 32);
  *reinterpret_cast<bf16x8*>(out + f * 8) = *reinterpret_cast<const bf16x8*>(w4 + row * HID + k);
}

__global__ __launch_bounds__(IT, 4) void k_infer_top4(const float* __restrict__ rays, uint64_t n, MlpWeights mw, int n_cls, int n_pad,
                                                      uint16_t* top) {
  __shared__ __align__(16) bf16 act[IR][ALD];
  // Per ray, the largest fourth-best logit any of its partial lists holds so
  // far (ordered bits): the ray's final fourth-best is at least that large, so
  // a candidate below it cannot enter the ray's top four and is skipped.
  __shared__ uint32_t thr[IR];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t ray0 = (uint64_t)blockIdx.x * IR;
  if (tid < IR) thr[tid] = ordered_bits(-INFINITY);
  // features: lane j of every wave computes frequency j for 16 rays (one ray
  // at a time for the whole wave: the ray's inputs are wave-uniform, the
  // Fourier column j stays in registers)
  {
    const int j = lane;
    const int w8 = __builtin_amdgcn_readfirstlane(wave);
    float bj[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) bj[k] = mw.rff[k * N_RFF + j];
#pragma unroll 1
    for (int i = 0; i < IR / 8; ++i) {
      const int rr = w8 * (IR / 8) + i;
      const uint64_t m = ray0 + rr;
      float x[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      if (m < n) ray_features(rays + m * 6, mw.box, x);
      float z = 0.0f;
#pragma unroll
      for (int k = 0; k < 6; ++k) z += x[k] * bj[k];
      const float a = 6.28318530717958647692f * z;
      float sv, cv;
      sincosf(a, &sv, &cv);                      // one argument reduction for both
      act[rr][j] = f2bf(m < n ? sv : 0.0f);
      act[rr][N_RFF + j] = f2bf(m < n ? cv : 0.0f);
    }
  }
  __syncthreads();
  dense_relu<F_IN>(act, mw.wf[0], mw.b[0]);
  dense_relu<HID>(act, mw.wf[1], mw.b[1]);
  dense_relu<HID>(act, mw.wf[2], mw.b[2]);
  // Last layer, transposed (classes as MFMA rows, rays as columns), on the
  // 32x32x16 MFMA: per 128-class tile wave w owns classes cr = 32 (w >> 1) ..
  // +32 and the two 32-ray tiles rr + 32 j, rr = 64 (w & 1); lane l holds, for
  // ray rr + 32 j + (l & 31), the classes m0 + cr + 8 (r >> 2) + 4 (l >> 5) +
  // (r & 3), r < 16, and keeps that ray's best four in registers (two rays per
  // lane).  Operands: lane l supplies class (or ray) l & 31 and K values
  // 8 (l >> 5) .. +8 of a 16-wide K slice.
  const int cr = (wave >> 1) * 32, rr = (wave & 1) * 64;
  const int l32 = lane & 31, kh = (lane >> 5) * 8;
  float tv[2][4];
  int ti[2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int k = 0; k < 4; ++k) { tv[j][k] = -INFINITY; ti[j][k] = 0x7fffffff; }
  const float* b4 = mw.b[3];
  // One step = one 32-wide K slice of one 128-class tile (two 16-wide MFMA
  // slices per ray tile).  The W4 stream from L2 is what bounds this loop, so
  // its fragments run ahead in a ring of four steps' registers, loaded two
  // steps at a time, two to three steps before use.  The ray
  // fragments come from LDS at the start of each step; each tile's 32 biases
  // of this wave arrive by scalar loads (the same for every lane).
  const int ksteps = HID / 32;
  const int nsteps = (n_pad / 128) * ksteps;     // a multiple of 8
  const bf16* W4f = mw.w4f + ((size_t)(cr / 32) * 2 * 64 + lane) * 8;
  auto w4 = [&](int step, bf16x8 (&a)[2]) {
    // fragment order: ((step * 4 + class group) * 2 + slice) * 64 + lane
    const bf16* f = W4f + (size_t)step * (4 * 2 * 64 * 8);
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) a[sl] = *reinterpret_cast<const bf16x8*>(f + sl * (64 * 8));
  };
  typedef const float __attribute__((address_space(4)))* const_fp;
  f32x16 acc[2];
  auto tile_step = [&](int step, const bf16x8 (&a)[2]) {
    const int kk = step % ksteps;
    const int tile = step / ksteps, m0 = tile * 128;
    bf16x8 b[2][2];
#pragma unroll
    for (int sl = 0; sl < 2; ++sl)
#pragma unroll
      for (int j = 0; j < 2; ++j) b[sl][j] = *reinterpret_cast<const bf16x8*>(&act[rr + 32 * j + l32][kk * 32 + 16 * sl + kh]);
    if (kk == 0) {
      // classes m0 + cr + 8 g + 4 h + r of lane half h = lane >> 5
      const const_fp bs = (const_fp)(b4 + m0 + cr);
      float bw[32];
#pragma unroll
      for (int k = 0; k < 32; ++k) bw[k] = bs[k];     // wave-uniform: scalar loads
      const bool hi = (lane >> 5) != 0;
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = hi ? bw[8 * g + 4 + r] : bw[8 * g + r];
          acc[0][4 * g + r] = v;
          acc[1][4 * g + r] = v;
        }
    }
#pragma unroll
    for (int sl = 0; sl < 2; ++sl)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[sl], b[sl][j], acc[j], 0, 0, 0);
    if (kk == ksteps - 1) {
      // Per ray, the best of its 16 candidates against th, the larger of this
      // list's fourth and the ray's shared threshold: below th nothing can
      // enter the ray's top four (a tie with another list's fourth needs >=),
      // so the insertion runs only for candidates at or above it.  A lane meets
      // its classes in increasing order (tiles, then r), so a candidate enters
      // its own list only above the list's fourth (a tie keeps the earlier,
      // smaller id) and ranks after the entries it ties: the branch-free
      // insertion top4_insert_after.  A list whose fourth rose publishes it.
      // Classes past n_cls exist only in the last tile.
      const bool tail = m0 + 128 > n_cls;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int ray = rr + 32 * j + l32;
        const float th = fmaxf(tv[j][3], from_ordered(thr[ray]));
        float best = acc[j][0];
#pragma unroll
        for (int r = 1; r < 16; ++r) best = fmaxf(best, acc[j][r]);
        if (best >= th) {
          const float before = tv[j][3];
          if (!tail) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
              const float v = acc[j][r];
              if (v >= th && v > tv[j][3]) top4_insert_after(tv[j], ti[j], v, m0 + cr + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3));
            }
          } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
              const int cls = m0 + cr + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
              const float v = acc[j][r];
              if (cls < n_cls && v >= th && v > tv[j][3]) top4_insert_after(tv[j], ti[j], v, cls);
            }
          }
          if (tv[j][3] > before) atomicMax(&thr[ray], ordered_bits(tv[j][3]));
        }
      }
    }
  };
  bf16x8 a0[2], a1[2], a2[2], a3[2];
  w4(0, a0); w4(1, a1); w4(2, a2); w4(3, a3);
#pragma unroll 1
  for (int step = 0; step < nsteps; step += 4) {
    tile_step(step, a0);
    tile_step(step + 1, a1);
    if (step + 4 < nsteps) { w4(step + 4, a0); w4(step + 5, a1); }
    tile_step(step + 2, a2);
    tile_step(step + 3, a3);
    if (step + 6 < nsteps) { w4(step + 6, a2); w4(step + 7, a3); }
  }
  __syncthreads();   // act is free: merge buffer
  constexpr int NL = 8 * 4;                                   // candidates per ray: 8 lists of 4
  float* cv = reinterpret_cast<float*>(&act[0][0]);          // [IR][NL]
  int* ci = reinterpret_cast<int*>(cv + IR * NL);             // [IR][NL]
  const int list = (wave >> 1) * 2 + (lane >> 5);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int ray = rr + 32 * j + l32;
#pragma unroll
    for (int k = 0; k < 4; ++k) { cv[ray * NL + list * 4 + k] = tv[j][k]; ci[ray * NL + list * 4 + k] = ti[j][k]; }
  }
  __syncthreads();
  if (tid < IR) {
    float bv[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    int bc[4] = {0x7fffffff, 0x7fffffff, 0x7fffffff, 0x7fffffff};
    for (int k = 0; k < NL; ++k) {
      const float v = cv[tid * NL + k];
      const int c = ci[tid * NL + k];
      if (c < n_cls && beats(v, c, bv[3], bc[3])) top4_insert(bv, bc, v, c);
    }
    const uint64_t m = ray0 + tid;
    if (m < n)
#pragma unroll
      for (int k = 0; k < 4; ++k) top[m * 4 + k] = bc[k] < n_cls ? (uint16_t)bc[k] : (uint16_t)0xFFFF;
  }
}

// Labels: subtree of the hit leaf (-1 on a miss or an id outside the scene).
__global__ void k_labels(const int32_t* __restrict__ leaf, uint64_t n, const int32_t* __restrict__ sub_of_leaf, uint64_t n_leaves,
                         int32_t* label) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t l = leaf[i];
  label[i] = l >= 0 && (uint64_t)l < n_leaves ? sub_of_leaf[l] : -1;
}

// Proposal-seeded exact nearest walk (triangles, D = 3): the proposed subtrees
// are walked first, each from its own ref, with the shrinking limit; then the
// full walk runs from the root with that limit and best hit.  The triangles
// are the guarded kind (SQ_HITS_GUARDED): every accepted hit lies inside its
// own box, so the answer does not depend on which subtrees go first.  top == NULL
// walks from the root at once (the plain walk).  Same pooled structure and
// pausing of descents as the plain nearest kernel.  COUNT accumulates into stats[0..2]: rays whose
// proposals held a hit, pair nodes fetched, primitives tested.
template <bool COUNT>
__global__ __launch_bounds__(256, COUNT ? 4 : 8) void k_nearest_proposed(SceneView s, const float* __restrict__ q, uint32_t n, float max_toi,
                                                             const uint16_t* __restrict__ top, const uint32_t* __restrict__ sub_ref,
                                                             uint32_t n_sub, int32_t* out_leaf, float* out_toi, uint32_t per_lane,
                                                             uint32_t S, SpillEntry* spill, uint32_t spill_stride,
                                                             unsigned long long* stats, uint32_t pause_below) {
  extern __shared__ uint32_t lds[];
  constexpr int D = 3, PS = 12;
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t gtid = blockIdx.x * B + tid;
  const uint32_t lane = tid & 63u;
  const uint64_t pool = 64ull * per_lane;
  uint64_t next = ((uint64_t)blockIdx.x * (B >> 6) + (tid >> 6)) * pool;
  const uint64_t end = next + pool < (uint64_t)n ? next + pool : (uint64_t)n;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill + gtid, spill_stride};
  const Filter nof{nullptr, ~0ull, -1};
  const uint32_t root_ref = fbits(s.root[2 * D]);
  RayPolicy<D, KIND_TRI, true> pol;
  WalkState w;
  WorkCount wc;
  uint32_t rid = NO_LEAF;
  int phase = 0;          // 0..3: a proposed subtree, 4: the full walk from the root
  bool exhausted = next >= end;
  uint32_t seeded = 0;
  // Refs of the subtrees this query has walked; the root ref (never a child)
  // fills the empty slots.  Later walks skip them: the limit only shrinks, so
  // a walked subtree holds no hit the query has not already seen.
  uint32_t done[4];
  // Start the first valid proposal at or after phase ph (ids >= n_sub mark
  // empty slots), else the full walk.
  auto start = [&](int ph) {
    w.sp = 0;
    if (ph == 0)
#pragma unroll
      for (int k = 0; k < 4; ++k) done[k] = root_ref;
    if (top)
      for (; ph < 4; ++ph) {
        const uint32_t c = top[(size_t)rid * 4 + ph];
        if (c < n_sub) {
          phase = ph;
          w.cur = sub_ref[c];
#pragma unroll
          for (int k = 0; k < 4; ++k) done[k] = k == ph ? w.cur : done[k];
          return;
        }
      }
    phase = 4;
    if (w.res.leaf != NO_LEAF) ++seeded;
    w.cur = root_ref;
  };
  while (true) {
    const bool need = rid == NO_LEAF && !exhausted;
    const uint64_t m = __ballot(need);
    if (m) {
      if (need) {
        const uint64_t my = next + __popcll(m & lt_mask);
        if (my < end) {
          rid = (uint32_t)my;
          pol.setup(q + (size_t)rid * 6);
          if (!walk_begin<D>(s, pol, max_toi, w)) {
            out_leaf[rid] = -1; out_toi[rid] = SQ_INF; rid = NO_LEAF;
          } else {
            start(0);
          }
        } else {
          exhausted = true;
        }
      }
      next += __popcll(m);
    }
    if (__ballot(!(exhausted && rid == NO_LEAF)) == 0) break;
    const uint32_t active = (uint32_t)__popcll(__ballot(rid != NO_LEAF));
    if (rid == NO_LEAF) continue;
    if (!walk_step<D, PS, false, false, COUNT, true>(s, pol, nof, st, w, &wc, done, WavePause{pause_below, active})) continue;
    if (phase < 4) { start(phase + 1); continue; }   // this subtree is exhausted
    out_leaf[rid] = w.res.leaf != NO_LEAF ? (int32_t)w.res.leaf : -1;
    out_toi[rid] = w.res.t;
    rid = NO_LEAF;
  }
  if constexpr (COUNT) {
    unsigned long long v[3] = {seeded, wc.nodes, wc.prims};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
      if (lane == 0 && v[k]) atomicAdd(stats + k, v[k]);
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------

struct Layer {
  int in, out, out_pad;            // out_pad: out rounded up to 128
  DevBuf w, m, v, g, gb, b, mb, vb;  // fp32 master [out_pad][in], Adam moments, grads; bias (+moments)
  DevBuf w_nk, w_kn;               // bf16 copies: [out_pad][in] and [in][out_pad]
};

struct sq_proposal {
  sq_ctx* ctx = nullptr;
  sq_bvh* scene = nullptr;         // the tree cut into subtrees (a Tlas's top tree); not owned
  sq_tlas* tlas = nullptr;         // the two-level scene, when the proposal serves one; not owned
  uint32_t n_sub = 0;
  int cut = 0;
  Layer L[4];
  DevBuf rff, box, sub_ref, sub_of_leaf;
  DevBuf w4f;                      // the last layer's weights in fragment order (inference)
  DevBuf wf[3];                    // the hidden layers' weights in fragment order (inference)
  // training temporaries (grow with the batch)
  uint64_t cap = 0;
  DevBuf F, H[3], Z, dZ, dZt, HT, dH, dHb, loss, lab, cnt, part, bpart, tmpT;
  int step = 0;
};

namespace {

int gemm(sq_ctx* c, int epi, const bf16* A, int lda, const bf16* B, int ldb, int M, int N, int K, float* Cf, bf16* Cb, int ldc,
         const float* bias, const bf16* mask, int ldm, int splits = 1, size_t split_stride = 0) {
  if (M % BM || N % BN || K % BK) return fail(SQ_ERR_INVALID, "gemm shape must be multiples of 128x128x32");
  const int k_chunk = ((K / splits + BK - 1) / BK) * BK;
  dim3 grid(N / BN, M / BM, splits);
#define SQ_GEMM(E) k_gemm_tn<E><<<grid, 256, 0, c->stream>>>(A, lda, B, ldb, K, k_chunk, Cf, Cb, ldc, split_stride, bias, mask, ldm)
  switch (epi) {
    case EPI_F32: SQ_GEMM(EPI_F32); break;
    case EPI_BIAS_RELU_BF16: SQ_GEMM(EPI_BIAS_RELU_BF16); break;
    case EPI_BIAS_F32: SQ_GEMM(EPI_BIAS_F32); break;
    case EPI_MASK_BF16: SQ_GEMM(EPI_MASK_BF16); break;
    default: SQ_GEMM(EPI_PARTIAL); break;
  }
#undef SQ_GEMM
  HIP_TRY(hipGetLastError());
  return SQ_OK;
}

int transpose(sq_ctx* c, const bf16* in, int rows, int cols, bf16* out) {
  dim3 grid(cols / 64, rows / 64), block(64, 4);
  k_transpose<<<grid, block, 0, c->stream>>>(in, rows, cols, out);
  HIP_TRY(hipGetLastError());
  return SQ_OK;
}

inline uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }

// Weight gradient dW[out_pad][in] = dY^T X, reduced over the batch with split-K.
int weight_grad(sq_proposal* p, const bf16* dYt, const bf16* Xt, uint64_t M, int out_pad, int in, float* g) {
  sq_ctx* c = p->ctx;
  const int tiles = (out_pad / BM) * (in / BN);
  int splits = std::max(1, std::min<int>(512 / std::max(1, tiles), (int)(M / 1024)));
  const size_t nout = (size_t)out_pad * in;
  int rc = p->part.ensure((size_t)splits * nout * 4);
  if (rc) return rc;
  if ((rc = gemm(c, EPI_PARTIAL, dYt, (int)M, Xt, (int)M, out_pad, in, (int)M, p->part.as<float>(), nullptr, in, nullptr, nullptr, 0,
                 splits, nout))) return rc;
  k_sum_splits<<<grid_for(nout, 256), 256, 0, c->stream>>>(p->part.as<float>(), splits, nout, g);
  HIP_TRY(hipGetLastError());
  return SQ_OK;
}

// Bias gradient: column sums of dY [M][cols].
int bias_grad(sq_proposal* p, const bf16* dY, uint64_t M, int cols, float* g) {
  sq_ctx* c = p->ctx;
  const uint64_t chunk = 512;
  const uint32_t nch = (uint32_t)((M + chunk - 1) / chunk);
  int rc = p->bpart.ensure((size_t)nch * cols * 4);
  if (rc) return rc;
  k_colsum<<<dim3(cols / 64, nch), 256, 0, c->stream>>>(dY, M, cols, chunk, p->bpart.as<float>());
  k_sum_splits<<<grid_for(cols, 256), 256, 0, c->stream>>>(p->bpart.as<float>(), (int)nch, (size_t)cols, g);
  HIP_TRY(hipGetLastError());
  return SQ_OK;
}

MlpWeights weights_of(const sq_proposal* p) {
  MlpWeights mw;
  for (int l = 0; l < 4; ++l) { mw.w[l] = p->L[l].w_nk.as<bf16>(); mw.b[l] = p->L[l].b.as<float>(); }
  mw.w4f = p->w4f.as<bf16>();
  for (int l = 0; l < 3; ++l) mw.wf[l] = p->wf[l].as<bf16>();
  mw.rff = p->rff.as<float>();
  mw.box = p->box.as<float>();
  return mw;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// A proposal over the subtrees of `scene` cut at `cut_depth` (the tree of a
// TRIANGLE sq_bvh, or a Tlas's top tree over its instances, tlas != NULL).
int proposal_create(sq_ctx* ctx, sq_bvh* scene, sq_tlas* tlas, int cut_depth, float rff_scale, uint64_t seed, sq_proposal** out) {
  if (scene->n_prims == 0) return fail(SQ_ERR_INVALID, "empty scene");
  if (cut_depth < 1 || cut_depth > 12) return fail(SQ_ERR_INVALID, "cut depth must be in [1, 12] (<= 4096 subtrees)");
  if (!(rff_scale > 0.0f && rff_scale < 1e3f)) return fail(SQ_ERR_INVALID, "rff_scale must be in (0, 1000)");
  int rh = ensure_host_tree(scene);
  if (rh) return rh;
  auto p = std::make_unique<sq_proposal>();
  p->ctx = ctx; p->scene = scene; p->tlas = tlas; p->cut = cut_depth;
  // Subtrees of the reference tree cut at `cut_depth`: nodes at that depth,
  // and leaves above it.  Their device refs and the subtree of every leaf.
  const sqh::RefTree& t = scene->tree;
  const size_t nn = t.node_count();
  std::vector<int> depth(nn, 0);
  std::vector<int32_t> sub_of_node(nn, -1);
  std::vector<uint32_t> refs;
  const int NF = (3 + 1) * 4;
  const auto& P = scene->packed;
  auto ref_of = [&](size_t node) -> uint32_t {
    const uint32_t slot = P.slot_of_ref[node];
    if (slot == 0xFFFFFFFFu) { uint32_t r; std::memcpy(&r, &P.root[6], 4); return r; }
    uint32_t r; std::memcpy(&r, &P.nodes[(size_t)(slot >> 1) * NF + 12 + (slot & 1)], 4);
    return r;
  };
  for (size_t i = 0; i < nn; ++i) {           // pre-order: parents first
    if (i == 0) depth[i] = 0;
    if (t.prim_count[i] == 0) { depth[t.child_a[i]] = depth[i] + 1; depth[t.child_b[i]] = depth[i] + 1; }
    if (sub_of_node[i] < 0 && (depth[i] == cut_depth || (t.prim_count[i] > 0 && depth[i] < cut_depth))) {
      sub_of_node[i] = (int32_t)refs.size();
      refs.push_back(ref_of(i));
    }
    if (sub_of_node[i] >= 0 && t.prim_count[i] == 0) {
      sub_of_node[t.child_a[i]] = sub_of_node[i];
      sub_of_node[t.child_b[i]] = sub_of_node[i];
    }
  }
  std::vector<int32_t> sol(scene->n_prims, -1);
  for (size_t i = 0; i < nn; ++i)
    if (t.prim_count[i] > 0)
      for (uint32_t k = t.child_a[i]; k < t.child_a[i] + t.prim_count[i]; ++k) sol[t.prim_indices[k]] = sub_of_node[i];
  p->n_sub = (uint32_t)refs.size();
  int rc;
  hipStream_t s = ctx->stream;
  if ((rc = upload(p->sub_ref, refs.data(), refs.size() * 4, s))) return rc;
  if ((rc = upload(p->sub_of_leaf, sol.data(), sol.size() * 4, s))) return rc;
  // scene box for the input normalisation
  if ((rc = upload(p->box, t.bounds.data(), 6 * 4, s))) return rc;
  // RFF matrix and weights from a SplitMix64 stream: Box-Muller normals
  uint64_t st = seed;
  auto next = [&]() { st += 0x9E3779B97F4A7C15ull; uint64_t z = st; z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; return z ^ (z >> 31); };
  auto unif = [&]() { return ((double)(next() >> 11) + 0.5) / 9007199254740992.0; };
  auto normal = [&]() { return std::sqrt(-2.0 * std::log(unif())) * std::cos(6.283185307179586 * unif()); };
  std::vector<float> B(6 * N_RFF);
  for (auto& x : B) x = (float)(rff_scale * normal());
  if ((rc = upload(p->rff, B.data(), B.size() * 4, s))) return rc;
  const int ins[4] = {F_IN, HID, HID, HID};
  const int outs[4] = {HID, HID, HID, (int)p->n_sub};
  for (int l = 0; l < 4; ++l) {
    Layer& Ly = p->L[l];
    Ly.in = ins[l]; Ly.out = outs[l]; Ly.out_pad = (int)round_up(outs[l], 128);
    const size_t nw = (size_t)Ly.out_pad * Ly.in;
    std::vector<float> w(nw, 0.0f), zero(nw, 0.0f), b(Ly.out_pad, 0.0f), zb(Ly.out_pad, 0.0f);
    const double sd = std::sqrt(2.0 / Ly.in);   // He initialisation
    for (int o = 0; o < Ly.out; ++o)
      for (int i = 0; i < Ly.in; ++i) w[(size_t)o * Ly.in + i] = (float)(sd * normal());
    std::vector<bf16> wnk(nw), wkn(nw);
    for (size_t i = 0; i < nw; ++i) { wnk[i] = (bf16)w[i]; }
    for (int o = 0; o < Ly.out_pad; ++o)
      for (int i = 0; i < Ly.in; ++i) wkn[(size_t)i * Ly.out_pad + o] = wnk[(size_t)o * Ly.in + i];
    if ((rc = upload(Ly.w, w.data(), nw * 4, s))) return rc;
    if ((rc = upload(Ly.m, zero.data(), nw * 4, s))) return rc;
    if ((rc = upload(Ly.v, zero.data(), nw * 4, s))) return rc;
    if ((rc = Ly.g.ensure(nw * 4))) return rc;
    if ((rc = upload(Ly.b, b.data(), b.size() * 4, s))) return rc;
    if ((rc = upload(Ly.mb, zb.data(), zb.size() * 4, s))) return rc;
    if ((rc = upload(Ly.vb, zb.data(), zb.size() * 4, s))) return rc;
    if ((rc = Ly.gb.ensure(zb.size() * 4))) return rc;
    if ((rc = upload(Ly.w_nk, wnk.data(), nw * 2, s))) return rc;
    if ((rc = upload(Ly.w_kn, wkn.data(), nw * 2, s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));   // the host vectors end with this iteration
  }
  HIP_TRY(hipStreamSynchronize(s));
  ctx_retain(ctx);
  *out = p.release();
  return SQ_OK;
}

}  // namespace

extern "C" {

int sq_proposal_create(sq_ctx* ctx, sq_bvh* scene, int cut_depth, float rff_scale, uint64_t seed, sq_proposal** out) {
  if (!ctx || !scene || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (scene->kind != KIND_TRI || scene->dim != 3) return fail(SQ_ERR_UNSUPPORTED, "the proposal serves TRIANGLE scenes (D = 3)");
  return proposal_create(ctx, scene, nullptr, cut_depth, rff_scale, seed, out);
}

int sq_proposal_create_tlas(sq_ctx* ctx, sq_tlas* scene, int cut_depth, float rff_scale, uint64_t seed, sq_proposal** out) {
  if (!ctx || !scene || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (tlas_dim(scene) != 3) return fail(SQ_ERR_UNSUPPORTED, "the proposal serves D = 3 scenes");
  return proposal_create(ctx, tlas_top_tree(scene), scene, cut_depth, rff_scale, seed, out);
}

int sq_proposal_destroy(sq_proposal* p) {
  if (!p) return SQ_OK;
  sq_ctx* c = p->ctx;
  (void)hipStreamSynchronize(c->stream);
  delete p;
  ctx_release(c);
  return SQ_OK;
}

int sq_proposal_num_subtrees(const sq_proposal* p, uint32_t* out) {
  if (!p || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = p->n_sub;
  return SQ_OK;
}

int sq_proposal_labels(sq_proposal* p, const int32_t* leaf, uint64_t n, int32_t* label) {
  if (!p || (n && (!leaf || !label))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (!n) return SQ_OK;
  k_labels<<<grid_for(n, 256), 256, 0, p->ctx->stream>>>(leaf, n, p->sub_of_leaf.as<int32_t>(), p->scene->n_prims, label);
  HIP_TRY(hipGetLastError());
  return SQ_OK;
}

// One Adam step on a batch (rays [n][6], labels [n]; label < 0 rows are
// ignored).  When grads_out != NULL the step stops after the backward pass (no
// update) and copies the gradients [dW1, db1, ..., dW4, db4] (fp32, unpadded
// [out][in] and [out]) there: the gradient check.  loss_out, when given,
// receives the mean loss over labelled rows (synchronises).
int sq_proposal_train_step(sq_proposal* p, const float* rays, const int32_t* labels, uint64_t n, float lr, float* loss_out,
                           float* grads_out) {
  if (!p || !rays || !labels || !n) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n > (1ull << 24)) return fail(SQ_ERR_INVALID, "at most 2^24 rays per training batch");
  sq_ctx* c = p->ctx;
  const uint64_t M = round_up(n, 128);
  const int K4 = p->L[3].out_pad;
  int rc;
  if (M > p->cap) {
    if ((rc = p->F.ensure(M * F_IN * 2))) return rc;
    for (int i = 0; i < 3; ++i) if ((rc = p->H[i].ensure(M * HID * 2))) return rc;
    if ((rc = p->Z.ensure(M * K4 * 4))) return rc;
    if ((rc = p->dZ.ensure(M * K4 * 2))) return rc;
    if ((rc = p->dZt.ensure(M * K4 * 2))) return rc;
    if ((rc = p->HT.ensure(M * HID * 2))) return rc;
    if ((rc = p->dH.ensure(M * HID * 2))) return rc;
    if ((rc = p->dHb.ensure(M * HID * 2))) return rc;
    if ((rc = p->tmpT.ensure(M * HID * 2))) return rc;
    if ((rc = p->loss.ensure(M * 4))) return rc;
    if ((rc = p->lab.ensure(M * 4))) return rc;
    p->cap = M;
  }
  if ((rc = p->cnt.ensure(64))) return rc;
  hipStream_t s = c->stream;
  // labels padded with -1, and their count
  int32_t* lab = p->lab.as<int32_t>();
  uint32_t* cnt = p->cnt.as<uint32_t>();
  k_fill_u32<<<grid_for(M, 256), 256, 0, s>>>(reinterpret_cast<uint32_t*>(lab), M, 0xFFFFFFFFu);
  k_fill_u32<<<1, 64, 0, s>>>(cnt, 1, 0u);
  HIP_TRY(hipMemcpyAsync(lab, labels, n * 4, hipMemcpyDeviceToDevice, s));
  k_count_labels<<<grid_for(n, 256), 256, 0, s>>>(lab, n, p->L[3].out, cnt);
  HIP_TRY(hipGetLastError());
  bf16* F = p->F.as<bf16>();
  bf16* H[3] = {p->H[0].as<bf16>(), p->H[1].as<bf16>(), p->H[2].as<bf16>()};
  // forward
  k_features<<<grid_for(M, 256), 256, 0, s>>>(rays, n, M, p->rff.as<float>(), p->box.as<float>(), F);
  HIP_TRY(hipGetLastError());
  const bf16* X = F;
  int in = F_IN;
  for (int l = 0; l < 3; ++l) {
    if ((rc = gemm(c, EPI_BIAS_RELU_BF16, X, in, p->L[l].w_nk.as<bf16>(), in, (int)M, HID, in, nullptr, H[l], HID, p->L[l].b.as<float>(),
                   nullptr, 0))) return rc;
    X = H[l];
    in = HID;
  }
  if ((rc = gemm(c, EPI_BIAS_F32, H[2], HID, p->L[3].w_nk.as<bf16>(), HID, (int)M, K4, HID, p->Z.as<float>(), nullptr, K4,
                 p->L[3].b.as<float>(), nullptr, 0))) return rc;
  // loss and dZ = (softmax - onehot) / count
  k_softmax_ce<<<(uint32_t)((M + 3) / 4), 256, 0, s>>>(p->Z.as<float>(), K4, p->L[3].out, lab, M, cnt, p->dZ.as<bf16>(), p->loss.as<float>());
  HIP_TRY(hipGetLastError());
  // backward, last layer
  bf16* dZ = p->dZ.as<bf16>();
  if ((rc = transpose(c, dZ, (int)M, K4, p->dZt.as<bf16>()))) return rc;
  if ((rc = transpose(c, H[2], (int)M, HID, p->HT.as<bf16>()))) return rc;
  if ((rc = weight_grad(p, p->dZt.as<bf16>(), p->HT.as<bf16>(), M, K4, HID, p->L[3].g.as<float>()))) return rc;
  if ((rc = bias_grad(p, dZ, M, K4, p->L[3].gb.as<float>()))) return rc;
  // dH3 = (dZ W4) masked by the ReLU of H3
  if ((rc = gemm(c, EPI_MASK_BF16, dZ, K4, p->L[3].w_kn.as<bf16>(), K4, (int)M, HID, K4, nullptr, p->dH.as<bf16>(), HID, nullptr, H[2],
                 HID))) return rc;
  bf16* dcur = p->dH.as<bf16>();
  bf16* dnext = p->dHb.as<bf16>();
  for (int l = 2; l >= 0; --l) {
    const bf16* Xin = l > 0 ? H[l - 1] : F;
    const int xin = l > 0 ? HID : F_IN;
    if ((rc = transpose(c, dcur, (int)M, HID, p->tmpT.as<bf16>()))) return rc;   // dY^T [256][M]
    if ((rc = transpose(c, Xin, (int)M, xin, p->HT.as<bf16>()))) return rc;      // X^T [xin][M]
    if ((rc = weight_grad(p, p->tmpT.as<bf16>(), p->HT.as<bf16>(), M, HID, xin, p->L[l].g.as<float>()))) return rc;
    if ((rc = bias_grad(p, dcur, M, HID, p->L[l].gb.as<float>()))) return rc;
    if (l > 0) {
      if ((rc = gemm(c, EPI_MASK_BF16, dcur, HID, p->L[l].w_kn.as<bf16>(), HID, (int)M, HID, HID, nullptr, dnext, HID, nullptr, H[l - 1],
                     HID))) return rc;
      std::swap(dcur, dnext);
    }
  }
  if (grads_out) {
    HIP_TRY(hipStreamSynchronize(s));
    float* o = grads_out;
    for (int l = 0; l < 4; ++l) {
      const Layer& Ly = p->L[l];
      HIP_TRY(hipMemcpy(o, Ly.g.p, (size_t)Ly.out * Ly.in * 4, hipMemcpyDeviceToHost));
      o += (size_t)Ly.out * Ly.in;
      HIP_TRY(hipMemcpy(o, Ly.gb.p, (size_t)Ly.out * 4, hipMemcpyDeviceToHost));
      o += Ly.out;
    }
  } else {
    // Adam (beta1 0.9, beta2 0.999, eps 1e-8), bias-corrected
    ++p->step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)p->step), c2 = 1.0f - std::pow(b2, (float)p->step);
    for (int l = 0; l < 4; ++l) {
      Layer& Ly = p->L[l];
      const size_t nw = (size_t)Ly.out_pad * Ly.in;
      k_adam<<<grid_for(nw, 256), 256, 0, s>>>(Ly.w.as<float>(), Ly.m.as<float>(), Ly.v.as<float>(), Ly.g.as<float>(), nw, lr, b1, b2, 1e-8f,
                                               c1, c2, Ly.w_nk.as<bf16>(), Ly.w_kn.as<bf16>(), Ly.out_pad, Ly.in);
      k_adam<<<grid_for(Ly.out_pad, 256), 256, 0, s>>>(Ly.b.as<float>(), Ly.mb.as<float>(), Ly.vb.as<float>(), Ly.gb.as<float>(), Ly.out_pad,
                                                       lr, b1, b2, 1e-8f, c1, c2, nullptr, nullptr, 0, 1);
      HIP_TRY(hipGetLastError());
    }
  }
  if (loss_out) {
    std::vector<float> hl(n);
    uint32_t hc = 0;
    HIP_TRY(hipMemcpyAsync(hl.data(), p->loss.p, n * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&hc, cnt, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double sum = 0.0;
    for (float v : hl) sum += v;
    *loss_out = hc ? (float)(sum / (double)hc) : 0.0f;
  }
  return SQ_OK;
}

// Weights (fp32 masters, unpadded): [W1 [out][in], b1, ..., W4, b4]; get or set.
int sq_proposal_weights(sq_proposal* p, float* get, const float* set) {
  if (!p) return fail(SQ_ERR_INVALID, "NULL argument");
  sq_ctx* c = p->ctx;
  HIP_TRY(hipStreamSynchronize(c->stream));
  size_t off = 0;
  for (int l = 0; l < 4; ++l) {
    Layer& Ly = p->L[l];
    const size_t nw = (size_t)Ly.out * Ly.in;
    if (get) {
      HIP_TRY(hipMemcpy(get + off, Ly.w.p, nw * 4, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(get + off + nw, Ly.b.p, (size_t)Ly.out * 4, hipMemcpyDeviceToHost));
    }
    if (set) {
      HIP_TRY(hipMemcpy(Ly.w.p, set + off, nw * 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(Ly.b.p, set + off + nw, (size_t)Ly.out * 4, hipMemcpyHostToDevice));
      std::vector<float> w(nw);
      std::memcpy(w.data(), set + off, nw * 4);
      std::vector<bf16> wnk((size_t)Ly.out_pad * Ly.in, (bf16)0.0f), wkn((size_t)Ly.out_pad * Ly.in, (bf16)0.0f);
      for (int o = 0; o < Ly.out; ++o)
        for (int i = 0; i < Ly.in; ++i) {
          wnk[(size_t)o * Ly.in + i] = (bf16)w[(size_t)o * Ly.in + i];
          wkn[(size_t)i * Ly.out_pad + o] = wnk[(size_t)o * Ly.in + i];
        }
      HIP_TRY(hipMemcpy(Ly.w_nk.p, wnk.data(), wnk.size() * 2, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(Ly.w_kn.p, wkn.data(), wkn.size() * 2, hipMemcpyHostToDevice));
    }
    off += nw + Ly.out;
  }
  return SQ_OK;
}

int sq_proposal_num_params(const sq_proposal* p, uint64_t* out) {
  if (!p || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  uint64_t n = 0;
  for (int l = 0; l < 4; ++l) n += (uint64_t)p->L[l].out * p->L[l].in + p->L[l].out;
  *out = n;
  return SQ_OK;
}

// The RFF matrix B [6][64] (fp32) and the scene box used for normalisation.
int sq_proposal_features_spec(const sq_proposal* p, float* rff, float* box) {
  if (!p) return fail(SQ_ERR_INVALID, "NULL argument");
  HIP_TRY(hipStreamSynchronize(p->ctx->stream));
  if (rff) HIP_TRY(hipMemcpy(rff, p->rff.p, 6 * N_RFF * 4, hipMemcpyDeviceToHost));
  if (box) HIP_TRY(hipMemcpy(box, p->box.p, 6 * 4, hipMemcpyDeviceToHost));
  return SQ_OK;
}

// Top-4 proposed subtrees per ray (uint16 [n][4], best first; 0xFFFF marks
// an empty slot when the scene has fewer than 4 subtrees): one fused launch.
int sq_proposal_infer(sq_proposal* p, const float* rays, uint64_t n, uint16_t* top) {
  if (!p || (n && (!rays || !top))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (!n) return SQ_OK;
  sq_ctx* c = p->ctx;
  const uint64_t blocks = (n + IR - 1) / IR;
  if (blocks > 0x7FFFFFFFull) return fail(SQ_ERR_INVALID, "too many rays");
  int rc = begin_timing(c);
  if (rc) return rc;
  const uint64_t pieces = (uint64_t)p->L[3].out_pad * HID / 8;
  if ((rc = p->w4f.ensure(pieces * 16))) return rc;
  k_w4_fragments<<<grid_for(pieces, 256), 256, 0, c->stream>>>(p->L[3].w_nk.as<bf16>(), (uint32_t)p->L[3].out_pad, p->w4f.as<bf16>());
  HIP_TRY(hipGetLastError());
  for (int l = 0; l < 3; ++l) {
    const uint32_t kin = (uint32_t)p->L[l].in, hp = HID * kin / 8;
    if ((rc = p->wf[l].ensure((size_t)hp * 16))) return rc;
    k_hidden_fragments<<<grid_for(hp, 256), 256, 0, c->stream>>>(p->L[l].w_nk.as<bf16>(), kin, p->wf[l].as<bf16>());
    HIP_TRY(hipGetLastError());
  }
  k_infer_top4<<<(uint32_t)blocks, IT, 0, c->stream>>>(rays, n, weights_of(p), p->L[3].out, p->L[3].out_pad, top);
  HIP_TRY(hipGetLastError());
  return end_timing(c);
}

// Exact nearest hit seeded by the proposal (top from sq_proposal_infer, or
// NULL for the plain walk).  Bit-identical to sq_raycast_nearest.  stats
// (device u64[3], may be NULL) accumulates: rays whose proposals held a hit,
// pair nodes fetched, primitives tested.
int sq_raycast_nearest_proposed(sq_proposal* p, const float* rays, uint64_t n, float max_toi, const uint16_t* top, int32_t* out_leaf,
                                float* out_toi, uint64_t* stats) {
  if (!p || (n && (!rays || !out_leaf || !out_toi))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (p->tlas) return fail(SQ_ERR_INVALID, "the proposal serves a Tlas: use sq_tlas_raycast_nearest_proposed");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 rays per call");
  if (!n) return SQ_OK;
  sq_ctx* c = p->ctx;
  sq_bvh* b = p->scene;
  const int B = std::min(c->block, 256);
  const uint32_t S = (uint32_t)c->lds_stack;
  const uint32_t per = pooled_per_lane(c, n);
  const uint64_t pool = 64ull * per;
  const uint64_t waves = (n + pool - 1) / pool;
  const uint64_t blocks = (waves + (B / 64) - 1) / (B / 64);
  const uint64_t lanes = blocks * (uint64_t)B;
  const uint32_t stack_n = b->packed.max_stack;
  const uint32_t rows = stack_n > S ? stack_n - S : 0;
  int rc;
  if ((rc = c->spill.ensure((size_t)rows * lanes * sizeof(SpillEntry) + 64))) return rc;
  if ((rc = begin_timing(c))) return rc;
  auto* st = reinterpret_cast<unsigned long long*>(stats);
  if (stats)
    k_nearest_proposed<true><<<(uint32_t)blocks, B, (size_t)S * B * 8, c->stream>>>(view_of(b), rays, (uint32_t)n, max_toi, top,
        p->sub_ref.as<uint32_t>(), p->n_sub, out_leaf, out_toi, per, S, c->spill.as<SpillEntry>(), (uint32_t)lanes, st,
        (uint32_t)c->descend_min);
  else
    k_nearest_proposed<false><<<(uint32_t)blocks, B, (size_t)S * B * 8, c->stream>>>(view_of(b), rays, (uint32_t)n, max_toi, top,
        p->sub_ref.as<uint32_t>(), p->n_sub, out_leaf, out_toi, per, S, c->spill.as<SpillEntry>(), (uint32_t)lanes, st,
        (uint32_t)c->descend_min);
  HIP_TRY(hipGetLastError());
  return end_timing(c);
}

// Exact nearest hit over a Tlas seeded by its proposal (the proposed
// top-tree subtrees first): the guarded two-level walk's answer.
int sq_tlas_raycast_nearest_proposed(sq_proposal* p, const float* rays, uint64_t n, float max_toi, const uint16_t* top,
                                     int32_t* out_inst, float* out_toi, int32_t* out_sub, uint64_t* stats) {
  if (!p || (n && (!rays || !out_inst || !out_toi))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (!p->tlas) return fail(SQ_ERR_INVALID, "the proposal serves a single-level tree: use sq_raycast_nearest_proposed");
  return tlas_nearest_proposed(p->tlas, rays, n, max_toi, top, p->sub_ref.as<uint32_t>(), p->n_sub, out_inst, out_toi, out_sub,
                               stats);
}

// Test hook for the GEMM: C[M][N] (fp32) = A[M][K] * B[N][K]^T, bf16 inputs
// given as fp32 (rounded to bf16 here), device pointers.  Shapes multiples of
// 128 x 128 x 32; splits > 1 exercises the split-K reduction.
int sq_debug_gemm(sq_ctx* c, const float* A, const float* B, int M, int N, int K, int splits, float* C) {
  if (!c || !A || !B || !C) return fail(SQ_ERR_INVALID, "NULL argument");
  DevBuf a, b, part;
  int rc;
  if ((rc = a.ensure((size_t)M * K * 2)) || (rc = b.ensure((size_t)N * K * 2))) return rc;
  std::vector<float> ha((size_t)M * K), hb((size_t)N * K);
  HIP_TRY(hipMemcpy(ha.data(), A, ha.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hb.data(), B, hb.size() * 4, hipMemcpyDeviceToHost));
  std::vector<bf16> ba(ha.size()), bb(hb.size());
  for (size_t i = 0; i < ha.size(); ++i) ba[i] = (bf16)ha[i];
  for (size_t i = 0; i < hb.size(); ++i) bb[i] = (bf16)hb[i];
  HIP_TRY(hipMemcpy(a.p, ba.data(), ba.size() * 2, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.p, bb.data(), bb.size() * 2, hipMemcpyHostToDevice));
  if (splits <= 1) return gemm(c, EPI_F32, a.as<bf16>(), K, b.as<bf16>(), K, M, N, K, C, nullptr, N, nullptr, nullptr, 0);
  const size_t nout = (size_t)M * N;
  if ((rc = part.ensure((size_t)splits * nout * 4))) return rc;
  if ((rc = gemm(c, EPI_PARTIAL, a.as<bf16>(), K, b.as<bf16>(), K, M, N, K, part.as<float>(), nullptr, N, nullptr, nullptr, 0, splits, nout))) return rc;
  k_sum_splits<<<grid_for(nout, 256), 256, 0, c->stream>>>(part.as<float>(), splits, nout, C);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return SQ_OK;
}

}  // extern "C"
