// MLP surrogate of a feature map: trained and evaluated with hand-written
// BF16 matrix-core (MFMA) kernels accumulating in FP32, for gfx950.
//
// No counterpart in the reference.  The network maps 2 normalised parameters
// to feature predictions:
//   encode   x (2) -> 16 fixed Fourier features, sin and cos    -> 32
//   layer 1  32  -> 256, SiLU
//   layer 2  256 -> 256, SiLU
//   layer 3  256 -> 256, SiLU
//   layer 4  256 -> 256, SiLU
//   heads    256 -> 4 heads x (log-period, amplitude, 16 class logits) = 72,
//            padded to 128 output columns (the padding stays exactly zero)
// Weights are kept transposed, Wt[out][in], as FP32 master copies (Adam) and
// BF16 copies (MFMA operands).  Activations are BF16, pre-activations FP32.
//
// The 4 heads share the trunk and are trained on different bootstrap subsets
// of every minibatch (a sample is seen by a head unless a hash of
// (sample, head) is 0 mod 4), so their disagreement measures the network's
// uncertainty.  Screening flags a grid cell for re-solving by the integrator
// when the heads disagree, when the predicted class differs from a
// neighbour's, and within a dilation radius of such cells.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../../include/clode_mlp.h"

// Staging memory of clode_mlp_screen_grid_noalloc: one byte per grid cell for
// the dilation's second flag buffer.  Each device holds its own copy of the
// array; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (clode_mlp_create records it).
static __device__ __align__(256) unsigned char scratch[CLODE_MLP_NOALLOC_MAX_CELLS];

namespace {

constexpr int NFF = 16;         // Fourier frequencies
constexpr int D0 = 2 * NFF;     // encoded input width
constexpr int H = 256;          // hidden width
constexpr int NHID = 4;         // hidden layers
constexpr int NHEAD = 4;        // heads
constexpr int NCLS = 16;        // classes
constexpr int HW = 2 + NCLS;    // outputs per head
constexpr int NOUT = NHEAD * HW;  // 72 used output columns
constexpr int NOUTP = 128;      // padded output columns
constexpr int NLAYER = NHID + 1;

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// round to nearest even (NaN kept quiet), identical on host and device
__host__ __device__ inline bf16_t f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
        return (bf16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// logistic sigmoid with the hardware exponential and reciprocal (v_exp_f32,
// v_rcp_f32; about 1 ulp each): the GEMM epilogues evaluate it once per
// output element, and an IEEE division there cost more than the products
__device__ inline float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigm(z); }
__device__ inline float dsilu(float z)
{
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}
// float -> BF16 on the device, round to nearest even: the gfx950 conversion
// instruction, which agrees with f2bf for every finite input
__device__ inline bf16_t f2bf_dev(float f) { return __builtin_bit_cast(bf16_t, (__bf16)f); }

// ---------------------------------------------------------------- GEMM
// C[M x N] = A[M x K] * B[K x N] with arbitrary strides on both operands,
// BF16 operands, FP32 accumulation on v_mfma_f32_16x16x32_bf16, and a fused
// epilogue.  Workgroup tile 64 x 64 (4 waves of 32 x 32, each 2 x 2 MFMA
// blocks), MFMA K step 32.  K is staged through LDS in chunks of 128 with k
// contiguous (each lane's 8-element operand is one 16-byte LDS read; the row
// pitch of 136 elements puts the 16 rows a read instruction touches on
// disjoint banks), and the global loads of the next chunk are issued before
// the MFMAs of the current one, so a 256-deep product waits on memory about
// once rather than once per K step.
enum Epi { EPI_FWD_SILU = 0, EPI_FWD_LINEAR = 1, EPI_BWD_DSILU = 2, EPI_WGRAD = 3 };

struct GemmArgs {
    int M, N, K;
    const bf16_t *A;
    long long lda_m, lda_k;
    const bf16_t *B;
    long long ldb_k, ldb_n;
    const float *bias;  // [N]                      (forward)
    const float *Zin;   // [M][ldz] pre-activation  (EPI_BWD_DSILU)
    float *Zout;        // [M][ldz] pre-activation  (forward; optional with SiLU)
    bf16_t *Yout;       // [M][ldy]                 (forward SiLU, dsilu)
    float *Cout;        // [M][ldc]                 (EPI_WGRAD)
    bf16_t *YTout;      // [N][ldt] transposed copy of Yout (optional)
    int ldz, ldy, ldc, ldt;
    int ksplit;         // K range of split z (blockIdx.z): [z * ksplit, (z + 1) * ksplit); 0 = all of K
    long long cstride;  // Cout offset between splits (EPI_WGRAD)
};

constexpr int TK = 32, KPAD = 8;

// Workgroup tiles: 64 x 64 (each of the 4 waves 32 x 32 = 2 x 2 MFMA blocks,
// K staged 128 deep) for the training products of a few thousand rows, where
// the number of workgroups matters; 128 x 128 (each wave 64 x 64 = 4 x 4
// blocks, K staged 64 deep) for the screening products, where every 16-byte
// LDS operand read then feeds four MFMAs instead of two.  The row pitch of
// KC + 8 elements keeps the 16 rows an operand read touches on disjoint banks.
template <int T>
struct Tile;
template <>
struct Tile<32> {
    static constexpr int TM = 32, TN = 32, KC = 128, WB = 1;
};
template <>
struct Tile<64> {
    static constexpr int TM = 64, TN = 64, KC = 128, WB = 2;
};
template <>
struct Tile<128> {
    static constexpr int TM = 128, TN = 128, KC = 64, WB = 4;
};

// rows [r0, r0 + TM) x k [k, k + KC) of a strided BF16 operand into registers
// (zero outside [0, rows) x [k, kend)); slot j of thread t is row
// (t + 256 j) / KG, group (t + 256 j) % KG, so consecutive threads read one
// contiguous row segment
template <int TM, int KC>
__device__ inline void gemm_load(const bf16_t *X, long long ld_r, long long ld_k, bool vec, int rows, int r0,
                                 int k, int kend, u16x8 (&v)[TM * (KC / 8) / 256])
{
    constexpr int KG = KC / 8;
#pragma unroll
    for (int j = 0; j < TM * KG / 256; ++j) {
        const int slot = (int)threadIdx.x + 256 * j;
        const int r = r0 + slot / KG, kk = k + (slot % KG) * 8;
        if (vec && r < rows && kk + 8 <= kend) {
            v[j] = *(const u16x8 *)(X + (long long)r * ld_r + kk);
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e)
                v[j][e] = (r < rows && kk + e < kend) ? X[(long long)r * ld_r + (long long)(kk + e) * ld_k] : 0;
        }
    }
}

template <int TM, int KC>
__device__ inline void gemm_stage(bf16_t *S, const u16x8 (&v)[TM * (KC / 8) / 256])
{
    constexpr int KG = KC / 8, KCP = KC + KPAD;
#pragma unroll
    for (int j = 0; j < TM * KG / 256; ++j) {
        const int slot = (int)threadIdx.x + 256 * j;
        *(u16x8 *)&S[(slot / KG) * KCP + (slot % KG) * 8] = v[j];
    }
}

// workgroup-tile product of one block (bx, by, bz) of the grid
// (ceil(N / TN), ceil(M / TM), splits); smem holds (TM + TN) x (KC + KPAD)
// BF16 elements
template <int EPI, int T>
__device__ __attribute__((always_inline)) inline void gemm_block(const GemmArgs &g, int bx, int by, int bz,
                                                                 bf16_t *smem)
{
    constexpr int TM = Tile<T>::TM, TN = Tile<T>::TN, KC = Tile<T>::KC, WB = Tile<T>::WB;
    constexpr int KCP = KC + KPAD, LDT = TM * (KC / 8) / 256;
    bf16_t *As = smem, *Bs = smem + TM * KCP;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = by * TM, n0 = bx * TN;
    const int kbeg = g.ksplit > 0 ? bz * g.ksplit : 0;
    const int kend = g.ksplit > 0 ? min(g.K, kbeg + g.ksplit) : g.K;
    f32x4 acc[WB][WB];
#pragma unroll
    for (int i = 0; i < WB; ++i)
#pragma unroll
        for (int j = 0; j < WB; ++j)
            acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const bool a_vec = g.lda_k == 1 && (g.lda_m % 8) == 0 && ((uintptr_t)g.A & 15) == 0;
    const bool b_vec = g.ldb_k == 1 && (g.ldb_n % 8) == 0 && ((uintptr_t)g.B & 15) == 0;
    u16x8 va[LDT], vb[LDT];
    if (kbeg < kend) {
        gemm_load<TM, KC>(g.A, g.lda_m, g.lda_k, a_vec, g.M, m0, kbeg, kend, va);
        gemm_load<TN, KC>(g.B, g.ldb_n, g.ldb_k, b_vec, g.N, n0, kbeg, kend, vb);
    }
    for (int kc = kbeg; kc < kend; kc += KC) {
        gemm_stage<TM, KC>(As, va);
        gemm_stage<TN, KC>(Bs, vb);
        __syncthreads();
        if (kc + KC < kend) {  // next chunk in flight during this chunk's MFMAs
            gemm_load<TM, KC>(g.A, g.lda_m, g.lda_k, a_vec, g.M, m0, kc + KC, kend, va);
            gemm_load<TN, KC>(g.B, g.ldb_n, g.ldb_k, b_vec, g.N, n0, kc + KC, kend, vb);
        }
        const int nks = (min(KC, kend - kc) + TK - 1) / TK;
        const int r = lane & 15;
        for (int ks = 0; ks < nks; ++ks) {
            const int kk = ks * TK + (lane >> 4) * 8;
            bf16x8 bfr[WB];
#pragma unroll
            for (int j = 0; j < WB; ++j)
                bfr[j] = *(const bf16x8 *)&Bs[(wn * (TN / 2) + j * 16 + r) * KCP + kk];
#pragma unroll
            for (int i = 0; i < WB; ++i) {
                const bf16x8 a = *(const bf16x8 *)&As[(wm * (TM / 2) + i * 16 + r) * KCP + kk];
#pragma unroll
                for (int j = 0; j < WB; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bfr[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();  // before the next chunk overwrites the tiles
    }
    // epilogue: lane holds rows (lane >> 4) * 4 + q, column lane & 15 of each
    // block.  The transposed copy (YTout) is first gathered in LDS (the
    // tiles are free once the K loop's last barrier has passed) and then
    // written a row of YT at a time with 16-byte stores, instead of one
    // scattered 2-byte store per element.
    constexpr int TP = TM + 8;  // pitch of the transposed tile
    static_assert(TN * TP <= (TM + TN) * KCP, "the transposed tile fits in the staging LDS");
    bf16_t *Tt = smem;
    const bool stage_t = (EPI == EPI_FWD_SILU || EPI == EPI_BWD_DSILU) && g.YTout != nullptr;
#pragma unroll
    for (int i = 0; i < WB; ++i)
#pragma unroll
        for (int j = 0; j < WB; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int ml = wm * (TM / 2) + i * 16 + (lane >> 4) * 4 + q;
                const int nl = wn * (TN / 2) + j * 16 + (lane & 15);
                const int m = m0 + ml;
                const int n = n0 + nl;
                if (m >= g.M || n >= g.N)
                    continue;
                float v = acc[i][j][q];
                if (EPI == EPI_FWD_SILU || EPI == EPI_FWD_LINEAR) {
                    v += g.bias[n];
                    if (EPI == EPI_FWD_LINEAR || g.Zout)
                        g.Zout[(long long)m * g.ldz + n] = v;
                    if (EPI == EPI_FWD_SILU) {
                        const bf16_t y = f2bf_dev(silu(v));
                        g.Yout[(long long)m * g.ldy + n] = y;
                        if (stage_t)
                            Tt[nl * TP + ml] = y;
                    }
                } else if (EPI == EPI_BWD_DSILU) {
                    const bf16_t y = f2bf_dev(v * dsilu(g.Zin[(long long)m * g.ldz + n]));
                    g.Yout[(long long)m * g.ldy + n] = y;
                    if (stage_t)
                        Tt[nl * TP + ml] = y;
                } else {
                    g.Cout[(long long)bz * g.cstride + (long long)m * g.ldc + n] = v;
                }
            }
    if (stage_t) {  // uniform across the workgroup
        __syncthreads();
        const bool vec = (g.ldt % 8) == 0 && ((uintptr_t)g.YTout & 15) == 0;
#pragma unroll
        for (int slot = (int)threadIdx.x; slot < TN * TM / 8; slot += 256) {
            const int nl = slot / (TM / 8), ml = (slot % (TM / 8)) * 8;
            const int n = n0 + nl, m = m0 + ml;
            if (n >= g.N)
                continue;
            bf16_t *dst = g.YTout + (long long)n * g.ldt + m;
            if (vec && m + 8 <= g.M) {
                *(u16x8 *)dst = *(const u16x8 *)&Tt[nl * TP + ml];
            } else {
                for (int e = 0; e < 8 && m + e < g.M; ++e)
                    dst[e] = Tt[nl * TP + ml + e];
            }
        }
    }
}

template <int EPI, int T>
__global__ void __launch_bounds__(256) mlp_gemm(const GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) bf16_t smem[(Tile<T>::TM + Tile<T>::TN) * (Tile<T>::KC + KPAD)];
    gemm_block<EPI, T>(g, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, smem);
}

// two independent products in one launch (a layer's weight gradient and its
// input gradient): blocks [0, n1) are the first product's grid (gx1 x gy1 x
// splits, x fastest), the rest the second's (gx2 wide); each block computes
// exactly what it computes in a launch of its own
template <int EPI1, int EPI2, int T>
__global__ void __launch_bounds__(256) mlp_gemm_pair(const GemmArgs g1, int gx1, int gy1, int n1, const GemmArgs g2,
                                                     int gx2)
{
    __shared__ __attribute__((aligned(16))) bf16_t smem[(Tile<T>::TM + Tile<T>::TN) * (Tile<T>::KC + KPAD)];
    const int b = (int)blockIdx.x;
    if (b < n1)
        gemm_block<EPI1, T>(g1, b % gx1, (b / gx1) % gy1, b / (gx1 * gy1), smem);
    else
        gemm_block<EPI2, T>(g2, (b - n1) % gx2, (b - n1) / gx2, 0, smem);
}

// ---------------------------------------------------------------- encode
// Fourier feature f of the input (x0, x1): sin and cos of
// 2 pi (B[0][f] x0 + B[1][f] x1), rounded to BF16
__device__ __attribute__((always_inline)) inline void encode_pair(float x0, float x1, int f, const float *Bff, bf16_t &s, bf16_t &c)
{
    const float p = 6.283185307179586f * (Bff[f] * x0 + Bff[NFF + f] * x1);
    s = f2bf(sinf(p));
    c = f2bf(cosf(p));
}

// row i of the encoded inputs; A0t (optional): transposed copy [32][ldt]
__device__ inline void encode_row(float x0, float x1, int i, const float *Bff, bf16_t *A0, bf16_t *A0t, int ldt)
{
    for (int f = 0; f < NFF; ++f) {
        bf16_t s, c;
        encode_pair(x0, x1, f, Bff, s, c);
        A0[(long long)i * D0 + 2 * f] = s;
        A0[(long long)i * D0 + 2 * f + 1] = c;
        if (A0t) {
            A0t[(long long)(2 * f) * ldt + i] = s;
            A0t[(long long)(2 * f + 1) * ldt + i] = c;
        }
    }
}

__global__ void __launch_bounds__(256) mlp_encode(const float *X, int n, const float *Bff, bf16_t *A0, bf16_t *A0t,
                                                  int ldt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    encode_row(X[2 * i], X[2 * i + 1], i, Bff, A0, A0t, ldt);
}

// grid inputs: cell c = ix * ny + iy at (lo + (hi - lo) * i / (n - 1)), normalised to [0, 1]
__device__ __attribute__((always_inline)) inline void grid_point(long long c, int nx, int ny, float &x0, float &x1)
{
    const int ix = (int)(c / ny), iy = (int)(c % ny);
    x0 = nx > 1 ? (float)ix / (float)(nx - 1) : 0.f;
    x1 = ny > 1 ? (float)iy / (float)(ny - 1) : 0.f;
}

__global__ void __launch_bounds__(256) mlp_grid_inputs(long long c0, int n, int nx, int ny, float *X)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    grid_point(c0 + i, nx, ny, X[2 * i], X[2 * i + 1]);
}

// ---------------------------------------------------------------- loss
__host__ __device__ inline bool head_keeps(unsigned long long sample, int head)
{
    unsigned long long z = sample * 4ull + (unsigned long long)head + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (z & 3ull) != 0ull;
}

// Per sample and head: squared error of the two normalised regression
// targets plus the softmax cross-entropy of the class, over the samples the
// head keeps; the loss is the mean over the batch of the sum over heads.
// dOut (BF16, [n][128]) is the gradient of that loss w.r.t. the head outputs,
// dOutT ([128][ldt]) its transpose.  One thread per (sample, head), sample
// fastest, so the transposed rows are written coalesced; the columns past the
// heads stay zero in dOut (cleared when the buffer is allocated) and are
// cleared here in dOutT, whose row pitch follows the batch size.
constexpr int PADROWS = (NOUTP - NOUT) / NHEAD;  // transposed padding rows each head thread clears
static_assert(PADROWS * NHEAD == NOUTP - NOUT, "padding rows split evenly over the heads");

__global__ void __launch_bounds__(256) mlp_loss(const float *Zh, int n, const float *T, const int *cls,
                                                const unsigned long long *ids, float inv_n, bf16_t *dOut,
                                                bf16_t *dOutT, int ldt, float *loss)
{
    __shared__ float red[256];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (t < (long long)NHEAD * n) {
        const int h = (int)(t / n), i = (int)(t % n);
        const float *zh = Zh + (long long)i * NOUTP + h * HW;
        bf16_t d[HW];
        const unsigned long long id = ids ? ids[i] : (unsigned long long)i;
        if (!head_keeps(id, h)) {
            for (int c = 0; c < HW; ++c)
                d[c] = 0;
        } else {
            const float e0 = zh[0] - T[2 * i], e1 = zh[1] - T[2 * i + 1];
            l += e0 * e0 + e1 * e1;
            d[0] = f2bf(2.f * e0 * inv_n);
            d[1] = f2bf(2.f * e1 * inv_n);
            float mx = zh[2];
            for (int c = 1; c < NCLS; ++c)
                mx = fmaxf(mx, zh[2 + c]);
            // each class's exponential once, for the sum and the gradient
            float ex[NCLS];
            float se = 0.f;
#pragma unroll
            for (int c = 0; c < NCLS; ++c) {
                ex[c] = __expf(zh[2 + c] - mx);
                se += ex[c];
            }
            const int y = cls[i];
            l += (mx + __logf(se)) - zh[2 + y];
#pragma unroll
            for (int c = 0; c < NCLS; ++c) {
                const float p = ex[c] / se;
                d[2 + c] = f2bf((p - (c == y ? 1.f : 0.f)) * inv_n);
            }
        }
        bf16_t *dh = dOut + (long long)i * NOUTP + h * HW;
        for (int c = 0; c < HW; ++c) {
            dh[c] = d[c];
            dOutT[(long long)(h * HW + c) * ldt + i] = d[c];
        }
        for (int c = NOUT + h * PADROWS; c < NOUT + (h + 1) * PADROWS; ++c)
            dOutT[(long long)c * ldt + i] = 0;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(loss, red[0] * inv_n);
}

// one launch for both per-layer reductions of the backward pass: blocks
// [0, sum_blocks) sum the split-K partial weight gradients in a fixed order
// (gw[e] = sum_z part[z * count + e]); the remaining ncol blocks form the bias
// gradient db[c] = sum_m dZ[m][c] from the transposed copy dZt[c][m] (row
// stride ld), each block reading one contiguous row
__global__ void __launch_bounds__(256) mlp_layer_grads(const float *part, int splits, long long count, float *gw,
                                                       int sum_blocks, const bf16_t *dZt, int n, int ncol, int ld,
                                                       float *db)
{
    __shared__ float red[256];
    if ((int)blockIdx.x < sum_blocks) {
        const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        float s = 0.f;
        for (int z = 0; z < splits; ++z)
            s += part[(long long)z * count + e];
        gw[e] = s;
        return;
    }
    const int c = (int)blockIdx.x - sum_blocks;
    float s = 0.f;
    for (int m = threadIdx.x; m < n; m += blockDim.x)
        s += bf2f(dZt[(long long)c * ld + m]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
            red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0 && c < ncol)
        db[c] = red[0];
}

// layer table for the Adam kernels: weights of layer l occupy
// [woff, woff + in * out) as Wt[out][in], its biases the out entries after
// them; the transposed BF16 copy keeps W[in][out] at the same offsets
struct LayerTable {
    long long woff[NLAYER];
    int in[NLAYER], out[NLAYER];
};

struct AdamStep {
    float lr, b1, b2, eps, c1, c2;
};

// Adam (Kingma & Ba 2015) on one FP32 master parameter, refreshing both BF16
// copies of the weights
__device__ __forceinline__ void adam_update(long long i, float gi, float *p, float *m1, float *m2, bf16_t *pbf,
                                            bf16_t *pbfT, const LayerTable &lt, const AdamStep &as)
{
    const float a = as.b1 * m1[i] + (1.f - as.b1) * gi;
    const float b = as.b2 * m2[i] + (1.f - as.b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = p[i] - as.lr * (a * as.c1) / (sqrtf(b * as.c2) + as.eps);
    p[i] = v;
    const bf16_t b16 = f2bf(v);
    pbf[i] = b16;
    for (int l = 0; l < NLAYER; ++l) {
        const long long r = i - lt.woff[l];
        if (r >= 0 && r < (long long)lt.in[l] * lt.out[l]) {
            const int o = (int)(r / lt.in[l]), k = (int)(r % lt.in[l]);
            pbfT[lt.woff[l] + (long long)k * lt.out[l] + o] = b16;
        }
    }
}

// Adam from a finished gradient vector g; frozen entries (mask 0: head
// padding) stay put
__global__ void __launch_bounds__(256) mlp_adam(float *p, const float *g, float *m1, float *m2, const uint8_t *mask,
                                                bf16_t *pbf, bf16_t *pbfT, const LayerTable lt, long long n,
                                                const AdamStep as)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !mask[i])
        return;
    adam_update(i, g[i], p, m1, m2, pbf, pbfT, lt, as);
}

// Adam that also finishes the gradient, in place of the per-layer reduction
// launches of the backward pass (the training loop's optimizer step).
// Blocks [0, wblocks): one thread per parameter; a weight's gradient is the
// sum of its split-K partial products part[z * pstride + i], z = 0 .. splits-1
// in order (the sum mlp_layer_grads forms).  The following blocks take one
// bias each, layer by layer: db[c] = sum_m dZt_l[c][m] over the minibatch by
// the same strided sums and tree as mlp_layer_grads, then thread 0 updates it.
// The updated parameters equal those of the separate launches bit for bit.
struct GradSources {
    const float *part;
    long long pstride;
    int splits;
    const bf16_t *dzt[NLAYER];  // transposed output gradient of each layer, [out][ld]
    int n, ld;
};

__global__ void __launch_bounds__(256) mlp_adam_sum(float *p, float *m1, float *m2, const uint8_t *mask, bf16_t *pbf,
                                                    bf16_t *pbfT, const LayerTable lt, long long n, int wblocks,
                                                    const GradSources gs, const AdamStep as)
{
    if ((int)blockIdx.x < wblocks) {
        const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n || !mask[i])
            return;
        int l = 0;
        while (l < NLAYER - 1 && i >= lt.woff[l + 1])
            ++l;
        if (i - lt.woff[l] >= (long long)lt.in[l] * lt.out[l])
            return;  // a bias: the blocks below
        float s = 0.f;
        for (int z = 0; z < gs.splits; ++z)
            s += gs.part[(long long)z * gs.pstride + i];
        adam_update(i, s, p, m1, m2, pbf, pbfT, lt, as);
        return;
    }
    __shared__ float red[256];
    int c = (int)blockIdx.x - wblocks, l = 0;
    while (l < NLAYER - 1 && c >= lt.out[l]) {
        c -= lt.out[l];
        ++l;
    }
    const bf16_t *dzt = gs.dzt[l];
    float s = 0.f;
    for (int m = threadIdx.x; m < gs.n; m += blockDim.x)
        s += bf2f(dzt[(long long)c * gs.ld + m]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
            red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    const long long i = lt.woff[l] + (long long)lt.in[l] * lt.out[l] + c;
    if (threadIdx.x == 0 && mask[i])
        adam_update(i, red[0], p, m1, m2, pbf, pbfT, lt, as);
}

// minibatch gather: row r of the batch is sample perm[r] of the training
// set; its inputs are encoded (with the transposed copy) in the same pass.
// One thread per (row, frequency), row fastest, so that a minibatch of a
// few thousand rows spreads its sin/cos over 16x more lanes and the
// transposed rows are written coalesced.
__global__ void __launch_bounds__(256) mlp_gather(const unsigned int *perm, int nb, const float *Xa, const float *Ta,
                                                  const int *Ca, float *T, int *C, unsigned long long *ids,
                                                  const float *Bff, bf16_t *A0, bf16_t *A0t, int ldt)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb * NFF)
        return;
    const int r = idx % nb, f = idx / nb;
    const unsigned int j = perm[r];
    if (f == 0) {
        T[2 * r] = Ta[2 * j];
        T[2 * r + 1] = Ta[2 * j + 1];
        C[r] = Ca[j];
        ids[r] = j;
    }
    bf16_t sv, cv;
    encode_pair(Xa[2 * j], Xa[2 * j + 1], f, Bff, sv, cv);
    A0[(long long)r * D0 + 2 * f] = sv;
    A0[(long long)r * D0 + 2 * f + 1] = cv;
    A0t[(long long)(2 * f) * ldt + r] = sv;
    A0t[(long long)(2 * f + 1) * ldt + r] = cv;
}

// ---------------------------------------------------------------- screening
// Per cell: each head's (log-period, amplitude) back to feature units and
// class argmax; the ensemble prediction is the head mean (softmax mean for the
// class).  flags bit 0: the heads disagree (relative spread above tol, or a
// head's class differs from the ensemble class).
// z: the 72 head outputs of one cell
__device__ __attribute__((always_inline)) inline void predict_row(const float *z, const float *tmean,
                                                                  const float *tstd, float tol, float &period,
                                                                  float &amp, int &cls, uint8_t &flags)
{
    // every loop has a constant trip count and is unrolled, so the arrays
    // live in registers (the first-occurrence argmax keeps its best value
    // instead of indexing prob[] at a run-time index)
    float P[NHEAD], A[NHEAD], prob[NCLS];
    int hc[NHEAD];
#pragma unroll
    for (int c = 0; c < NCLS; ++c)
        prob[c] = 0.f;
    float pm = 0.f, am = 0.f;
#pragma unroll
    for (int h = 0; h < NHEAD; ++h) {
        const float *zh = z + h * HW;
        P[h] = expf(zh[0] * tstd[0] + tmean[0]);
        A[h] = zh[1] * tstd[1] + tmean[1];
        pm += P[h];
        am += A[h];
        float mx = zh[2];
        int am_i = 0;
#pragma unroll
        for (int c = 1; c < NCLS; ++c)
            if (zh[2 + c] > mx) {
                mx = zh[2 + c];
                am_i = c;
            }
        hc[h] = am_i;
        float se = 0.f;
#pragma unroll
        for (int c = 0; c < NCLS; ++c)
            se += __expf(zh[2 + c] - mx);
#pragma unroll
        for (int c = 0; c < NCLS; ++c)
            prob[c] += __expf(zh[2 + c] - mx) / se;
    }
    pm *= 1.f / NHEAD;
    am *= 1.f / NHEAD;
    int best = 0;
    float pbest = prob[0];
#pragma unroll
    for (int c = 1; c < NCLS; ++c)
        if (prob[c] > pbest) {
            pbest = prob[c];
            best = c;
        }
    bool dis = false;
#pragma unroll
    for (int h = 0; h < NHEAD; ++h) {
        dis |= fabsf(P[h] - pm) > tol * fabsf(pm);
        dis |= fabsf(A[h] - am) > tol * fabsf(am);
        dis |= hc[h] != best;
    }
    period = pm;
    amp = am;
    cls = best;
    flags = dis ? 1 : 0;
}

__global__ void __launch_bounds__(256) mlp_predict_cells(const float *Zh, int n, const float *tmean,
                                                         const float *tstd, float tol, float *period, float *amp,
                                                         int *cls, uint8_t *flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    predict_row(Zh + (long long)i * NOUTP, tmean, tstd, tol, period[i], amp[i], cls[i], flags[i]);
}

// flags bit 1: a 4-neighbour on the nx x ny grid has another class
__global__ void __launch_bounds__(256) mlp_boundary(const int *cls, int nx, int ny, uint8_t *flags)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)nx * ny)
        return;
    const int ix = (int)(c / ny), iy = (int)(c % ny);
    const int k = cls[c];
    bool b = (ix > 0 && cls[c - ny] != k) || (ix + 1 < nx && cls[c + ny] != k) || (iy > 0 && cls[c - 1] != k) ||
             (iy + 1 < ny && cls[c + 1] != k);
    if (b)
        flags[c] |= 2;
}

// one step of 3 x 3 dilation of "any flag" into bit 2 (out must not alias in)
__global__ void __launch_bounds__(256) mlp_dilate(const uint8_t *in, int nx, int ny, uint8_t *out)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long long)nx * ny)
        return;
    const int ix = (int)(c / ny), iy = (int)(c % ny);
    uint8_t f = in[c];
    bool any = false;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const int jx = ix + dx, jy = iy + dy;
            if (jx >= 0 && jx < nx && jy >= 0 && jy < ny && in[(long long)jx * ny + jy])
                any = true;
        }
    out[c] = (uint8_t)(f | (any && !f ? 4 : 0));
}

// ---------------------------------------------------------------- host
thread_local std::string g_err;

// ---------------------------------------------------------------- fused screening
// One workgroup per block of SROWS grid cells runs the whole network:
// encode, the four hidden layers, the heads and the per-cell prediction.
// The activations stay in LDS between layers; only the weights (0.4 MB,
// L2-resident) and the four per-cell results touch memory, where the
// layer-by-layer path writes and re-reads 0.5 KB of activations per cell and
// layer.  Each wave owns 64 output columns of a hidden layer (8 x 4 blocks of
// 16 x 16) or 32 of the heads (8 x 2), takes its weight fragments straight
// from global memory with the next K step's in flight, and reads the
// activations from LDS.  Operand layout, K order (32-deep steps, ascending)
// and epilogue arithmetic are those of mlp_gemm, so every output equals the
// layer-by-layer path bit for bit.
//
// LDS: the activations [SROWS][H] BF16, 64 KiB, two workgroups per CU.  The
// 16-byte chunk kc of row m is stored at chunk kc ^ (m & 15) of that row, so
// the 16 rows an operand read touches fall on 16 different chunk positions
// (no padding, which would not fit two workgroups).  After the last hidden
// layer the same bytes hold the FP32 head outputs [SROWS][ZPITCH].
constexpr int SROWS = 128;         // cells per workgroup
constexpr int ZPITCH = NOUT + 1;   // row pitch of the head outputs (floats)
static_assert(SROWS * ZPITCH * 4 <= SROWS * H * 2, "head outputs fit in the activation buffer");
static_assert(SROWS == 8 * 16 && H == 4 * 64 && NOUTP == 4 * 32, "wave tiling");

struct ScreenArgs {
    long long ncell;
    int nx, ny;
    const float *Bff;
    const bf16_t *W[NLAYER];  // Wt[out][in], BF16
    const float *bias[NLAYER];
    const float *tm;  // tmean[2], tstd[2]
    float tol;
    float *period, *amp;
    int *cls;
    uint8_t *flags;
};

// The lane index, recomputed where it is used: an opaque value keeps the
// compiler from computing every layer's epilogue LDS addresses once at the
// top of the kernel, where they would stay live across all the K loops and
// spill (two v_mbcnt instructions per epilogue instead).
__device__ __attribute__((always_inline)) inline int lane_here()
{
    int v;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(v));
    return v;
}

// element (m, k) of the swizzled activation tile
__device__ __attribute__((always_inline)) inline int sw(int m, int k) { return m * H + ((((k >> 3) ^ (m & 15))) << 3) + (k & 7); }

// acc[i][j] += act[rows of block i][0, K) * W[columns of block j][0, K)^T for
// the NB column blocks starting at column n0 (16 NB columns)
template <int K, int NB>
__device__ __attribute__((always_inline)) inline void screen_product(const bf16_t *act, const bf16_t *__restrict__ W, int n0, int lane,
                                      f32x4 (&acc)[8][NB])
{
    constexpr int NKS = K / TK;
    const int r = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
            acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 b[NB], bn[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j)
        b[j] = *(const bf16x8 *)(W + (long long)(n0 + j * 16 + r) * K + kq);
#pragma unroll 1
    for (int ks = 0; ks < NKS; ++ks) {
        if (ks + 1 < NKS) {
#pragma unroll
            for (int j = 0; j < NB; ++j)
                bn[j] = *(const bf16x8 *)(W + (long long)(n0 + j * 16 + r) * K + (ks + 1) * TK + kq);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const bf16x8 a = *(const bf16x8 *)&act[sw(i * 16 + r, ks * TK + kq)];
#pragma unroll
            for (int j = 0; j < NB; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
        if (ks + 1 < NKS) {
#pragma unroll
            for (int j = 0; j < NB; ++j)
                b[j] = bn[j];
        }
    }
}

// hidden layer in place: act <- SiLU(act W^T + bias), BF16
template <int K>
__device__ __attribute__((always_inline)) inline void screen_hidden(bf16_t *act, const bf16_t *__restrict__ W, const float *__restrict__ bias,
                                     int wave, int lane)
{
    f32x4 acc[8][4];
    screen_product<K, 4>(act, W, wave * 64, lane, acc);
    __syncthreads();  // every wave has read the layer's input
    lane = lane_here();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int n = wave * 64 + j * 16 + (lane & 15);
        const float bj = bias[n];
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float v = acc[i][j][q];
                v += bj;
                act[sw(i * 16 + (lane >> 4) * 4 + q, n)] = f2bf_dev(silu(v));
            }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) mlp_screen_fused(const ScreenArgs s)
{
    __shared__ __attribute__((aligned(16))) bf16_t act[SROWS * H];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const long long c0 = (long long)blockIdx.x * SROWS;
    {  // encode: thread t -> cell row t / 2, frequencies (t % 2) * 8 .. + 8
        const int m = t >> 1, f0 = (t & 1) * (NFF / 2);
        const long long c = c0 + m < s.ncell ? c0 + m : s.ncell - 1;
        float x0, x1;
        grid_point(c, s.nx, s.ny, x0, x1);
        for (int f = f0; f < f0 + NFF / 2; ++f) {
            bf16_t sv, cv;
            encode_pair(x0, x1, f, s.Bff, sv, cv);
            act[sw(m, 2 * f)] = sv;
            act[sw(m, 2 * f + 1)] = cv;
        }
    }
    __syncthreads();
    screen_hidden<D0>(act, s.W[0], s.bias[0], wave, lane);
    screen_hidden<H>(act, s.W[1], s.bias[1], wave, lane);
    screen_hidden<H>(act, s.W[2], s.bias[2], wave, lane);
    screen_hidden<H>(act, s.W[3], s.bias[3], wave, lane);
    // heads: 128 columns, 32 per wave; the 72 used ones go to LDS in FP32
    f32x4 acc[8][2];
    screen_product<H, 2>(act, s.W[NHID], wave * 32, lane, acc);
    __syncthreads();  // the activations are no longer needed: z overlays them
    float *z = reinterpret_cast<float *>(act);
    const int hl = lane_here();
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = wave * 32 + j * 16 + (hl & 15);
        if (n < NOUT) {
            const float bj = s.bias[NHID][n];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float v = acc[i][j][q];
                    v += bj;
                    z[(i * 16 + (hl >> 4) * 4 + q) * ZPITCH + n] = v;
                }
        }
    }
    __syncthreads();
    if (t < SROWS && c0 + t < s.ncell) {
        const long long c = c0 + t;
        predict_row(z + t * ZPITCH, s.tm, s.tm + 2, s.tol, s.period[c], s.amp[c], s.cls[c], s.flags[c]);
    }
}

struct MlpError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw MlpError(std::string(what) + ": " + hipGetErrorString(e));
}

// Device address of the static `scratch` array, per device: looked up once
// per device, with that device current, by the first clode_mlp_create on it.
std::mutex g_scratch_mu;
std::vector<unsigned char *> g_scratch_addr;

unsigned char *scratch_address(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)device >= g_scratch_addr.size())
        g_scratch_addr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_addr[device] = static_cast<unsigned char *>(p);
    }
    return g_scratch_addr[device];
}

template <class T>
struct Dev {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t k)
    {
        if (k <= n && p)
            return;
        free();
        check(hipMalloc((void **)&p, std::max<size_t>(k, 1) * sizeof(T)), "hipMalloc");
        n = k;
    }
    void free()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~Dev() { free(); }
};

struct Layer {
    int in, out;        // out x in weights, stored transposed Wt[out][in]
    size_t woff, boff;  // offsets in the flat parameter vector
};

}  // namespace

struct clode_mlp {
    int device = 0;
    int cus = 256;  // compute units of the device (tile-size choice)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Layer L[NLAYER];
    size_t nparam = 0;
    std::vector<float> Bff;       // [2][16] Fourier frequencies
    float tmean[2] = {0, 0}, tstd[2] = {1, 1};
    Dev<float> p, g, m1, m2, dBff, dT, dtm;
    Dev<uint8_t> mask;
    Dev<bf16_t> pbf, pbfT;  // BF16 weights Wt[out][in] and W[in][out]
    // activations for a chunk of rows, and transposed copies [width][ldt]
    // (ldt = rows rounded up to 8) for the weight-gradient products
    Dev<bf16_t> A0, Y[NHID], dY[NHID], dOut;
    Dev<bf16_t> A0t, Yt[NHID], dYt[NHID], dOutT;
    Dev<float> Z[NHID], Zh, X, T, loss, wsp;
    Dev<float> wall;  // split-K partial weight gradients of all layers (training), KSPLIT_MAX x nparam
    Dev<int> C;
    Dev<unsigned long long> ids;
    long long adamStep = 0;
    double lastMs = 0;
    size_t rows = 0;
    unsigned char *stage = nullptr;  // this device's static scratch (no-allocation screening)

    void ensureRows(size_t n)
    {
        if (n <= rows)
            return;
        const size_t nt = (n + 7) & ~(size_t)7;
        A0.alloc(n * D0);
        A0t.alloc(nt * D0);
        for (int l = 0; l < NHID; ++l) {
            Y[l].alloc(n * H);
            dY[l].alloc(n * H);
            Yt[l].alloc(nt * H);
            dYt[l].alloc(nt * H);
            Z[l].alloc(n * H);
        }
        dOut.alloc(n * NOUTP);
        // the loss kernel never writes the padding columns: they stay zero
        check(hipMemsetAsync(dOut.p, 0, n * NOUTP * sizeof(bf16_t), stream), "hipMemsetAsync");
        dOutT.alloc(nt * NOUTP);
        Zh.alloc(n * NOUTP);
        X.alloc(n * 2);
        T.alloc(n * 2);
        C.alloc(n);
        ids.alloc(n);
        rows = n;
    }
};

namespace {

template <int T>
void gemm_tiled(clode_mlp *h, int epi, const GemmArgs &a)
{
    constexpr int TM = Tile<T>::TM, TN = Tile<T>::TN;
    const int splits = a.ksplit > 0 ? (a.K + a.ksplit - 1) / a.ksplit : 1;
    dim3 grid((a.N + TN - 1) / TN, (a.M + TM - 1) / TM, splits);
    switch (epi) {
    case EPI_FWD_SILU: hipLaunchKernelGGL((mlp_gemm<EPI_FWD_SILU, T>), grid, dim3(256), 0, h->stream, a); break;
    case EPI_FWD_LINEAR: hipLaunchKernelGGL((mlp_gemm<EPI_FWD_LINEAR, T>), grid, dim3(256), 0, h->stream, a); break;
    case EPI_BWD_DSILU: hipLaunchKernelGGL((mlp_gemm<EPI_BWD_DSILU, T>), grid, dim3(256), 0, h->stream, a); break;
    default: hipLaunchKernelGGL((mlp_gemm<EPI_WGRAD, T>), grid, dim3(256), 0, h->stream, a); break;
    }
    check(hipGetLastError(), "mlp_gemm launch");
}

// 128 x 128 tiles once the product has enough rows to keep every CU busy
// with them (screening chunks); 64 x 64 below that; 32 x 32 when 64 x 64
// tiles would leave CUs idle (the training minibatch products: 2048 x 256
// is 128 tiles of 64 x 64 on 256 CUs).  Every tile size accumulates each
// output element over the same K chunks in the same order, so the results
// are identical.
int tile_for(const clode_mlp *h, const GemmArgs &a)
{
    const long long splits = a.ksplit > 0 ? (a.K + a.ksplit - 1) / a.ksplit : 1;
    const long long tiles64 = (long long)((a.M + 63) / 64) * ((a.N + 63) / 64) * splits;
    if (a.ksplit == 0 && a.M >= 32768)
        return 128;
    return tiles64 < h->cus ? 32 : 64;
}

void gemm(clode_mlp *h, int epi, const GemmArgs &a)
{
    switch (tile_for(h, a)) {
    case 128: gemm_tiled<128>(h, epi, a); break;
    case 32: gemm_tiled<32>(h, epi, a); break;
    default: gemm_tiled<64>(h, epi, a); break;
    }
}

// a layer's weight gradient w (split-K) and input gradient d (x SiLU') in
// one launch when both use 32 x 32 tiles (the training minibatches), as two
// launches otherwise; the same products either way
void gemm_pair(clode_mlp *h, const GemmArgs &w, const GemmArgs &d)
{
    if (tile_for(h, w) != 32 || tile_for(h, d) != 32) {
        gemm(h, EPI_WGRAD, w);
        gemm(h, EPI_BWD_DSILU, d);
        return;
    }
    const int gx1 = (w.N + 31) / 32, gy1 = (w.M + 31) / 32;
    const int sp1 = w.ksplit > 0 ? (w.K + w.ksplit - 1) / w.ksplit : 1;
    const int gx2 = (d.N + 31) / 32, gy2 = (d.M + 31) / 32;
    const int n1 = gx1 * gy1 * sp1;
    hipLaunchKernelGGL((mlp_gemm_pair<EPI_WGRAD, EPI_BWD_DSILU, 32>), dim3((unsigned int)(n1 + gx2 * gy2)),
                       dim3(256), 0, h->stream, w, gx1, gy1, n1, d, gx2);
    check(hipGetLastError(), "mlp_gemm_pair launch");
}

GemmArgs ga() { GemmArgs a; std::memset(&a, 0, sizeof(a)); return a; }

inline int ldT(int n) { return (n + 7) & ~7; }

constexpr int KSPLIT_MAX = 32;    // partial products per weight gradient
constexpr int KSPLIT_ROWS = 256;  // rows per partial product (at least)

// forward over n rows already encoded in h->A0; with train, the activations'
// transposed copies are written for backward
void forward(clode_mlp *h, int n, bool train = false)
{
    const bf16_t *in = h->A0.p;
    int kin = D0;
    for (int l = 0; l < NHID; ++l) {
        GemmArgs a = ga();
        a.M = n; a.N = H; a.K = kin;
        a.A = in; a.lda_m = kin; a.lda_k = 1;
        a.B = h->pbf.p + h->L[l].woff; a.ldb_k = 1; a.ldb_n = kin;
        a.bias = h->p.p + h->L[l].boff;
        a.ldz = H;
        a.Yout = h->Y[l].p; a.ldy = H;
        if (train) {  // backward needs the pre-activations and the transposed activations
            a.Zout = h->Z[l].p;
            a.YTout = h->Yt[l].p;
            a.ldt = ldT(n);
        }
        gemm(h, EPI_FWD_SILU, a);
        in = h->Y[l].p;
        kin = H;
    }
    GemmArgs a = ga();
    a.M = n; a.N = NOUTP; a.K = H;
    a.A = in; a.lda_m = H; a.lda_k = 1;
    a.B = h->pbf.p + h->L[NHID].woff; a.ldb_k = 1; a.ldb_n = H;
    a.bias = h->p.p + h->L[NHID].boff;
    a.Zout = h->Zh.p; a.ldz = NOUTP;
    gemm(h, EPI_FWD_LINEAR, a);
}

// backward from h->dOut / h->dOutT (BF16) into h->g; every product reads
// operands that are contiguous along its reduction index.  With gs (the
// training loop) the split-K partial weight gradients of every layer stay in
// h->wall (split z of parameter i at z * nparam + i) and the sums, with the
// bias gradients, are left to the optimizer launch (mlp_adam_sum); gs
// receives where they are.  Returns nothing to sum otherwise.
void backward(clode_mlp *h, int n, GradSources *gs = nullptr)
{
    const int lt = ldT(n);
    const bf16_t *dz = h->dOut.p, *dzt = h->dOutT.p;
    int nout = NOUTP;
    for (int l = NHID; l >= 0; --l) {
        const Layer &Ly = h->L[l];
        const bf16_t *xt = l > 0 ? h->Yt[l - 1].p : h->A0t.p;
        // dWt[o][i] = sum_m dz[m][o] x[m][i] = sum_m dzt[o][m] xt[i][m]
        // split the batch (the reduction) so that the product has enough
        // workgroups, then sum the partial products in a fixed order
        GemmArgs w = ga();
        w.M = Ly.out; w.N = Ly.in; w.K = n;
        w.A = dzt; w.lda_m = lt; w.lda_k = 1;
        w.B = xt; w.ldb_k = 1; w.ldb_n = lt;
        const int splits = std::max(1, std::min(KSPLIT_MAX, n / KSPLIT_ROWS));
        w.ksplit = ((n + splits - 1) / splits + TK - 1) / TK * TK;
        const int used = (n + w.ksplit - 1) / w.ksplit;
        const long long count = (long long)Ly.out * Ly.in;
        if (gs) {
            w.Cout = h->wall.p + Ly.woff; w.ldc = Ly.in; w.cstride = (long long)h->nparam;
            gs->splits = used;
            gs->dzt[l] = dzt;
            if (l == 0) {
                gemm(h, EPI_WGRAD, w);
                break;
            }
            // the weight gradient and the input gradient read the same
            // output gradient and nothing the other writes: one launch
            GemmArgs d = ga();
            d.M = n; d.N = Ly.in; d.K = Ly.out;
            d.A = dz; d.lda_m = nout; d.lda_k = 1;
            d.B = h->pbfT.p + Ly.woff; d.ldb_k = 1; d.ldb_n = Ly.out;
            d.Zin = h->Z[l - 1].p; d.ldz = H;
            d.Yout = h->dY[l - 1].p; d.ldy = H;
            d.YTout = h->dYt[l - 1].p; d.ldt = lt;
            gemm_pair(h, w, d);
            dz = h->dY[l - 1].p;
            dzt = h->dYt[l - 1].p;
            nout = H;
            continue;
        } else {
            w.Cout = h->wsp.p; w.ldc = Ly.in; w.cstride = count;
            gemm(h, EPI_WGRAD, w);
            const int sum_blocks = (int)((count + 255) / 256);
            hipLaunchKernelGGL(mlp_layer_grads, dim3((unsigned int)(sum_blocks + Ly.out)), dim3(256), 0, h->stream,
                               h->wsp.p, used, count, h->g.p + Ly.woff, sum_blocks, dzt, n, Ly.out, lt,
                               h->g.p + Ly.boff);
            check(hipGetLastError(), "mlp_layer_grads launch");
        }
        if (l == 0)
            break;
        // dY[l-1][m][i] = (sum_o dz[m][o] W[i][o]) * silu'(Z[l-1][m][i])
        GemmArgs d = ga();
        d.M = n; d.N = Ly.in; d.K = Ly.out;
        d.A = dz; d.lda_m = nout; d.lda_k = 1;
        d.B = h->pbfT.p + Ly.woff; d.ldb_k = 1; d.ldb_n = Ly.out;
        d.Zin = h->Z[l - 1].p; d.ldz = H;
        d.Yout = h->dY[l - 1].p; d.ldy = H;
        d.YTout = h->dYt[l - 1].p; d.ldt = lt;
        gemm(h, EPI_BWD_DSILU, d);
        dz = h->dY[l - 1].p;
        dzt = h->dYt[l - 1].p;
        nout = H;
    }
}

void encode(clode_mlp *h, const float *Xdev, int n, bool train = false)
{
    hipLaunchKernelGGL(mlp_encode, dim3((n + 255) / 256), dim3(256), 0, h->stream, Xdev, n, h->dBff.p, h->A0.p,
                       train ? h->A0t.p : nullptr, ldT(n));
    check(hipGetLastError(), "mlp_encode launch");
}

void upload(clode_mlp *h, void *dst, const void *src, size_t bytes)
{
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream), "hipMemcpy H2D");
}
void download(clode_mlp *h, void *dst, const void *src, size_t bytes)
{
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream), "hipMemcpy D2H");
    check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
}

void refresh_bf16(clode_mlp *h)
{
    std::vector<float> p(h->nparam);
    download(h, p.data(), h->p.p, h->nparam * 4);
    std::vector<bf16_t> b(h->nparam), bt(h->nparam, 0);
    for (size_t i = 0; i < b.size(); ++i)
        b[i] = f2bf(p[i]);
    for (int l = 0; l < NLAYER; ++l) {
        const Layer &Ly = h->L[l];
        for (int o = 0; o < Ly.out; ++o)
            for (int k = 0; k < Ly.in; ++k)
                bt[Ly.woff + (size_t)k * Ly.out + o] = b[Ly.woff + (size_t)o * Ly.in + k];
    }
    upload(h, h->pbf.p, b.data(), b.size() * 2);
    upload(h, h->pbfT.p, bt.data(), bt.size() * 2);
    check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
}

// loss and gradient over n rows (inputs X, targets T, classes C already on the device)
float loss_grad(clode_mlp *h, int n)
{
    encode(h, h->X.p, n, true);
    forward(h, n, true);
    check(hipMemsetAsync(h->loss.p, 0, 4, h->stream), "hipMemset");
    hipLaunchKernelGGL(mlp_loss, dim3((NHEAD * n + 255) / 256), dim3(256), 0, h->stream, h->Zh.p, n, h->T.p, h->C.p, h->ids.p,
                       1.0f / (float)n, h->dOut.p, h->dOutT.p, ldT(n), h->loss.p);
    check(hipGetLastError(), "mlp_loss launch");
    backward(h, n);
    float l = 0;
    download(h, &l, h->loss.p, 4);
    return l;
}

// a caller's pointer that must not be null
template <class T>
T *need(T *p, const char *what)
{
    if (!p)
        throw std::invalid_argument(std::string("null ") + what);
    return p;
}

#define MLP_TRY try {
#define MLP_CATCH                        \
    }                                    \
    catch (const std::invalid_argument &e) \
    {                                    \
        g_err = e.what();                \
        return CLODE_MLP_ERR_INVALID;    \
    }                                    \
    catch (const std::exception &e)      \
    {                                    \
        g_err = e.what();                \
        return CLODE_MLP_ERR_RUNTIME;    \
    }                                    \
    return CLODE_MLP_OK;

#define CHECK_H(h)                        \
    if (!(h)) {                           \
        g_err = "null surrogate handle";  \
        return CLODE_MLP_ERR_INVALID;     \
    }

// ---- screening steps shared by the allocating and the no-allocation entry points

void upload_target_scaling(clode_mlp *h)
{
    float tm[4] = {h->tmean[0], h->tmean[1], h->tstd[0], h->tstd[1]};
    upload(h, h->dtm.p, tm, 16);
}

// the fused screening kernel over the nx x ny grid into device outputs
void screen_fused(clode_mlp *h, int nx, int ny, float tol, float *P, float *A, int *C, uint8_t *F)
{
    const long long ncell = (long long)nx * ny;
    ScreenArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ncell = ncell;
    a.nx = nx;
    a.ny = ny;
    a.Bff = h->dBff.p;
    for (int l = 0; l < NLAYER; ++l) {
        a.W[l] = h->pbf.p + h->L[l].woff;
        a.bias[l] = h->p.p + h->L[l].boff;
        if (((uintptr_t)a.W[l] & 15) != 0)  // the kernel reads 16-byte weight fragments
            throw MlpError("screening weights are not 16-byte aligned");
    }
    a.tm = h->dtm.p;
    a.tol = tol;
    a.period = P;
    a.amp = A;
    a.cls = C;
    a.flags = F;
    const unsigned int nblk = (unsigned int)((ncell + SROWS - 1) / SROWS);
    hipLaunchKernelGGL(mlp_screen_fused, dim3(nblk), dim3(256), 0, h->stream, a);
    check(hipGetLastError(), "mlp_screen_fused launch");
}

// flags bit 1 (class boundary) into F, then `dilate` rounds of dilation
// between F and F2 (two buffers of nx * ny bytes); returns the buffer that
// holds the final flags
const uint8_t *screen_flags(clode_mlp *h, int nx, int ny, int dilate, const int *C, uint8_t *F, uint8_t *F2)
{
    const long long ncell = (long long)nx * ny;
    const unsigned int gb = (unsigned int)((ncell + 255) / 256);
    hipLaunchKernelGGL(mlp_boundary, dim3(gb), dim3(256), 0, h->stream, C, nx, ny, F);
    check(hipGetLastError(), "mlp_boundary launch");
    for (int r = 0; r < dilate; ++r) {
        hipLaunchKernelGGL(mlp_dilate, dim3(gb), dim3(256), 0, h->stream, F, nx, ny, F2);
        check(hipGetLastError(), "mlp_dilate launch");
        std::swap(F, F2);
    }
    return F;
}

// a caller's output pointer that must be device memory of the handle's device
template <class T>
void need_device(clode_mlp *h, T *p, const char *what)
{
    need(p, what);
    hipPointerAttribute_t at;
    std::memset(&at, 0, sizeof(at));
    const hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // leave no error behind for the next HIP call
        throw std::invalid_argument(std::string(what) + " is not device memory");
    }
    if ((at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != h->device)
        throw std::invalid_argument(std::string(what) + " is not device memory of device " +
                                    std::to_string(h->device));
}

}  // namespace

extern "C" {

const char *clode_mlp_last_error(void) { return g_err.c_str(); }

int clode_mlp_create(clode_mlp **out, int device, unsigned long long seed)
{
    if (!out) {
        g_err = "null output handle";
        return CLODE_MLP_ERR_INVALID;
    }
    *out = nullptr;
    MLP_TRY
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0)
        throw MlpError("no HIP device is available");
    if (device < 0 || device >= nd)
        throw std::invalid_argument("device ordinal out of range");
    check(hipSetDevice(device), "hipSetDevice");
    auto *h = new clode_mlp;
    try {
        h->device = device;
        check(hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
        check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&h->ev0), "hipEventCreate");
        check(hipEventCreate(&h->ev1), "hipEventCreate");
        size_t off = 0;
        int in = D0;
        for (int l = 0; l < NLAYER; ++l) {
            int o = l < NHID ? H : NOUTP;
            h->L[l] = Layer{in, o, off, off + (size_t)in * o};
            off += (size_t)in * o + o;
            in = o;
        }
        h->nparam = off;
        // He-style initialisation for SiLU layers, small heads; Fourier frequencies ~ N(0, 3^2)
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd01(0.f, 1.f);
        std::vector<float> p(h->nparam, 0.f);
        std::vector<uint8_t> mask(h->nparam, 1);
        for (int l = 0; l < NLAYER; ++l) {
            const Layer &Ly = h->L[l];
            float sd = l < NHID ? std::sqrt(2.f / Ly.in) : std::sqrt(1.f / Ly.in);
            for (int o = 0; o < Ly.out; ++o)
                for (int i = 0; i < Ly.in; ++i) {
                    size_t k = Ly.woff + (size_t)o * Ly.in + i;
                    if (l == NHID && o >= NOUT) {
                        mask[k] = 0;  // padding columns stay zero
                        continue;
                    }
                    p[k] = sd * nd01(rng);
                }
            if (l == NHID)
                for (int o = NOUT; o < NOUTP; ++o)
                    mask[Ly.boff + o] = 0;
        }
        h->Bff.resize(2 * NFF);
        for (auto &b : h->Bff)
            b = 3.f * nd01(rng);
        h->p.alloc(h->nparam);
        h->g.alloc(h->nparam);
        h->m1.alloc(h->nparam);
        h->m2.alloc(h->nparam);
        h->mask.alloc(h->nparam);
        h->pbf.alloc(h->nparam);
        h->pbfT.alloc(h->nparam);
        h->wsp.alloc((size_t)KSPLIT_MAX * H * H);  // the largest weight gradient, per split
        h->wall.alloc((size_t)KSPLIT_MAX * h->nparam);  // every layer's, per split (30 MB)
        h->dBff.alloc(2 * NFF);
        h->dtm.alloc(4);
        h->loss.alloc(1);
        h->stage = scratch_address(device);  // device is current (hipSetDevice above)
        upload(h, h->p.p, p.data(), p.size() * 4);
        upload(h, h->mask.p, mask.data(), mask.size());
        upload(h, h->dBff.p, h->Bff.data(), h->Bff.size() * 4);
        check(hipMemsetAsync(h->g.p, 0, h->nparam * 4, h->stream), "hipMemset");
        check(hipMemsetAsync(h->m1.p, 0, h->nparam * 4, h->stream), "hipMemset");
        check(hipMemsetAsync(h->m2.p, 0, h->nparam * 4, h->stream), "hipMemset");
        refresh_bf16(h);
    } catch (...) {
        clode_mlp_destroy(h);
        throw;
    }
    *out = h;
    MLP_CATCH
}

int clode_mlp_destroy(clode_mlp *h)
{
    if (!h)
        return CLODE_MLP_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
    }
    if (h->ev0)
        (void)hipEventDestroy(h->ev0);
    if (h->ev1)
        (void)hipEventDestroy(h->ev1);
    delete h;
    return CLODE_MLP_OK;
}

int clode_mlp_n_params(clode_mlp *h, size_t *n)
{
    CHECK_H(h) MLP_TRY *need(n, "n") = h->nparam;
    MLP_CATCH
}

int clode_mlp_get_params(clode_mlp *h, float *out, size_t n)
{
    CHECK_H(h) MLP_TRY
    if (n != h->nparam)
        throw std::invalid_argument("parameter vector has the wrong length");
    need(out, "output buffer");
    check(hipSetDevice(h->device), "hipSetDevice");
    download(h, out, h->p.p, n * 4);
    MLP_CATCH
}

int clode_mlp_set_params(clode_mlp *h, const float *in, size_t n)
{
    CHECK_H(h) MLP_TRY
    if (n != h->nparam)
        throw std::invalid_argument("parameter vector has the wrong length");
    need(in, "parameter vector");
    check(hipSetDevice(h->device), "hipSetDevice");
    upload(h, h->p.p, in, n * 4);
    refresh_bf16(h);
    MLP_CATCH
}

int clode_mlp_get_fourier(clode_mlp *h, float *out, size_t n)
{
    CHECK_H(h) MLP_TRY
    if (n != 2 * NFF)
        throw std::invalid_argument("Fourier matrix has 32 entries");
    std::copy(h->Bff.begin(), h->Bff.end(), need(out, "output buffer"));
    MLP_CATCH
}

int clode_mlp_set_target_scaling(clode_mlp *h, const float *mean2, const float *std2)
{
    CHECK_H(h) MLP_TRY
    need(mean2, "mean2");
    need(std2, "std2");
    for (int k = 0; k < 2; ++k) {
        if (!(std2[k] > 0))
            throw std::invalid_argument("target scales must be positive");
        h->tmean[k] = mean2[k];
        h->tstd[k] = std2[k];
    }
    MLP_CATCH
}

int clode_mlp_forward(clode_mlp *h, const float *X, size_t n, float *out, unsigned short *encoded)
{
    CHECK_H(h) MLP_TRY
    if (n == 0 || n > (1u << 18))
        throw std::invalid_argument("forward takes 1 to 262144 rows");
    need(X, "input rows");
    need(out, "output buffer");
    check(hipSetDevice(h->device), "hipSetDevice");
    h->ensureRows(n);
    upload(h, h->X.p, X, n * 8);
    encode(h, h->X.p, (int)n);
    forward(h, (int)n);
    std::vector<float> z(n * NOUTP);
    download(h, z.data(), h->Zh.p, z.size() * 4);
    for (size_t i = 0; i < n; ++i)
        std::copy(z.begin() + i * NOUTP, z.begin() + i * NOUTP + NOUT, out + i * NOUT);
    if (encoded)
        download(h, encoded, h->A0.p, n * D0 * 2);
    MLP_CATCH
}

int clode_mlp_loss_grad(clode_mlp *h, const float *X, const float *T, const int *cls, size_t n, float *loss,
                        float *grad)
{
    CHECK_H(h) MLP_TRY
    if (n == 0 || n > (1u << 16))
        throw std::invalid_argument("loss_grad takes 1 to 65536 rows");
    need(X, "input rows");
    need(T, "targets");
    need(cls, "class labels");
    need(loss, "loss output");
    for (size_t i = 0; i < n; ++i)
        if (cls[i] < 0 || cls[i] >= NCLS)
            throw std::invalid_argument("class labels must be in 0..15");
    check(hipSetDevice(h->device), "hipSetDevice");
    h->ensureRows(n);
    upload(h, h->X.p, X, n * 8);
    upload(h, h->T.p, T, n * 8);
    upload(h, h->C.p, cls, n * 4);
    std::vector<unsigned long long> ids(n);
    for (size_t i = 0; i < n; ++i)
        ids[i] = i;
    upload(h, h->ids.p, ids.data(), n * 8);
    *loss = loss_grad(h, (int)n);
    if (grad)
        download(h, grad, h->g.p, h->nparam * 4);
    MLP_CATCH
}

int clode_mlp_train(clode_mlp *h, const float *X, const float *T, const int *cls, size_t n, int epochs, int batch,
                    float lr, unsigned long long seed, float *epoch_loss)
{
    CHECK_H(h) MLP_TRY
    if (n < 1 || n > (1u << 24) || epochs < 1 || batch < 64 || batch > 65536)
        throw std::invalid_argument("train: 1..16777216 samples, epochs >= 1, batch 64..65536");
    need(X, "input rows");
    need(T, "targets");
    need(cls, "class labels");
    for (size_t i = 0; i < n; ++i)
        if (cls[i] < 0 || cls[i] >= NCLS)
            throw std::invalid_argument("class labels must be in 0..15");
    check(hipSetDevice(h->device), "hipSetDevice");
    // the training set stays on the device; each step gathers a shuffled
    // minibatch into the activation buffers
    Dev<float> Xall, Tall, eloss;
    Dev<int> Call;
    Dev<unsigned int> dperm;
    Xall.alloc(n * 2);
    Tall.alloc(n * 2);
    Call.alloc(n);
    dperm.alloc(n);
    eloss.alloc((size_t)epochs);
    upload(h, Xall.p, X, n * 8);
    upload(h, Tall.p, T, n * 8);
    upload(h, Call.p, cls, n * 4);
    h->ensureRows((size_t)batch);
    std::mt19937_64 rng(seed);
    // epoch e's order is a shuffle of epoch e-1's; the next epoch's shuffle
    // runs on the host while the GPU trains on the current one.  Every
    // epoch's order has its own host buffer, untouched after its upload
    // is queued, and the epoch losses stay on the device until the end, so
    // the host waits between epochs only for an upload two epochs back.
    std::vector<unsigned int> perms[2] = {std::vector<unsigned int>(n), std::vector<unsigned int>(n)};
    hipEvent_t uploaded[2];
    for (int b = 0; b < 2; ++b)
        check(hipEventCreateWithFlags(&uploaded[b], hipEventDisableTiming), "hipEventCreate");
    for (size_t i = 0; i < n; ++i)
        perms[0][i] = (unsigned int)i;
    std::shuffle(perms[0].begin(), perms[0].end(), rng);
    const size_t steps_per_epoch = (n + batch - 1) / batch;
    const double total = (double)steps_per_epoch * epochs;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    LayerTable lt;
    for (int l = 0; l < NLAYER; ++l) {
        lt.woff[l] = (long long)h->L[l].woff;
        lt.in[l] = h->L[l].in;
        lt.out[l] = h->L[l].out;
    }
    // by default each layer's weight and input gradients share a launch and
    // the gradient sums run inside the optimizer launch (13 launches per
    // step); CLODE_MLP_FUSED_TRAIN=0 launches every product on its own and
    // sums in the backward pass (22 launches, same parameters bit for bit)
    const char *fa = std::getenv("CLODE_MLP_FUSED_TRAIN");
    const bool fused = !(fa && std::strcmp(fa, "0") == 0);
    int nbias = 0;
    for (int l = 0; l < NLAYER; ++l)
        nbias += h->L[l].out;
    long long step = 0;
    check(hipMemsetAsync(eloss.p, 0, (size_t)epochs * 4, h->stream), "hipMemset");
    check(hipEventRecord(h->ev0, h->stream), "hipEventRecord");
    for (int e = 0; e < epochs; ++e) {
        upload(h, dperm.p, perms[e % 2].data(), n * 4);
        check(hipEventRecord(uploaded[e % 2], h->stream), "hipEventRecord");
        for (size_t s = 0; s < steps_per_epoch; ++s, ++step) {
            const size_t b0 = s * (size_t)batch;
            const int nb = (int)std::min<size_t>((size_t)batch, n - b0);
            hipLaunchKernelGGL(mlp_gather, dim3((nb * NFF + 255) / 256), dim3(256), 0, h->stream, dperm.p + b0, nb,
                               Xall.p, Tall.p, Call.p, h->T.p, h->C.p, h->ids.p, h->dBff.p, h->A0.p, h->A0t.p,
                               ldT(nb));
            check(hipGetLastError(), "mlp_gather launch");
            forward(h, nb, true);
            hipLaunchKernelGGL(mlp_loss, dim3((NHEAD * nb + 255) / 256), dim3(256), 0, h->stream, h->Zh.p, nb, h->T.p,
                               h->C.p, h->ids.p, 1.0f / (float)nb, h->dOut.p, h->dOutT.p, ldT(nb), eloss.p + e);
            check(hipGetLastError(), "mlp_loss launch");
            GradSources gs;
            std::memset(&gs, 0, sizeof(gs));
            gs.part = h->wall.p;
            gs.pstride = (long long)h->nparam;
            gs.n = nb;
            gs.ld = ldT(nb);
            backward(h, nb, fused ? &gs : nullptr);
            // cosine decay to 5 % of the initial rate
            const double frac = (double)step / total;
            AdamStep as;
            as.lr = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac))));
            as.b1 = b1;
            as.b2 = b2;
            as.eps = eps;
            ++h->adamStep;
            as.c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)h->adamStep)));
            as.c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)h->adamStep)));
            const int wblocks = (int)((h->nparam + 255) / 256);
            if (fused)
                hipLaunchKernelGGL(mlp_adam_sum, dim3((unsigned int)(wblocks + nbias)), dim3(256), 0, h->stream,
                                   h->p.p, h->m1.p, h->m2.p, h->mask.p, h->pbf.p, h->pbfT.p, lt,
                                   (long long)h->nparam, wblocks, gs, as);
            else
                hipLaunchKernelGGL(mlp_adam, dim3((unsigned int)wblocks), dim3(256), 0, h->stream, h->p.p, h->g.p,
                                   h->m1.p, h->m2.p, h->mask.p, h->pbf.p, h->pbfT.p, lt, (long long)h->nparam, as);
            check(hipGetLastError(), "mlp_adam launch");
        }
        if (e + 1 < epochs) {
            std::vector<unsigned int> &nx = perms[(e + 1) % 2];
            if (e >= 1)  // epoch e-1's upload read this buffer
                check(hipEventSynchronize(uploaded[(e + 1) % 2]), "hipEventSynchronize");
            nx = perms[e % 2];
            std::shuffle(nx.begin(), nx.end(), rng);
        }
    }
    for (int b = 0; b < 2; ++b)
        (void)hipEventDestroy(uploaded[b]);
    std::vector<float> el((size_t)epochs);
    download(h, el.data(), eloss.p, (size_t)epochs * 4);
    if (epoch_loss)
        for (int e = 0; e < epochs; ++e)
            epoch_loss[e] = el[(size_t)e] / (float)steps_per_epoch;
    check(hipEventRecord(h->ev1, h->stream), "hipEventRecord");
    check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->lastMs = ms;
    MLP_CATCH
}

int clode_mlp_screen_grid(clode_mlp *h, int nx, int ny, float tol, int dilate, float *period, float *amp, int *cls,
                          unsigned char *flags)
{
    CHECK_H(h) MLP_TRY
    if (nx < 1 || ny < 1 || (long long)nx * ny > (1ll << 26))
        throw std::invalid_argument("grid must have 1 to 2^26 cells");
    if (dilate < 0 || dilate > 64)
        throw std::invalid_argument("dilation radius must be 0..64");
    need(period, "period output");
    need(amp, "amplitude output");
    need(cls, "class output");
    need(flags, "flags output");
    check(hipSetDevice(h->device), "hipSetDevice");
    const long long ncell = (long long)nx * ny;
    // the fused kernel by default; CLODE_MLP_FUSED_SCREEN=0 runs the network
    // layer by layer over chunks of 2^18 cells (same results, bit for bit)
    const char *fe = std::getenv("CLODE_MLP_FUSED_SCREEN");
    const bool fused = !(fe && std::strcmp(fe, "0") == 0);
    const int chunk = (int)std::min<long long>(ncell, 1 << 18);
    if (!fused)
        h->ensureRows((size_t)chunk);
    Dev<float> dP, dA;
    Dev<int> dC;
    Dev<uint8_t> dF, dF2;
    dP.alloc(ncell);
    dA.alloc(ncell);
    dC.alloc(ncell);
    dF.alloc(ncell);
    dF2.alloc(ncell);
    upload_target_scaling(h);
    check(hipEventRecord(h->ev0, h->stream), "hipEventRecord");
    if (fused)
        screen_fused(h, nx, ny, tol, dP.p, dA.p, dC.p, dF.p);
    for (long long c0 = 0; !fused && c0 < ncell; c0 += chunk) {
        const int nb = (int)std::min<long long>(chunk, ncell - c0);
        hipLaunchKernelGGL(mlp_grid_inputs, dim3((nb + 255) / 256), dim3(256), 0, h->stream, c0, nb, nx, ny, h->X.p);
        check(hipGetLastError(), "mlp_grid_inputs launch");
        encode(h, h->X.p, nb);
        forward(h, nb);
        hipLaunchKernelGGL(mlp_predict_cells, dim3((nb + 255) / 256), dim3(256), 0, h->stream, h->Zh.p, nb, h->dtm.p,
                           h->dtm.p + 2, tol, dP.p + c0, dA.p + c0, dC.p + c0, dF.p + c0);
        check(hipGetLastError(), "mlp_predict_cells launch");
    }
    const uint8_t *res = screen_flags(h, nx, ny, dilate, dC.p, dF.p, dF2.p);
    check(hipEventRecord(h->ev1, h->stream), "hipEventRecord");
    download(h, period, dP.p, ncell * 4);
    download(h, amp, dA.p, ncell * 4);
    download(h, cls, dC.p, ncell * 4);
    download(h, flags, res, ncell);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->lastMs = ms;
    MLP_CATCH
}

int clode_mlp_screen_grid_noalloc(clode_mlp *h, int nx, int ny, float tol, int dilate, float *d_period, float *d_amp,
                                  int *d_cls, unsigned char *d_flags)
{
    CHECK_H(h) MLP_TRY
    if (nx < 1 || ny < 1)
        throw std::invalid_argument("grid must have at least one cell");
    const long long ncell = (long long)nx * ny;
    if (ncell > CLODE_MLP_NOALLOC_MAX_CELLS)
        throw std::invalid_argument("grid of " + std::to_string(ncell) + " cells exceeds the " +
                                    std::to_string(CLODE_MLP_NOALLOC_MAX_CELLS) +
                                    "-cell limit of the no-allocation screening (its static staging buffer)");
    if (dilate < 0 || dilate > 64)
        throw std::invalid_argument("dilation radius must be 0..64");
    check(hipSetDevice(h->device), "hipSetDevice");
    need_device(h, d_period, "period output");
    need_device(h, d_amp, "amplitude output");
    need_device(h, d_cls, "class output");
    need_device(h, d_flags, "flags output");
    if (!h->stage)
        throw MlpError("the static scratch of this device has no address");
    upload_target_scaling(h);
    check(hipEventRecord(h->ev0, h->stream), "hipEventRecord");
    screen_fused(h, nx, ny, tol, d_period, d_amp, d_cls, d_flags);
    const uint8_t *res = screen_flags(h, nx, ny, dilate, d_cls, d_flags, h->stage);
    if (res != d_flags)  // an odd number of dilation rounds ends in the staging buffer
        check(hipMemcpyAsync(d_flags, res, (size_t)ncell, hipMemcpyDeviceToDevice, h->stream), "hipMemcpy D2D");
    check(hipEventRecord(h->ev1, h->stream), "hipEventRecord");
    check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->lastMs = ms;
    MLP_CATCH
}

int clode_mlp_last_ms(clode_mlp *h, double *ms)
{
    CHECK_H(h) MLP_TRY *need(ms, "ms") = h->lastMs;
    MLP_CATCH
}

}  // extern "C"
