// Surrogate MLP on the CDNA4 matrix cores: forward, masked Gaussian NLL,
// backward and AdamW, bf16 operands with fp32 accumulation.
//
// Matrix products use one LDS-tiled kernel: a 256-thread block computes a
// 64 x 64 tile of C = A * B^T, each 64-wide wave a 32 x 32 quarter with
// v_mfma_f32_32x32x16_bf16.  Operands are staged through LDS as [row][k]
// tiles of 32 k; an operand whose reduction index is not the contiguous one
// (the transposed products of the backward pass) is loaded with 16-byte
// coalesced reads along its contiguous index and transposed on the LDS store,
// so no transposed copies of weights or activations are ever written.
// Epilogues are fused: bias + GELU (keeping the pre-activation), bias only
// (the head, fp32 out), GELU' times the incoming gradient, or a raw fp32
// split-K partial of a weight gradient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "xmcml_surrogate.h"

namespace xmc {
int set_error(int code, const std::string& msg);
}

namespace xmlp {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x16 __attribute__((__vector_size__(16 * sizeof(float))));
typedef uint32_t u32x4 __attribute__((__vector_size__(16)));
typedef float f32x4 __attribute__((__vector_size__(16)));

constexpr int BM = 64, BN = 64, BK = 64;
constexpr int kAnswerRows = 65536;  // rows per chunk of xmc_mlp_answer
constexpr int LDS_STRIDE = BK + 8;  // halfwords per LDS tile row (144 bytes: 16-byte aligned rows)
constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kGeluA = 0.044715f;
constexpr float kLogVarClamp = 10.0f;

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even (finite inputs)
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// bf16 rounding on the device: v_cvt_pk_bf16_f32, round to nearest even,
// two values per instruction (the same bits as f2bf for finite inputs)
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ inline uint32_t bf16x2_rne(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

__device__ inline uint16_t bf16_rne(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// GELU, tanh form: 0.5 z (1 + tanh y) with y = sqrt(2/pi) (z + 0.044715 z^3),
// evaluated as z / (1 + e^{-2y}) with e^{-2y} = 2^{z (g1 + g3 z^2)} on the
// hardware exponential and reciprocal (v_exp_f32, v_rcp_f32, about 1 ulp
// each).  The epilogues run once per output element: five VALU operations
// and two transcendentals here, against a product of hidden-width length.
constexpr float kGelu1 = -2.0f * kGeluC * 1.4426950408889634f;  // -2 sqrt(2/pi) log2(e)
constexpr float kGelu3 = kGelu1 * kGeluA;

__device__ inline float gelu_sigmoid(float z) {  // 1 / (1 + e^{-2y}) = 0.5 (1 + tanh y)
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(z * z, kGelu3, kGelu1)));
}

__device__ inline float gelu(float z) { return z * gelu_sigmoid(z); }

// d/dz of z s(z), s = 0.5 (1 + tanh y): s + z s (1 - s) 2 y'(z)
__device__ inline float gelu_grad(float z) {
    float s = gelu_sigmoid(z);
    return s + z * s * (1.0f - s) * (2.0f * kGeluC) * (1.0f + 3.0f * kGeluA * z * z);
}

enum Epi { EPI_GELU = 0, EPI_LINEAR = 1, EPI_DGELU = 2, EPI_PARTIAL = 3 };

// C[m][n] = sum_k A(m,k) B(k,n) over k in this block's split-K chunk.
// A(m,k) = A[m*lda + k] if AK else A[k*lda + m];
// B(k,n) = B[n*ldb + k] if BKC else B[k*ldb + n].
struct GemmArgs {
    const uint16_t* A;
    long lda;
    const uint16_t* B;
    long ldb;
    int M, N, kchunk;
    const float* bias;   // EPI_GELU / EPI_LINEAR: [N]
    const uint16_t* zin; // EPI_DGELU: pre-activation Z[m*ldo + n] of the layer the gradient enters
    uint16_t* out;       // EPI_GELU: activation; EPI_DGELU: gradient; [m*ldo + n]
    uint16_t* zout;      // EPI_GELU: pre-activation [m*ldo + n]
    long ldo;
    float* outf;         // EPI_LINEAR: [m*ldo + n]; EPI_PARTIAL: outf + z*M*ldo + m*ldo + n
};

// A 64 x 64 (rows x k) operand tile is 512 16-byte pieces: two per thread.
// AK: piece = (row, 8 consecutive k); otherwise piece = (k, 8 consecutive
// rows), transposed on the LDS store.
template <bool AK>
__device__ inline void load_pieces(u32x4 (&v)[2], const uint16_t* src, long ld, int r0, int k0, int t) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int p = t + 256 * j;
        int a = p >> 3, c = (p & 7) * 8;
        v[j] = AK ? *(const u32x4*)(src + (long)(r0 + a) * ld + k0 + c)
                  : *(const u32x4*)(src + (long)(k0 + a) * ld + r0 + c);
    }
}

template <bool AK>
__device__ inline void store_pieces(uint16_t* lds, const u32x4 (&v)[2], int t) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int p = t + 256 * j;
        int a = p >> 3, c = (p & 7) * 8;
        if constexpr (AK) {
            *(u32x4*)(lds + a * LDS_STRIDE + c) = v[j];
        } else {
            uint16_t h[8];
            std::memcpy(h, &v[j], 16);
#pragma unroll
            for (int i = 0; i < 8; i++) lds[(c + i) * LDS_STRIDE + a] = h[i];
        }
    }
}

// The accumulator tile of one wave (32 x 32, v_mfma_f32_32x32x16_bf16 layout:
// element v of a lane is row 8*(v/4) + 4*(lane/32) + v%4, column lane%32)
// through the epilogue, at rows m_base.., columns n_base..
// Accumulators start at the bias (EPI_GELU, EPI_LINEAR: the bias is the first
// term of every output's sum) or at zero.  Column n_base + lane % 32 in the
// v_mfma_f32_32x32x16_bf16 layout.
template <int EPI>
__device__ __forceinline__ float acc_start(const GemmArgs& g, int n_base, int lane) {
    if constexpr (EPI == EPI_GELU || EPI == EPI_LINEAR) return g.bias[n_base + (lane & 31)];
    else return 0.0f;
}

// The accumulator tile of one wave (32 x 32, v_mfma_f32_32x32x16_bf16 layout:
// element v of a lane is row 8*(v/4) + 4*(lane/32) + v%4, column lane%32)
// through the epilogue, at rows m_base.., columns n_base..
template <int EPI>
__device__ __forceinline__ void epilogue(const GemmArgs& g, const f32x16& acc, int m_base, int n_base, int lane) {
    const int n = n_base + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int m = m_base + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const long o = (long)m * g.ldo + n;
        float x = acc[v];
        if constexpr (EPI == EPI_GELU) {
            g.zout[o] = f2bf(x);
            g.out[o] = bf16_rne(gelu(x));
        } else if constexpr (EPI == EPI_LINEAR) {
            g.outf[o] = x;
        } else if constexpr (EPI == EPI_DGELU) {
            g.out[o] = f2bf(x * gelu_grad(bf2f(g.zin[o])));
        } else {
            g.outf[(long)blockIdx.z * g.M * g.ldo + o] = x;
        }
    }
}

template <bool AK, bool BKC, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    // two LDS stages: the tile being multiplied and the next one being stored
    __shared__ __attribute__((aligned(16))) uint16_t As[2][BM * LDS_STRIDE];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN * LDS_STRIDE];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kb = blockIdx.z * g.kchunk, ke = kb + g.kchunk;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int lr = lane & 31, lk = (lane >> 5) * 8;
    f32x16 acc;
    const float a0 = acc_start<EPI>(g, n0 + wn, lane);
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = a0;

    u32x4 ra[2], rb[2];
    load_pieces<AK>(ra, g.A, g.lda, m0, kb, t);
    load_pieces<BKC>(rb, g.B, g.ldb, n0, kb, t);
    store_pieces<AK>(As[0], ra, t);
    store_pieces<BKC>(Bs[0], rb, t);
    __syncthreads();
    int stage = 0;
    for (int k0 = kb; k0 < ke; k0 += BK) {
        const bool more = k0 + BK < ke;
        if (more) {  // next tile's global reads in flight during this tile's products
            load_pieces<AK>(ra, g.A, g.lda, m0, k0 + BK, t);
            load_pieces<BKC>(rb, g.B, g.ldb, n0, k0 + BK, t);
        }
        const uint16_t* as = As[stage];
        const uint16_t* bs = Bs[stage];
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            // lane holds row (lane % 32), k = kk + 8*(lane / 32) .. +7 of each operand
            u32x4 av = *(const u32x4*)(as + (wm + lr) * LDS_STRIDE + kk + lk);
            u32x4 bv = *(const u32x4*)(bs + (wn + lr) * LDS_STRIDE + kk + lk);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                          __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
        }
        if (more) {
            store_pieces<AK>(As[stage ^ 1], ra, t);
            store_pieces<BKC>(Bs[stage ^ 1], rb, t);
        }
        __syncthreads();
        stage ^= 1;
    }

    epilogue<EPI>(g, acc, m0 + wm, n0 + wn, lane);
}

// 128 x 128 tiles for the large products (both dimensions multiples of 128):
// each wave computes a 64 x 64 quarter as 2 x 2 MFMA tiles, so every
// operand fragment read from LDS feeds two MFMAs.  32-deep k stages, two LDS
// buffers (41 KB), next stage's global reads in flight during the products.
constexpr int BM2 = 128, BN2 = 128, BK2 = 32, LDS_STRIDE2 = BK2 + 8;

template <bool AK>
__device__ inline void load_pieces2(u32x4 (&v)[2], const uint16_t* src, long ld, int r0, int k0, int t) {
    // 128 rows x 32 k = 512 pieces of 8: two per thread
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int p = t + 256 * j;
        if constexpr (AK) {
            int a = p >> 2, c = (p & 3) * 8;
            v[j] = *(const u32x4*)(src + (long)(r0 + a) * ld + k0 + c);
        } else {
            int a = p >> 4, c = (p & 15) * 8;  // 32 k-rows of 16 pieces along the rows
            v[j] = *(const u32x4*)(src + (long)(k0 + a) * ld + r0 + c);
        }
    }
}

template <bool AK>
__device__ inline void store_pieces2(uint16_t* lds, const u32x4 (&v)[2], int t) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int p = t + 256 * j;
        if constexpr (AK) {
            int a = p >> 2, c = (p & 3) * 8;
            *(u32x4*)(lds + a * LDS_STRIDE2 + c) = v[j];
        } else {
            int a = p >> 4, c = (p & 15) * 8;
            uint16_t h[8];
            std::memcpy(h, &v[j], 16);
#pragma unroll
            for (int i = 0; i < 8; i++) lds[(c + i) * LDS_STRIDE2 + a] = h[i];
        }
    }
}

template <bool AK, bool BKC, int EPI>
__global__ __launch_bounds__(256) void gemm128_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[2][BM2 * LDS_STRIDE2];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN2 * LDS_STRIDE2];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * BM2, n0 = blockIdx.x * BN2;
    const int kb = blockIdx.z * g.kchunk, ke = kb + g.kchunk;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
    const int lr = lane & 31, lk = (lane >> 5) * 8;
    f32x16 acc[2][2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const float a0 = acc_start<EPI>(g, n0 + wn + 32 * j, lane);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[i][j][v] = a0;
    }

    u32x4 ra[2], rb[2];
    load_pieces2<AK>(ra, g.A, g.lda, m0, kb, t);
    load_pieces2<BKC>(rb, g.B, g.ldb, n0, kb, t);
    store_pieces2<AK>(As[0], ra, t);
    store_pieces2<BKC>(Bs[0], rb, t);
    __syncthreads();
    int stage = 0;
    for (int k0 = kb; k0 < ke; k0 += BK2) {
        const bool more = k0 + BK2 < ke;
        if (more) {
            load_pieces2<AK>(ra, g.A, g.lda, m0, k0 + BK2, t);
            load_pieces2<BKC>(rb, g.B, g.ldb, n0, k0 + BK2, t);
        }
        const uint16_t* as = As[stage];
        const uint16_t* bs = Bs[stage];
#pragma unroll
        for (int kk = 0; kk < BK2; kk += 16) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                a[i] = __builtin_bit_cast(bf16x8, *(const u32x4*)(as + (wm + 32 * i + lr) * LDS_STRIDE2 + kk + lk));
                b[i] = __builtin_bit_cast(bf16x8, *(const u32x4*)(bs + (wn + 32 * i + lr) * LDS_STRIDE2 + kk + lk));
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            store_pieces2<AK>(As[stage ^ 1], ra, t);
            store_pieces2<BKC>(Bs[stage ^ 1], rb, t);
        }
        __syncthreads();
        stage ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) epilogue<EPI>(g, acc[i][j], m0 + wm + 32 * i, n0 + wn + 32 * j, lane);
}

// x[b][0..nin) fp32 -> bf16 rows of width ldx, zero padding beyond nin and batch
__global__ void pack_input_kernel(const float* x, int batch, int nin, uint16_t* a0, int ldx, int rows) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)rows * ldx) return;
    int r = (int)(i / ldx), c = (int)(i % ldx);
    a0[i] = (r < batch && c < nin) ? bf16_rne(x[(long)r * nin + c]) : (uint16_t)0;
}

// masked Gaussian NLL of the head output; writes the bf16 gradient of the
// head (zero for padded rows) and one partial loss sum per block
__global__ void loss_kernel(const float* out, const float* y, const float* mask, int batch, int nout, int rows,
                            float scale, uint16_t* dout, float* partial) {
    __shared__ float red[256];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < (long)rows * nout) {
        int r = (int)(i / nout), j = (int)(i % nout);
        long oi = (long)r * 2 * nout;
        float dmu = 0.0f, ds = 0.0f;
        if (r < batch) {
            float mk = mask[i];
            float mu = out[oi + j], s = out[oi + nout + j];
            float sc = fminf(fmaxf(s, -kLogVarClamp), kLogVarClamp);
            float e = expf(-sc), d = y[i] - mu;
            float ed2 = e * d * d;
            l = mk * 0.5f * (ed2 + sc);
            dmu = -scale * mk * e * d;
            ds = (s >= -kLogVarClamp && s <= kLogVarClamp) ? scale * mk * 0.5f * (1.0f - ed2) : 0.0f;
        }
        dout[oi + j] = f2bf(dmu);
        dout[oi + nout + j] = f2bf(ds);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// part[z][n] = sum of src[r][n] (bf16 in, fp32 out) over the rows of slice
// z = blockIdx.y, in a fixed order: 64 columns per block, four row lanes
// (coalesced 128-byte rows), then a fixed LDS combine
constexpr int kColsumSlices = 32;
__global__ void colsum_kernel(const uint16_t* src, int rows, int cols, float* part) {
    __shared__ float red[4][64];
    const int c = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    const int per = (rows + gridDim.y - 1) / gridDim.y;
    const int r0 = blockIdx.y * per, r1 = min(rows, r0 + per);
    float s = 0.0f;
    if (n < cols)
        for (int r = r0 + ty; r < r1; r += 4) s += bf2f(src[(long)r * cols + n]);
    red[ty][c] = s;
    __syncthreads();
    if (ty == 0 && n < cols)
        part[(long)blockIdx.y * cols + n] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// g[i] = sum_z p[z*count + i], fixed order
__global__ void reduce_partials_kernel(const float* p, int splits, long count, float* g) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int z = 0; z < splits; z++) s += p[z * count + i];
    g[i] = s;
}

__global__ void adam_kernel(float* w, const float* g, float* m1, float* m2, uint16_t* mirror, long count,
                            float lr, float b1, float b2, float eps, float c1, float c2, float decay) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float gi = g[i];
    float a = b1 * m1[i] + (1.0f - b1) * gi;
    float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    float x = w[i];
    x -= lr * (a / c1) / (sqrtf(b / c2) + eps) + lr * decay * x;
    w[i] = x;
    if (mirror) mirror[i] = f2bf(x);
}

__device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// batch row r <- data row splitmix64(key + r) mod rows; one thread per
// copied value, so consecutive lanes move consecutive columns of a row
__global__ void gather_kernel(const float* X, const float* Y, const float* M, long rows, int nin, int nout,
                              int batch, uint64_t key, float* xb, float* yb, float* mb) {
    const int width = nin + 2 * nout;
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)batch * width) return;
    int r = (int)(e / width), c = (int)(e % width);
    long i = (long)(splitmix64(key + (uint64_t)r) % (uint64_t)rows);
    if (c < nin) xb[(long)r * nin + c] = X[i * nin + c];
    else if (c < nin + nout) yb[(long)r * nout + (c - nin)] = Y[i * nout + (c - nin)];
    else mb[(long)r * nout + (c - nin - nout)] = M[i * nout + (c - nin - nout)];
}

// Feature rows of layered configurations: layer tuples (n, mua, mus, g, d) ->
// per used slot the box-mapped (mua, mus (1 - g), g, n, d) in [-1, 1] (log
// scale where the box says), unused slots 0, then the slots' occupied flags;
// and whether the configuration lies in the box.  The same double operations
// as the host encoding, rounded to float at the end.
struct EncodeArgs {
    const double* layers;
    const int32_t* count;
    long n;
    int slots, min_layers;
    double lo[5], span[5], loglo[5], logspan[5], lo_tol[5], hi_tol[5];
    int logscale[5];
    float* x;
    uint8_t* in_range;
};

__global__ void encode_kernel(EncodeArgs a) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int c = a.count[i];
    bool ok = c >= a.min_layers && c <= a.slots;
    const int nf = 6 * a.slots;
    float* x = a.x + i * nf;
    for (int sl = 0; sl < a.slots; sl++) {
        const bool used = sl < c;
        const double* t = a.layers + (i * a.slots + sl) * 5;
        const double v[5] = {t[1], t[2] * (1.0 - t[3]), t[3], t[0], t[4]};
        for (int k = 0; k < 5; k++) {
            double u = 0.0;
            if (used) {
                ok = ok && v[k] >= a.lo_tol[k] && v[k] <= a.hi_tol[k];
                u = a.logscale[k] ? 2.0 * (log(v[k]) - a.loglo[k]) / a.logspan[k] - 1.0
                                  : 2.0 * (v[k] - a.lo[k]) / a.span[k] - 1.0;
            }
            x[sl * 5 + k] = (float)u;
        }
        x[5 * a.slots + sl] = used ? 1.0f : 0.0f;
    }
    a.in_range[i] = ok ? 1 : 0;
}

// values[b][j] = exp(out[b][j] * scale[j] + shift[j]), j < outputs
__global__ void answer_values_kernel(const float* out, long rows, int head, int outputs, const float* scale,
                                     float* values) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * outputs) return;
    const long b = e / outputs;
    const int j = (int)(e - b * outputs);
    values[e] = expf(out[b * head + j] * scale[j] + scale[outputs + j]);
}

// uncertain[b] = mean over j < flag_cols of exp(0.5 clamp(out[b][outputs + j], -10, 10)) * scale[j]
// above sigma_max
__global__ void answer_flags_kernel(const float* out, long rows, int head, int outputs, int flag_cols,
                                    const float* scale, float sigma_max, uint8_t* uncertain) {
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= rows) return;
    const float* o = out + b * head + outputs;
    float sum = 0.0f;
    for (int j = 0; j < flag_cols; j++) sum += expf(0.5f * fminf(fmaxf(o[j], -10.0f), 10.0f)) * scale[j];
    uncertain[b] = (sum / (float)flag_cols > sigma_max) ? 1 : 0;
}

__global__ void to_bf16_kernel(const float* src, uint16_t* dst, long count) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = f2bf(src[i]);
}

// ---------------------------------------------------------------------------
// Fused inference: the whole network on one row tile per block.
//
// The tile's activations stay in LDS as bf16 [rows][HMAX + 8] for the whole
// pass (input columns 0..63 first, each hidden layer overwriting them after a
// barrier); nothing but the fp32 head rows goes back to HBM.  Weights are read
// from L2 in the matrix-core fragment order they are packed into (one
// coalesced 1 KB read per fragment): fragment (c, s) of a layer holds, for
// lane l, W[32c + l%32][16s + 8(l/32) .. +7].
//
// Products are computed transposed, D[n][m] = sum_k W[n][k] X[m][k], so a
// lane's accumulator holds four consecutive output columns of one row: the
// epilogue writes 8-byte bf16 quads to LDS and 16-byte fp32 quads of the head.
// Accumulators start at the bias, the k order is ascending (16 per
// instruction) and the epilogue arithmetic is that of the per-layer GEMMs, so
// both give the same bits.
//
// A block holds R = kTileRows/32 row tiles and 2R waves.  A layer with
// NT = fout/32 column tiles gives each wave NT R / (2R) = NT/2 accumulator
// tiles: all R row tiles of NT/(2R) column tiles when 2R divides NT, else one
// row tile (wave % R) of NT/2 column tiles.
//
// Overlap of the GELU epilogue with the matrix products: a hidden layer whose
// waves own an even number of column tiles runs them in two passes.  The
// second pass's k loop also finishes the first pass's accumulators (bias was
// their start; GELU and bf16 rounding remain), a few elements per k step,
// into packed registers, so that VALU work issues while the matrix core is
// busy.  Only the second pass's epilogue runs after the layer's barrier, and
// both halves are written to LDS then (the layer is computed in place: every
// wave must have read its input before any output is stored).

#ifndef XMC_INFER_ROWS
#define XMC_INFER_ROWS 64
#endif
constexpr int kTileRows = XMC_INFER_ROWS;     // rows per block: kTileRows / 32 row tiles
constexpr int kInferWaves = kTileRows / 16;
constexpr int kInferThreads = 64 * kInferWaves;
constexpr int kRowTiles = kTileRows / 32;

struct InferArgs {
    const float* x;           // [batch][nin] fp32
    int batch, nin;
    const uint16_t* wp;       // packed bf16 weights, layer l at woff[l]
    const float* w;           // fp32 parameters (biases at boff[l])
    float* out;               // [batch][head] fp32
    int layers, head;
    long woff[9], boff[9];
};

// accumulator tiles acc[j][i] (column tile ct + j, any row tile) start at the bias
template <int RT, int CT>
__device__ __forceinline__ void infer_start(f32x16 (&acc)[CT][RT], const float* bias, int ct, int lane) {
#pragma unroll
    for (int j = 0; j < CT; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 b = *reinterpret_cast<const float4*>(bias + 32 * (ct + j) + 8 * q + 4 * (lane >> 5));
#pragma unroll
            for (int i = 0; i < RT; i++) {
                acc[j][i][4 * q] = b.x;
                acc[j][i][4 * q + 1] = b.y;
                acc[j][i][4 * q + 2] = b.z;
                acc[j][i][4 * q + 3] = b.w;
            }
        }
}

// GELU + bf16 of elements e, e + 1 (flattened over pend[j][i][v]) into held
template <int RT, int CP>
__device__ __forceinline__ void infer_finish_pair(const f32x16 (&pend)[CP][RT], uint32_t (&held)[CP * RT * 8], int e) {
    const int p = e >> 4, v = e & 15;
    held[e >> 1] = bf16x2_rne(gelu(pend[p / RT][p % RT][v]), gelu(pend[p / RT][p % RT][v + 1]));
    // pins the value here: otherwise the compiler sinks the arithmetic to the
    // value's only use, the LDS store after the layer's barrier
    asm volatile("" : "+v"(held[e >> 1]));
}

// One pass of a layer's products over KS k steps: acc[j][i] += W(col tile
// ct + j) X(row tile rt0 + i).  A weight fragment is requested three steps
// before its products (a ring of four register sets), an activation fragment
// one step before (a ring of two); the scheduling barriers keep each step's
// requests ahead of its products.  CP > 0: also finishes pend into held.
template <int HS, int RT, int CT, int KS, int CP, int RTP>
__device__ __forceinline__ void infer_pass(f32x16 (&acc)[CT][RT], const uint16_t* wl, const uint16_t* act, int rt0,
                                           int ct, int lane, const f32x16 (&pend)[CP > 0 ? CP : 1][RTP],
                                           uint32_t (&held)[CP > 0 ? CP * RTP * 8 : 1]) {
    constexpr int E = CP * RTP * 16 / KS;  // pending elements finished per k step
    static_assert(CP == 0 || (E % 2 == 0 && E * KS == CP * RTP * 16), "pending elements must split into pairs per step");
    // a compiler-level fence: this pass's first requests are not hoisted into
    // the previous pass, where they would only occupy registers
    asm volatile("" ::: "memory");
    const bf16x8* wf = reinterpret_cast<const bf16x8*>(wl) + (long)ct * KS * 64 + lane;
    const uint16_t* ap = act + (32 * rt0 + (lane & 31)) * HS + 8 * (lane >> 5);
    bf16x8 w[4][CT], a[2][RT];
#pragma unroll
    for (int s = 0; s < 3 && s < KS; s++)
#pragma unroll
        for (int j = 0; j < CT; j++) w[s][j] = wf[(j * KS + s) * 64];
#pragma unroll
    for (int i = 0; i < RT; i++) a[0][i] = *reinterpret_cast<const bf16x8*>(ap + 32 * i * HS);
#pragma unroll
    for (int s = 0; s < KS; s++) {
        if (s + 3 < KS) {
#pragma unroll
            for (int j = 0; j < CT; j++) w[(s + 3) & 3][j] = wf[(j * KS + s + 3) * 64];
        }
        if (s + 1 < KS) {
#pragma unroll
            for (int i = 0; i < RT; i++)
                a[(s + 1) & 1][i] = *reinterpret_cast<const bf16x8*>(ap + 32 * i * HS + (s + 1) * 16);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < CT; j++)
#pragma unroll
            for (int i = 0; i < RT; i++)
                acc[j][i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[s & 3][j], a[s & 1][i], acc[j][i], 0, 0, 0);
        if constexpr (CP > 0) {
#pragma unroll
            for (int e = s * E; e < (s + 1) * E; e += 2) infer_finish_pair<RTP, CP>(pend, held, e);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// bf16 quads of held (pairs of acc[j][i][v]) into the LDS activations
template <int HS, int RT, int CT>
__device__ __forceinline__ void infer_store_held(const uint32_t (&held)[CT * RT * 8], uint16_t* act, int rt0, int ct,
                                                 int lane) {
#pragma unroll
    for (int j = 0; j < CT; j++)
#pragma unroll
        for (int i = 0; i < RT; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int n = 32 * (ct + j) + 8 * q + 4 * (lane >> 5);
                const int m = 32 * (rt0 + i) + (lane & 31);
                const int h = (j * RT + i) * 8 + 2 * q;
                *reinterpret_cast<uint2*>(act + m * HS + n) = make_uint2(held[h], held[h + 1]);
            }
}

// fp32 head rows (accumulators already hold bias + products)
template <int RT, int CT>
__device__ __forceinline__ void infer_store_head(const f32x16 (&acc)[CT][RT], const InferArgs& g, int r0, int rt0,
                                                 int ct, int lane) {
#pragma unroll
    for (int j = 0; j < CT; j++)
#pragma unroll
        for (int i = 0; i < RT; i++) {
            const int m = 32 * (rt0 + i) + (lane & 31);
            if (r0 + m < g.batch) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int n = 32 * (ct + j) + 8 * q + 4 * (lane >> 5);
                    f32x4 o = {acc[j][i][4 * q], acc[j][i][4 * q + 1], acc[j][i][4 * q + 2], acc[j][i][4 * q + 3]};
                    __builtin_nontemporal_store(o, reinterpret_cast<f32x4*>(g.out + (long)(r0 + m) * g.head + n));
                }
            }
        }
}

// A hidden layer (fin = 16 KS, GELU into LDS) or the head (fp32 rows out)
template <int HS, int RT, int CT, int KS, bool HEAD>
__device__ __forceinline__ void infer_layer(const InferArgs& g, int l, uint16_t* act, int r0, int w, int lane) {
    const int rt0 = RT == kRowTiles ? 0 : w % kRowTiles;
    const int ct0 = RT == kRowTiles ? w * CT : (w / kRowTiles) * CT;
    const uint16_t* wl = g.wp + g.woff[l];
    const float* bias = g.w + g.boff[l];
    f32x16 none[1][RT];
    uint32_t none_held[1];
    if constexpr (HEAD || CT % 2) {
        f32x16 acc[CT][RT];
        infer_start<RT, CT>(acc, bias, ct0, lane);
        infer_pass<HS, RT, CT, KS, 0, RT>(acc, wl, act, rt0, ct0, lane, none, none_held);
        if constexpr (HEAD) {
            infer_store_head<RT, CT>(acc, g, r0, rt0, ct0, lane);
        } else {
            uint32_t held[CT * RT * 8];
#pragma unroll
            for (int e = 0; e < CT * RT * 16; e += 2) infer_finish_pair<RT, CT>(acc, held, e);
            __syncthreads();  // every wave has read the layer's input before it is overwritten
            infer_store_held<HS, RT, CT>(held, act, rt0, ct0, lane);
            __syncthreads();
        }
    } else {
        constexpr int CH = CT / 2;
        f32x16 first[CH][RT], second[CH][RT];
        uint32_t held1[CH * RT * 8], held2[CH * RT * 8];
        infer_start<RT, CH>(first, bias, ct0, lane);
        infer_pass<HS, RT, CH, KS, 0, RT>(first, wl, act, rt0, ct0, lane, none, none_held);
        infer_start<RT, CH>(second, bias, ct0 + CH, lane);
        infer_pass<HS, RT, CH, KS, CH, RT>(second, wl, act, rt0, ct0 + CH, lane, first, held1);
#pragma unroll
        for (int e = 0; e < CH * RT * 16; e += 2) infer_finish_pair<RT, CH>(second, held2, e);
        __syncthreads();
        infer_store_held<HS, RT, CH>(held1, act, rt0, ct0, lane);
        infer_store_held<HS, RT, CH>(held2, act, rt0, ct0 + CH, lane);
        __syncthreads();
    }
}

// H: hidden width; RTH x CTH accumulator tiles per wave in the hidden layers,
// RTO x CTO in the head
template <int H, int RTH, int CTH, int RTO, int CTO>
__global__ __launch_bounds__(kInferThreads) __attribute__((amdgpu_waves_per_eu(2))) void infer_kernel(InferArgs g) {
    constexpr int HS = H + 8;  // halfwords per LDS row: 16-byte aligned, consecutive rows on rotating banks
    __shared__ __attribute__((aligned(16))) uint16_t act[kTileRows * HS];
    // the wave index as a scalar: weight addresses are then a scalar base plus the lane's 16 bytes
    const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int r0 = blockIdx.x * kTileRows;
    // input tile: columns 0..63, zero beyond nin and beyond the batch
    for (int e = t; e < kTileRows * 64; e += kInferThreads) {
        const int r = e >> 6, c = e & 63;
        float v = (c < g.nin && r0 + r < g.batch) ? g.x[(long)(r0 + r) * g.nin + c] : 0.0f;
        act[r * HS + c] = bf16_rne(v);
    }
    __syncthreads();
    infer_layer<HS, RTH, CTH, 4, false>(g, 0, act, r0, w, lane);
    for (int l = 1; l < g.layers - 1; l++) infer_layer<HS, RTH, CTH, H / 16, false>(g, l, act, r0, w, lane);
    infer_layer<HS, RTO, CTO, H / 16, true>(g, g.layers - 1, act, r0, w, lane);
}

// row-major bf16 weights [fout][fin] -> fragment order (see above)
__global__ void pack_fragments_kernel(const uint16_t* src, uint16_t* dst, int fout, int fin) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)fout * fin) return;
    const int ksn = fin / 16;
    const int e = (int)(p & 7), lane = (int)((p >> 3) & 63);
    const long f = p >> 9;  // fragment index c * ksn + s
    const int c = (int)(f / ksn), s = (int)(f % ksn);
    dst[p] = src[(long)(32 * c + (lane & 31)) * fin + 16 * s + 8 * (lane >> 5) + e];
}

}  // namespace xmlp

using namespace xmlp;

namespace {

int fail(int code, const std::string& msg) { return xmc::set_error(code, msg); }

#define MLP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail(-2, std::string(#expr) + ": " + hipGetErrorString(_e));                \
    } while (0)

inline int round_up(int x, int q) { return (x + q - 1) / q * q; }

// 128 x 128 tiles when both dimensions allow and there are enough tiles to
// fill the CUs, 64 x 64 otherwise
template <bool AK, bool BKC, int EPI>
hipError_t launch_gemm(const GemmArgs& g, int splits, hipStream_t s, bool force128 = false) {
    if (EPI != EPI_PARTIAL && g.M % BM2 == 0 && g.N % BN2 == 0 && (g.kchunk % BK2) == 0 &&
        (long)(g.M / BM2) * (g.N / BN2) * splits >= (force128 ? 1 : 256)) {
        dim3 grid(g.N / BN2, g.M / BM2, splits);
        hipLaunchKernelGGL((gemm128_kernel<AK, BKC, EPI>), grid, dim3(256), 0, s, g);
    } else {
        dim3 grid(g.N / BN, g.M / BM, splits);
        hipLaunchKernelGGL((gemm_kernel<AK, BKC, EPI>), grid, dim3(256), 0, s, g);
    }
    return hipGetLastError();
}

int split_count(int k) {
    for (int s : {8, 4, 2})
        if (k % (s * BK) == 0 && k / s >= 256) return s;
    return 1;
}

// every reduction length handed to the kernel is a multiple of BK
static_assert(BK == 64, "padded widths (input 64, head 2*outputs, hidden, batch rows) assume 64");

}  // namespace

struct xmc_mlp {
    xmc_mlp_desc d{};
    int in_pad = 0, head = 0, rows = 0;  // padded input width, head width, padded batch rows
    int layers = 0;
    std::vector<int> fin, fout;          // padded fan-in / fan-out per layer
    std::vector<long> woff, boff;        // offsets in the flat padded parameter array
    long nparam = 0;
    float* w = nullptr;   // fp32 master parameters, padded layout
    float* g = nullptr;   // gradients
    float* m1 = nullptr;
    float* m2 = nullptr;
    uint16_t* wb = nullptr;  // bf16 mirror (same offsets as w; only weight blocks used)
    uint16_t* wp = nullptr;  // the mirror's weight blocks in matrix-core fragment order (fused inference)
    bool packed = false;     // wp matches wb
    uint16_t* a0 = nullptr;  // packed input [rows][in_pad]
    std::vector<uint16_t*> z, a;  // per hidden layer [rows][hidden]
    float* out = nullptr;         // [rows][head]
    uint16_t* dout = nullptr;     // [rows][head]
    uint16_t* dz[2] = {nullptr, nullptr};
    float* part = nullptr;        // split-K partials
    long part_count = 0;
    float* lossp = nullptr;       // loss partial sums
    float *xb = nullptr, *yb = nullptr, *mb = nullptr;  // sampled batch (xmc_mlp_train_sampled)
    int loss_blocks = 0;
    int step = 0;
    float* ans_out = nullptr;    // head outputs of one answer chunk [kAnswerRows][head]
    float* ans_scale = nullptr;  // per output: scale, then shift
    std::vector<void*> allocs;
};

namespace {

hipError_t dalloc(xmc_mlp* m, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) m->allocs.push_back(*p);
    return e;
}

// calls f(index in caller order, index in the padded device layout) for every parameter
void for_each_param(const xmc_mlp* m, const std::function<void(long, long)>& f) {
    long u = 0;
    for (int l = 0; l < m->layers; l++) {
        int in_user = l == 0 ? m->d.inputs : m->d.hidden;
        int out_user = l == m->layers - 1 ? m->head : m->d.hidden;
        for (int r = 0; r < out_user; r++)
            for (int c = 0; c < in_user; c++) f(u++, m->woff[l] + (long)r * m->fin[l] + c);
        for (int r = 0; r < out_user; r++) f(u++, m->boff[l] + r);
    }
}

long user_count(const xmc_mlp* m) {
    long n = 0;
    for (int l = 0; l < m->layers; l++) {
        int in_user = l == 0 ? m->d.inputs : m->d.hidden;
        int out_user = l == m->layers - 1 ? m->head : m->d.hidden;
        n += (long)out_user * in_user + out_user;
    }
    return n;
}

int refresh_mirror(xmc_mlp* m, hipStream_t s) {
    long n = m->nparam;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m->w, m->wb, n);
    MLP_TRY(hipGetLastError());
    m->packed = false;
    return 0;
}

// The fused kernel for this model's widths, or nullptr: [kTileRows][hidden + 8]
// bf16 activations in LDS, each wave's accumulator tiles fixed at compile time
typedef void (*InferKernel)(InferArgs);

// accumulator tiles per wave of a layer with nt column tiles
constexpr int shape_rt(int nt) { return nt % kInferWaves == 0 ? kRowTiles : 1; }
constexpr int shape_ct(int nt) { return nt % kInferWaves == 0 ? nt / kInferWaves : nt / 2; }

template <int H, int NTO>
InferKernel pick() {
    return infer_kernel<H, shape_rt(H / 32), shape_ct(H / 32), shape_rt(NTO), shape_ct(NTO)>;
}

InferKernel fused_kernel(const xmc_mlp* m) {
    const int h = m->d.hidden, o = m->head;
#define XMC_INFER(H)                                                          \
    switch (o) {                                                              \
        case 64: return pick<H, 2>();                                         \
        case 128: return pick<H, 4>();                                        \
        case 192: return pick<H, 6>();                                        \
        case 256: return pick<H, 8>();                                        \
        default: return nullptr;                                              \
    }
    if (h == 128) XMC_INFER(128)
    if (h == 256) XMC_INFER(256)
    if (h == 512) XMC_INFER(512)
#undef XMC_INFER
    return nullptr;
}

// whole-network forward of `batch` rows of x into out, one launch
int forward_fused(xmc_mlp* m, const float* x, int batch, float* out, hipStream_t s) {
    if (!m->packed) {
        for (int l = 0; l < m->layers; l++) {
            long n = (long)m->fout[l] * m->fin[l];
            hipLaunchKernelGGL(pack_fragments_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               m->wb + m->woff[l], m->wp + m->woff[l], m->fout[l], m->fin[l]);
            MLP_TRY(hipGetLastError());
        }
        m->packed = true;
    }
    if (batch == 0) return 0;
    InferArgs g{};
    g.x = x;
    g.batch = batch;
    g.nin = m->d.inputs;
    g.wp = m->wp;
    g.w = m->w;
    g.out = out;
    g.layers = m->layers;
    g.head = m->head;
    for (int l = 0; l < m->layers; l++) {
        g.woff[l] = m->woff[l];
        g.boff[l] = m->boff[l];
    }
    dim3 grid((unsigned)((batch + kTileRows - 1) / kTileRows));
    hipLaunchKernelGGL(fused_kernel(m), grid, dim3(kInferThreads), 0, s, g);
    MLP_TRY(hipGetLastError());
    return 0;
}

// forward of `batch` rows (<= rows) already packed into a0
// head_out: where the fp32 head rows go (the model's buffer, or the caller's
// when every one of the `rows` rows is the caller's)
int forward_packed(xmc_mlp* m, int rows, hipStream_t s, float* head_out) {
    const uint16_t* prev = m->a0;
    long ldp = m->in_pad;
    for (int l = 0; l < m->layers; l++) {
        GemmArgs g{};
        g.A = prev;
        g.lda = ldp;
        g.B = m->wb + m->woff[l];
        g.ldb = m->fin[l];
        g.M = rows;
        g.N = m->fout[l];
        g.kchunk = m->fin[l];
        g.bias = m->w + m->boff[l];
        if (l < m->layers - 1) {
            g.out = m->a[l];
            g.zout = m->z[l];
            g.ldo = m->d.hidden;
            MLP_TRY((launch_gemm<true, true, EPI_GELU>(g, 1, s)));
            prev = m->a[l];
            ldp = m->d.hidden;
        } else {
            g.outf = head_out;
            g.ldo = m->head;
            MLP_TRY((launch_gemm<true, true, EPI_LINEAR>(g, 1, s)));
        }
    }
    return 0;
}

int pack(xmc_mlp* m, const float* x, int batch, int rows, hipStream_t s) {
    long n = (long)rows * m->in_pad;
    hipLaunchKernelGGL(pack_input_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, batch,
                       m->d.inputs, m->a0, m->in_pad, rows);
    MLP_TRY(hipGetLastError());
    return 0;
}

// dW_l (+ db_l) of layer l from the gradient `dz` [rows][fout[l]] and the
// layer's input `in` [rows][fin[l]]
int weight_grad(xmc_mlp* m, int l, const uint16_t* dz, const uint16_t* in, int rows, hipStream_t s) {
    int splits = split_count(rows);
    GemmArgs g{};
    g.A = dz;
    g.lda = m->fout[l];
    g.B = in;
    g.ldb = m->fin[l];
    g.M = m->fout[l];
    g.N = m->fin[l];
    g.kchunk = rows / splits;
    g.outf = m->part;
    g.ldo = m->fin[l];
    MLP_TRY((launch_gemm<false, false, EPI_PARTIAL>(g, splits, s)));
    long cnt = (long)m->fout[l] * m->fin[l];
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m->part,
                       splits, cnt, m->g + m->woff[l]);
    MLP_TRY(hipGetLastError());
    // bias gradient: column sums in slices (reusing the partial buffer), then their fixed-order sum
    hipLaunchKernelGGL(colsum_kernel, dim3((m->fout[l] + 63) / 64, kColsumSlices), dim3(256), 0, s, dz, rows,
                       m->fout[l], m->part);
    MLP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((m->fout[l] + 255) / 256), dim3(256), 0, s, m->part,
                       kColsumSlices, (long)m->fout[l], m->g + m->boff[l]);
    MLP_TRY(hipGetLastError());
    return 0;
}

int backward_impl(xmc_mlp* m, const float* x, const float* y, const float* mask, int batch, float scale,
                  float* loss, hipStream_t s) {
    if (batch <= 0 || batch > m->d.maxBatch) return fail(-1, "batch outside 1..maxBatch");
    if (!x || !y || !mask) return fail(-1, "NULL device array");
    int rows = round_up(batch, BM);
    int rc = pack(m, x, batch, rows, s);
    if (rc) return rc;
    if ((rc = forward_packed(m, rows, s, m->out))) return rc;
    long n = (long)rows * m->d.outputs;
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(loss_kernel, dim3(blocks), dim3(256), 0, s, m->out, y, mask, batch, m->d.outputs, rows,
                       scale, m->dout, m->lossp);
    MLP_TRY(hipGetLastError());

    const int L = m->layers;
    // head
    const uint16_t* head_in = m->a[L - 2];
    if ((rc = weight_grad(m, L - 1, m->dout, head_in, rows, s))) return rc;
    int cur = 0;
    {
        GemmArgs g{};
        g.A = m->dout;
        g.lda = m->head;
        g.B = m->wb + m->woff[L - 1];
        g.ldb = m->fin[L - 1];
        g.M = rows;
        g.N = m->fin[L - 1];
        g.kchunk = m->head;
        g.zin = m->z[L - 2];
        g.out = m->dz[cur];
        g.ldo = m->d.hidden;
        MLP_TRY((launch_gemm<true, false, EPI_DGELU>(g, 1, s)));
    }
    for (int l = L - 2; l >= 0; l--) {
        const uint16_t* in = l == 0 ? m->a0 : m->a[l - 1];
        if ((rc = weight_grad(m, l, m->dz[cur], in, rows, s))) return rc;
        if (l > 0) {
            GemmArgs g{};
            g.A = m->dz[cur];
            g.lda = m->d.hidden;
            g.B = m->wb + m->woff[l];
            g.ldb = m->fin[l];
            g.M = rows;
            g.N = m->fin[l];
            g.kchunk = m->d.hidden;
            g.zin = m->z[l - 1];
            g.out = m->dz[cur ^ 1];
            g.ldo = m->d.hidden;
            MLP_TRY((launch_gemm<true, false, EPI_DGELU>(g, 1, s)));
            cur ^= 1;
        }
    }
    if (loss) {
        std::vector<float> h(blocks);
        MLP_TRY(hipMemcpyAsync(h.data(), m->lossp, blocks * sizeof(float), hipMemcpyDeviceToHost, s));
        MLP_TRY(hipStreamSynchronize(s));
        double acc = 0.0;
        for (float v : h) acc += v;
        *loss = (float)(acc * scale);
    }
    return 0;
}

int adam_impl(xmc_mlp* m, float lr, float b1, float b2, float eps, float decay, hipStream_t s) {
    m->step++;
    float c1 = 1.0f - std::pow(b1, (float)m->step);
    float c2 = 1.0f - std::pow(b2, (float)m->step);
    for (int l = 0; l < m->layers; l++) {
        long nw = (long)m->fout[l] * m->fin[l];
        m->packed = false;
        hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, m->w + m->woff[l],
                           m->g + m->woff[l], m->m1 + m->woff[l], m->m2 + m->woff[l], m->wb + m->woff[l], nw, lr,
                           b1, b2, eps, c1, c2, decay);
        MLP_TRY(hipGetLastError());
        long nb = m->fout[l];
        hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, m->w + m->boff[l],
                           m->g + m->boff[l], m->m1 + m->boff[l], m->m2 + m->boff[l], (uint16_t*)nullptr, nb, lr,
                           b1, b2, eps, c1, c2, 0.0f);
        MLP_TRY(hipGetLastError());
    }
    return 0;
}

// SplitMix64, for the initial weights
uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int xmc_mlp_destroy(xmc_mlp* m) {
    if (!m) return 0;
    (void)hipSetDevice(m->d.device);
    (void)hipDeviceSynchronize();
    for (void* p : m->allocs) (void)hipFree(p);
    delete m;
    return 0;
}

int xmc_mlp_create(const xmc_mlp_desc* desc, uint64_t seed, xmc_mlp** out) {
    if (!desc || !out) return fail(-1, "NULL argument");
    *out = nullptr;
    const xmc_mlp_desc& d = *desc;
    if (d.inputs < 1 || d.inputs > 64) return fail(-1, "inputs must be in 1..64");
    if (d.hidden < 64 || d.hidden % 64) return fail(-1, "hidden must be a positive multiple of 64");
    if (d.depth < 1 || d.depth > 8) return fail(-1, "depth must be in 1..8");
    if (d.outputs < 32 || d.outputs % 32) return fail(-1, "outputs must be a positive multiple of 32");
    if (d.maxBatch < 1 || d.maxBatch > (1 << 20)) return fail(-1, "maxBatch must be in 1..2^20");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(-4, "no GPU");
    if (d.device < 0 || d.device >= ndev) return fail(-1, "device index out of range");
    MLP_TRY(hipSetDevice(d.device));

    auto* m = new xmc_mlp();
    std::unique_ptr<xmc_mlp, int (*)(xmc_mlp*)> guard(m, xmc_mlp_destroy);
    m->d = d;
    m->in_pad = 64;
    m->head = 2 * d.outputs;
    m->rows = round_up(d.maxBatch, BM);
    m->layers = d.depth + 1;
    long off = 0;
    for (int l = 0; l < m->layers; l++) {
        int fi = l == 0 ? m->in_pad : d.hidden;
        int fo = l == m->layers - 1 ? m->head : d.hidden;
        m->fin.push_back(fi);
        m->fout.push_back(fo);
        m->woff.push_back(off);
        off += (long)fi * fo;
        m->boff.push_back(off);
        off += fo;
        off = (off + 63) / 64 * 64;  // keep every block 256-byte aligned
    }
    m->nparam = off;
    const size_t pb = off * sizeof(float);
    MLP_TRY(dalloc(m, (void**)&m->w, pb));
    MLP_TRY(dalloc(m, (void**)&m->g, pb));
    MLP_TRY(dalloc(m, (void**)&m->m1, pb));
    MLP_TRY(dalloc(m, (void**)&m->m2, pb));
    MLP_TRY(dalloc(m, (void**)&m->wb, off * sizeof(uint16_t)));
    MLP_TRY(dalloc(m, (void**)&m->wp, off * sizeof(uint16_t)));
    MLP_TRY(dalloc(m, (void**)&m->a0, (size_t)m->rows * m->in_pad * 2));
    for (int l = 0; l < d.depth; l++) {
        uint16_t *zp, *ap;
        MLP_TRY(dalloc(m, (void**)&zp, (size_t)m->rows * d.hidden * 2));
        MLP_TRY(dalloc(m, (void**)&ap, (size_t)m->rows * d.hidden * 2));
        m->z.push_back(zp);
        m->a.push_back(ap);
    }
    MLP_TRY(dalloc(m, (void**)&m->out, (size_t)m->rows * m->head * 4));
    MLP_TRY(dalloc(m, (void**)&m->dout, (size_t)m->rows * m->head * 2));
    MLP_TRY(dalloc(m, (void**)&m->dz[0], (size_t)m->rows * d.hidden * 2));
    MLP_TRY(dalloc(m, (void**)&m->dz[1], (size_t)m->rows * d.hidden * 2));
    long maxw = 0;
    for (int l = 0; l < m->layers; l++) maxw = std::max(maxw, (long)m->fin[l] * m->fout[l]);
    m->part_count = 8 * maxw;
    MLP_TRY(dalloc(m, (void**)&m->part, m->part_count * 4));
    MLP_TRY(dalloc(m, (void**)&m->xb, (size_t)d.maxBatch * d.inputs * 4));
    MLP_TRY(dalloc(m, (void**)&m->yb, (size_t)d.maxBatch * d.outputs * 4));
    MLP_TRY(dalloc(m, (void**)&m->mb, (size_t)d.maxBatch * d.outputs * 4));
    m->loss_blocks = (int)(((long)m->rows * d.outputs + 255) / 256);
    MLP_TRY(dalloc(m, (void**)&m->lossp, (size_t)m->loss_blocks * 4));

    // Glorot-uniform weights, zero biases (padded columns stay zero)
    std::vector<float> host(off, 0.0f);
    uint64_t st = seed;
    for (int l = 0; l < m->layers; l++) {
        int in_user = l == 0 ? d.inputs : d.hidden;
        int out_user = m->fout[l];
        float lim = std::sqrt(6.0f / (float)(in_user + out_user));
        for (int r = 0; r < out_user; r++)
            for (int c = 0; c < in_user; c++) {
                double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
                host[m->woff[l] + (long)r * m->fin[l] + c] = (float)((2.0 * u - 1.0) * lim);
            }
    }
    MLP_TRY(hipMemcpy(m->w, host.data(), pb, hipMemcpyHostToDevice));
    MLP_TRY(hipMemset(m->g, 0, pb));
    MLP_TRY(hipMemset(m->m1, 0, pb));
    MLP_TRY(hipMemset(m->m2, 0, pb));
    int rc = refresh_mirror(m, nullptr);
    if (rc) return rc;
    MLP_TRY(hipDeviceSynchronize());
    *out = guard.release();
    return 0;
}

int xmc_mlp_param_count(const xmc_mlp* m, int64_t* count) {
    if (!m || !count) return fail(-1, "NULL argument");
    *count = user_count(m);
    return 0;
}

static int copy_out(xmc_mlp* m, const float* src, float* dst, int64_t count) {
    if (!m || !dst) return fail(-1, "NULL argument");
    if (count != user_count(m)) return fail(-1, "parameter count mismatch");
    MLP_TRY(hipSetDevice(m->d.device));
    MLP_TRY(hipDeviceSynchronize());
    std::vector<float> h(m->nparam);
    MLP_TRY(hipMemcpy(h.data(), src, m->nparam * 4, hipMemcpyDeviceToHost));
    for_each_param(m, [&](long u, long p) { dst[u] = h[p]; });
    return 0;
}

int xmc_mlp_get_params(xmc_mlp* m, float* params, int64_t count) {
    return m ? copy_out(m, m->w, params, count) : fail(-1, "NULL model");
}

int xmc_mlp_get_grads(xmc_mlp* m, float* grads, int64_t count) {
    return m ? copy_out(m, m->g, grads, count) : fail(-1, "NULL model");
}

int xmc_mlp_set_params(xmc_mlp* m, const float* params, int64_t count) {
    if (!m || !params) return fail(-1, "NULL argument");
    if (count != user_count(m)) return fail(-1, "parameter count mismatch");
    MLP_TRY(hipSetDevice(m->d.device));
    MLP_TRY(hipDeviceSynchronize());
    std::vector<float> h(m->nparam, 0.0f);
    for_each_param(m, [&](long u, long p) { h[p] = params[u]; });
    MLP_TRY(hipMemcpy(m->w, h.data(), m->nparam * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemset(m->m1, 0, m->nparam * 4));
    MLP_TRY(hipMemset(m->m2, 0, m->nparam * 4));
    m->step = 0;
    int rc = refresh_mirror(m, nullptr);
    if (rc) return rc;
    MLP_TRY(hipDeviceSynchronize());
    return 0;
}

int xmc_mlp_forward_layers(xmc_mlp* m, const float* x, int batch, float* out, void* stream) {
    if (!m || !x || !out) return fail(-1, "NULL argument");
    if (batch < 0) return fail(-1, "negative batch");
    MLP_TRY(hipSetDevice(m->d.device));
    hipStream_t s = (hipStream_t)stream;
    for (int b0 = 0; b0 < batch; b0 += m->d.maxBatch) {
        int nb = std::min(m->d.maxBatch, batch - b0);
        int rows = round_up(nb, BM);
        int rc = pack(m, x + (long)b0 * m->d.inputs, nb, rows, s);
        if (rc) return rc;
        float* dst = out + (long)b0 * m->head;
        if ((rc = forward_packed(m, rows, s, nb == rows ? dst : m->out))) return rc;
        if (nb != rows)
            MLP_TRY(hipMemcpyAsync(dst, m->out, (size_t)nb * m->head * 4, hipMemcpyDeviceToDevice, s));
    }
    return 0;
}

int xmc_mlp_forward(xmc_mlp* m, const float* x, int batch, float* out, void* stream) {
    if (!m || !x || !out) return fail(-1, "NULL argument");
    if (batch < 0) return fail(-1, "negative batch");
    if (!fused_kernel(m) || ((uintptr_t)out & 15)) return xmc_mlp_forward_layers(m, x, batch, out, stream);
    MLP_TRY(hipSetDevice(m->d.device));
    return forward_fused(m, x, batch, out, (hipStream_t)stream);
}

int xmc_mlp_backward(xmc_mlp* m, const float* x, const float* y, const float* mask, int batch, float scale,
                     float* loss, void* stream) {
    if (!m) return fail(-1, "NULL model");
    MLP_TRY(hipSetDevice(m->d.device));
    return backward_impl(m, x, y, mask, batch, scale, loss, (hipStream_t)stream);
}

int xmc_mlp_adam(xmc_mlp* m, float lr, float b1, float b2, float eps, float decay, void* stream) {
    if (!m) return fail(-1, "NULL model");
    MLP_TRY(hipSetDevice(m->d.device));
    return adam_impl(m, lr, b1, b2, eps, decay, (hipStream_t)stream);
}

int xmc_mlp_train_step(xmc_mlp* m, const float* x, const float* y, const float* mask, int batch, float scale,
                       float lr, float b1, float b2, float eps, float decay, void* stream) {
    if (!m) return fail(-1, "NULL model");
    MLP_TRY(hipSetDevice(m->d.device));
    int rc = backward_impl(m, x, y, mask, batch, scale, nullptr, (hipStream_t)stream);
    if (rc) return rc;
    return adam_impl(m, lr, b1, b2, eps, decay, (hipStream_t)stream);
}

int xmc_mlp_train_sampled(xmc_mlp* m, const float* X, const float* Y, const float* M, int64_t rows, int batch,
                          uint64_t seed, int64_t step, float scale, float lr, float b1, float b2, float eps,
                          float decay, float* loss, void* stream) {
    if (!m || !X || !Y || !M) return fail(-1, "NULL argument");
    if (rows <= 0) return fail(-1, "empty data set");
    if (batch <= 0 || batch > m->d.maxBatch) return fail(-1, "batch outside 1..maxBatch");
    MLP_TRY(hipSetDevice(m->d.device));
    hipStream_t s = (hipStream_t)stream;
    uint64_t key = seed * 0x9e3779b97f4a7c15ull + (uint64_t)step * 0xd1b54a32d192ed03ull;
    long elems = (long)batch * (m->d.inputs + 2 * m->d.outputs);
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((elems + 255) / 256)), dim3(256), 0, s, X, Y, M, (long)rows,
                       m->d.inputs, m->d.outputs, batch, key, m->xb, m->yb, m->mb);
    MLP_TRY(hipGetLastError());
    int rc = backward_impl(m, m->xb, m->yb, m->mb, batch, scale, loss, s);
    if (rc) return rc;
    return adam_impl(m, lr, b1, b2, eps, decay, s);
}

int xmc_surrogate_encode(const double* layers, const int32_t* count, int64_t n, const xmc_layer_box* box, float* x,
                         uint8_t* in_range, int device, void* stream) {
    if (!layers || !count || !box || !x || !in_range) return fail(-1, "NULL argument");
    if (n < 0) return fail(-1, "negative count");
    if (box->slots < 1 || box->slots > 8) return fail(-1, "slots must be 1..8");
    EncodeArgs a{};
    a.layers = layers;
    a.count = count;
    a.n = n;
    a.slots = box->slots;
    a.min_layers = box->minLayers;
    for (int k = 0; k < 5; k++) {
        if (!(box->lo[k] > 0.0 && box->hi[k] > box->lo[k])) return fail(-1, "box bounds must be 0 < lo < hi");
        a.lo[k] = box->lo[k];
        a.span[k] = box->hi[k] - box->lo[k];
        a.loglo[k] = std::log(box->lo[k]);
        a.logspan[k] = std::log(box->hi[k]) - std::log(box->lo[k]);
        a.lo_tol[k] = box->lo[k] * (1.0 - box->tol);
        a.hi_tol[k] = box->hi[k] * (1.0 + box->tol);
        a.logscale[k] = box->logScale[k] != 0;
    }
    a.x = x;
    a.in_range = in_range;
    if (n == 0) return 0;
    MLP_TRY(hipSetDevice(device));
    hipLaunchKernelGGL(encode_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a);
    MLP_TRY(hipGetLastError());
    return 0;
}

int xmc_mlp_answer(xmc_mlp* m, const float* x, int64_t n, const float* scale, const float* shift, int flagColumns,
                   float sigmaMax, float* values, uint8_t* uncertain, void* stream) {
    if (!m || !x || !scale || !shift || !values || !uncertain) return fail(-1, "NULL argument");
    if (n < 0) return fail(-1, "negative count");
    const int O = m->d.outputs;
    if (flagColumns < 1 || flagColumns > O) return fail(-1, "flagColumns must be 1..outputs");
    MLP_TRY(hipSetDevice(m->d.device));
    hipStream_t s = (hipStream_t)stream;
    if (!m->ans_out) {
        MLP_TRY(dalloc(m, (void**)&m->ans_out, (size_t)kAnswerRows * m->head * 4));
        MLP_TRY(dalloc(m, (void**)&m->ans_scale, (size_t)2 * O * 4));
    }
    MLP_TRY(hipMemcpyAsync(m->ans_scale, scale, (size_t)O * 4, hipMemcpyHostToDevice, s));
    MLP_TRY(hipMemcpyAsync(m->ans_scale + O, shift, (size_t)O * 4, hipMemcpyHostToDevice, s));
    for (int64_t r0 = 0; r0 < n; r0 += kAnswerRows) {
        const int nb = (int)std::min<int64_t>(kAnswerRows, n - r0);
        int rc = xmc_mlp_forward(m, x + r0 * m->d.inputs, nb, m->ans_out, s);
        if (rc) return rc;
        const long e = (long)nb * O;
        hipLaunchKernelGGL(answer_values_kernel, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, m->ans_out,
                           (long)nb, m->head, O, m->ans_scale, values + r0 * O);
        MLP_TRY(hipGetLastError());
        hipLaunchKernelGGL(answer_flags_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, m->ans_out,
                           (long)nb, m->head, O, flagColumns, m->ans_scale, sigmaMax, uncertain + r0);
        MLP_TRY(hipGetLastError());
    }
    return 0;
}

int xmc_dev_alloc(int device, uint64_t bytes, void** ptr) {
    if (!ptr) return fail(-1, "NULL argument");
    *ptr = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(-4, "no GPU");
    MLP_TRY(hipSetDevice(device));
    MLP_TRY(hipMalloc(ptr, bytes > 0 ? bytes : 1));
    return 0;
}

int xmc_dev_free(void* ptr) {
    if (ptr) MLP_TRY(hipFree(ptr));
    return 0;
}

int xmc_dev_copy(void* dst, const void* src, uint64_t bytes, int kind) {
    if ((!dst || !src) && bytes) return fail(-1, "NULL argument");
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                    : kind == 3 ? hipMemcpyDeviceToDevice : hipMemcpyDefault;
    if (kind < 1 || kind > 3) return fail(-1, "copy kind must be 1, 2 or 3");
    if (bytes) MLP_TRY(hipMemcpy(dst, src, bytes, k));
    return 0;
}

int xmc_dev_sync(int device) {
    MLP_TRY(hipSetDevice(device));
    MLP_TRY(hipDeviceSynchronize());
    return 0;
}

int xmc_mlp_gemm_check(const float* a, const float* b, int M, int N, int K, float* c, int tile, int device) {
    if (!a || !b || !c) return fail(-1, "NULL argument");
    if (M <= 0 || N <= 0 || K <= 0 || M % BM || N % BN || K % BK)
        return fail(-1, "M, N and K must be multiples of 64");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(-4, "no GPU");
    MLP_TRY(hipSetDevice(device));
    std::vector<uint16_t> ha((size_t)M * K), hb((size_t)N * K);
    for (size_t i = 0; i < ha.size(); i++) ha[i] = f2bf(a[i]);
    for (size_t i = 0; i < hb.size(); i++) hb[i] = f2bf(b[i]);
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr, *bias = nullptr;
    int rc = 0;
    hipError_t e = hipMalloc((void**)&da, ha.size() * 2);
    if (e == hipSuccess) e = hipMalloc((void**)&db, hb.size() * 2);
    if (e == hipSuccess) e = hipMalloc((void**)&dc, (size_t)M * N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&bias, (size_t)N * 4);
    if (e == hipSuccess) e = hipMemcpy(da, ha.data(), ha.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, hb.data(), hb.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(bias, 0, (size_t)N * 4);
    if (e == hipSuccess) {
        GemmArgs g{};
        g.A = da;
        g.lda = K;
        g.B = db;
        g.ldb = K;
        g.M = M;
        g.N = N;
        g.kchunk = K;
        g.bias = bias;
        g.outf = dc;
        g.ldo = N;
        if (tile == 128 && (M % BM2 || N % BN2)) e = hipErrorInvalidValue;
        else e = launch_gemm<true, true, EPI_LINEAR>(g, 1, nullptr, tile == 128);
    }
    if (e == hipSuccess) e = hipMemcpy(c, dc, (size_t)M * N * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(-2, std::string("gemm check: ") + hipGetErrorString(e));
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    (void)hipFree(bias);
    return rc;
}

}  // extern "C"
