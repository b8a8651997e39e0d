// mlp.hip -- the TL surrogate's network: an MLP ensemble trained on tracer
// fields, forward and backward on gfx950 matrix cores (BF16 inputs, FP32
// accumulate).
//
// Network (one ensemble member; E members are trained side by side):
//   16 normalised inputs -> 64 features -> 6 hidden layers of 256 units, GELU
//   (erf form) -> n_out outputs (linear). Features: front 0, [sin, cos](2 pi B x)
//   of 32 projections B x; front 1, the 16 inputs themselves followed by
//   [sin, cos] of the first 24 projections. n_out = 1 gives a coordinate
//   network (TL at one receiver); n_out = K (a multiple of 32, <= 256) gives
//   the K coefficients of a field expanded in a fixed basis. Weights are kept
//   in FP32 (master, Adam state) and as a BF16 copy that the matrix products
//   read; hidden activations are stored in BF16, pre-activations in FP32.
//
// Matrix products use v_mfma_f32_32x32x16_bf16: a workgroup of 4 wavefronts
// computes a 64 x 64 tile of C = A * B, each wavefront one 32 x 32 block, with
// A and B staged through LDS 32 columns of K at a time (converted to BF16 on
// the way in, B transposed so that every lane reads its 8 K values as one
// 16-byte LDS load). Operand element (m, k) of A is A[m * sam + k * sak] and
// (k, n) of B is B[k * sbk + n * sbn], so the backward products read the
// forward tensors transposed without copies. Members are the grid's z axis.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/bhc_hip.h"

namespace {

thread_local std::string g_mlp_error;
int mfail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_mlp_error = buf;
    return -1;
}
#define MCHECK(x) \
    do { \
        hipError_t _e = (x); \
        if(_e != hipSuccess) return mfail("%s failed: %s (%s:%d)", #x, hipGetErrorString(_e), __FILE__, __LINE__); \
    } while(0)

constexpr int D_IN = 16, N_FREQ = 32, D_FF = 2 * N_FREQ, WIDTH = 256, N_HID = 6;
constexpr int N_FREQ_RAW = (D_FF - D_IN) / 2; // projections used by front 1
constexpr int TILE = 64, KSTEP = 32;

typedef uint16_t bf16_t;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// parameter layout of one member (floats): W0 [D_FF][WIDTH], b0 [WIDTH],
// W1..W5 [WIDTH][WIDTH], b1..b5 [WIDTH], Wout [WIDTH][n_out], bout [n_out]
constexpr size_t off_W(int l) { return l == 0 ? 0 : (size_t)(D_FF + 1) * WIDTH + (size_t)(l - 1) * (WIDTH + 1) * WIDTH; }
constexpr size_t off_b(int l) { return off_W(l) + (size_t)(l == 0 ? D_FF : WIDTH) * WIDTH; }
constexpr size_t OFF_WOUT = off_W(N_HID);
constexpr size_t off_bout(int n_out) { return OFF_WOUT + (size_t)WIDTH * n_out; }
constexpr size_t n_param(int n_out) { return off_bout(n_out) + n_out; }

__host__ __device__ inline bf16_t f2bf(float f)
{ // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    if((u & 0x7F800000u) == 0x7F800000u) return (bf16_t)(u >> 16); // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__device__ inline bf16_t to_bf(float v) { return f2bf(v); }
__device__ inline bf16_t to_bf(bf16_t v) { return v; }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum : int { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_ATOMIC = 2, EPI_BIAS = 3 };

struct GemmArgs {
    const void *A;
    const void *B;
    float *C;
    long long sam, sak, sbk, sbn, ldc;
    long long zA, zB, zC;     // member strides (elements)
    int M, N, K;
    const float *bias;        // EPI_BIAS_GELU, EPI_BIAS: bias[n] (member stride zbias)
    long long zbias;
    bf16_t *H;                // EPI_BIAS_GELU: GELU(C) as BF16, same layout as C
    int ksplit;               // EPI_ATOMIC: K is split over ksplit slices (grid z = members x ksplit)
};

// C = A * B (FP32); EPI_BIAS: C = A * B + bias; EPI_BIAS_GELU: C = A * B + bias
// and H = GELU(C); EPI_ATOMIC: C += A * B over K slices.
// M, N arbitrary (guarded), K a multiple of 32.
template<typename TA, typename TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g)
{
    __shared__ __align__(16) bf16_t As[TILE][KSTEP + 8];
    __shared__ __align__(16) bf16_t Bt[TILE][KSTEP + 8];
    const int ks = EPI == EPI_ATOMIC ? (int)(blockIdx.z % g.ksplit) : 0;
    const int z = EPI == EPI_ATOMIC ? (int)(blockIdx.z / g.ksplit) : (int)blockIdx.z;
    const TA *A = static_cast<const TA *>(g.A) + z * g.zA;
    const TB *B = static_cast<const TB *>(g.B) + z * g.zB;
    float *C = g.C + z * g.zC;
    const int m0 = blockIdx.y * TILE, n0 = blockIdx.x * TILE;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    int kbeg = 0, kend = g.K;
    if constexpr(EPI == EPI_ATOMIC) {
        int per = (g.K / KSTEP + g.ksplit - 1) / g.ksplit * KSTEP;
        kbeg = ks * per;
        kend = min(g.K, kbeg + per);
    }
    f32x16 acc = {};
    for(int k0 = kbeg; k0 < kend; k0 += KSTEP) {
        // stage A[m0 .. m0+63][k0 .. k0+31] and B[k0 .. +31][n0 .. +63] (transposed)
        for(int i = t; i < TILE * KSTEP; i += 256) {
            int r = i / KSTEP, c = i % KSTEP;
            int m = m0 + r;
            As[r][c] = m < g.M ? to_bf(A[m * g.sam + (long long)(k0 + c) * g.sak]) : (bf16_t)0;
        }
        for(int i = t; i < TILE * KSTEP; i += 256) {
            int kk, nn;
            if(g.sbn == 1) { // consecutive threads along n: coalesced reads
                kk = i / TILE;
                nn = i % TILE;
            } else {
                nn = i / KSTEP;
                kk = i % KSTEP;
            }
            int n = n0 + nn;
            Bt[nn][kk] = n < g.N ? to_bf(B[(long long)(k0 + kk) * g.sbk + n * g.sbn]) : (bf16_t)0;
        }
        __syncthreads();
#pragma unroll
        for(int kk = 0; kk < KSTEP; kk += 16) {
            const int kq = kk + 8 * (lane >> 5);
            bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm + (lane & 31)][kq]);
            bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bt[wn + (lane & 31)][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout: lane -> column wn + lane % 32; register r -> row
    // wm + 8 (r / 4) + 4 (lane / 32) + r % 4
    const int n = n0 + wn + (lane & 31);
    if(n >= g.N) return;
    for(int r = 0; r < 16; ++r) {
        int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if(m >= g.M) continue;
        float v = acc[r];
        if constexpr(EPI == EPI_BIAS_GELU || EPI == EPI_BIAS) v += g.bias[z * g.zbias + n];
        if constexpr(EPI == EPI_BIAS_GELU) g.H[z * g.zC + m * g.ldc + n] = f2bf(gelu(v));
        if constexpr(EPI == EPI_ATOMIC) unsafeAtomicAdd(C + m * g.ldc + n, v);
        else C[m * g.ldc + n] = v;
    }
}

// Features of normalised inputs x[b][16], stored BF16 (shared by all members).
// Front 0: F[b][2j] = sin(2 pi <B_j, x_b>), F[b][2j+1] = cos(...), j < 32.
// Front 1: F[b][d] = x[b][d] for d < 16, then sin / cos of projections j < 24
// at F[b][16 + 2j], F[b][17 + 2j]. One thread per (b, j), j < 32; under
// front 1 threads j >= 24 copy two inputs each.
__global__ void k_features(const float *x, int n, const float *Bf, int front, bf16_t *F)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n * N_FREQ) return;
    int b = i / N_FREQ, j = i % N_FREQ;
    if(front == 1 && j >= N_FREQ_RAW) {
        int d = 2 * (j - N_FREQ_RAW);
        F[b * D_FF + d] = f2bf(x[b * D_IN + d]);
        F[b * D_FF + d + 1] = f2bf(x[b * D_IN + d + 1]);
        return;
    }
    float s = 0.f;
    for(int d = 0; d < D_IN; ++d) s += Bf[j * D_IN + d] * x[b * D_IN + d];
    float sn, cs;
    sincospif(2.0f * s, &sn, &cs);
    const int o = front == 1 ? D_IN : 0;
    F[b * D_FF + o + 2 * j] = f2bf(sn);
    F[b * D_FF + o + 2 * j + 1] = f2bf(cs);
}

// Single-output layer: out[e][b] = <H[e][b][:], Wout[e]> + bout[e] (BF16
// operands, FP32 sum). np = parameters per member.
__global__ void k_out_fwd(const bf16_t *H, const bf16_t *Wbf, const float *P, int n, long long zH, long long zO,
                          long long np, float *out)
{
    int e = blockIdx.y, b = blockIdx.x * blockDim.x + threadIdx.x;
    if(b >= n) return;
    const bf16_t *h = H + e * zH + (long long)b * WIDTH;
    const bf16_t *w = Wbf + e * np + OFF_WOUT;
    float s = 0.f;
    for(int k = 0; k < WIDTH; ++k) s += bf2f(h[k]) * bf2f(w[k]);
    out[(long long)e * zO + b] = s + P[e * np + off_bout(1)];
}

// Loss and single-output backward for a batch: dOut = 2 (out - y) / n,
// dH5[e][b][k] = dOut * Wout[k] (FP32), loss[e] += sum (out - y)^2 / n.
__global__ void k_out_bwd(const float *out, const float *y, int n, const bf16_t *Wbf, float *dH, long long zH,
                          long long zO, long long np, float *dout, float *loss)
{
    int e = blockIdx.y, b = blockIdx.x;
    if(b >= n) return;
    float d = out[(long long)e * zO + b] - y[b];
    float g = 2.0f * d / (float)n;
    if(threadIdx.x == 0) {
        dout[(long long)e * zO + b] = g;
        atomicAdd(loss + e, d * d / (float)n);
    }
    const bf16_t *w = Wbf + e * np + OFF_WOUT;
    for(int k = threadIdx.x; k < WIDTH; k += blockDim.x) dH[e * zH + (long long)b * WIDTH + k] = g * bf2f(w[k]);
}

// Loss and output gradient of a multi-output batch (K outputs per point):
// dOut = 2 (out - y) / (n K), loss[e] += sum (out - y)^2 / (n K). One block of
// 256 threads per (chunk of 256 elements, member).
__global__ void k_out_loss(const float *out, const float *y, int total, long long zO, float *dout, float *loss)
{
    __shared__ float red[256];
    int e = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.f;
    if(i < total) {
        float d = out[e * zO + i] - y[i];
        dout[e * zO + i] = 2.0f * d / (float)total;
        d2 = d * d / (float)total;
    }
    red[threadIdx.x] = d2;
    __syncthreads();
    for(int h = 128; h > 0; h >>= 1) {
        if((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if(threadIdx.x == 0) atomicAdd(loss + e, red[0]);
}

constexpr int RCHUNK = 128; // batch rows per block of the column reductions

// Single output: dWout[e][k] += sum over a chunk of b of H5[e][b][k] dOut[e][b];
// dbout likewise (G's output-layer slots are zeroed first).
__global__ void k_out_wgrad(const bf16_t *H, const float *dout, int n, long long zH, long long zO, long long np,
                            float *G)
{
    int e = blockIdx.y, k = threadIdx.x;
    int b0 = blockIdx.x * RCHUNK, b1 = min(n, b0 + RCHUNK);
    float s = 0.f;
    if(k < WIDTH) {
        for(int b = b0; b < b1; ++b) s += bf2f(H[e * zH + (long long)b * WIDTH + k]) * dout[(long long)e * zO + b];
        unsafeAtomicAdd(G + e * np + OFF_WOUT + k, s);
    } else if(k == WIDTH) {
        for(int b = b0; b < b1; ++b) s += dout[(long long)e * zO + b];
        unsafeAtomicAdd(G + e * np + off_bout(1), s);
    }
}

// zero the gradient slots that are accumulated by atomics: biases, output
// layer, and the weights of the split-K products
__global__ void k_zero(float *p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) p[i] = 0.0f;
}

// dZ = dH * GELU'(Z) in place of dH (FP32).
__global__ void k_gelu_bwd(float *dH, const float *Z, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i < total) dH[i] *= gelu_d(Z[i]);
}

// db[e][c] += sum over a chunk of b of dZ[e][b][c], c < cols (one thread per
// column, rows of `cols` values).
__global__ void k_bias_grad(const float *dZ, int n, int cols, long long zZ, long long np, float *G, size_t off)
{
    int e = blockIdx.y, c = threadIdx.x;
    int b0 = blockIdx.x * RCHUNK, b1 = min(n, b0 + RCHUNK);
    if(c >= cols) return;
    float s = 0.f;
    for(int b = b0; b < b1; ++b) s += dZ[e * zZ + (long long)b * cols + c];
    unsafeAtomicAdd(G + e * np + off + c, s);
}

// Adam (FP32 master weights) and the BF16 copy the products read.
__global__ void k_adam(float *P, float *M, float *V, const float *G, bf16_t *Pbf, long long total, float lr,
                       float b1, float b2, float eps, float c1, float c2)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= total) return;
    float g = G[i];
    float m = b1 * M[i] + (1.0f - b1) * g;
    float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    float p = P[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[i] = p;
    Pbf[i] = f2bf(p);
}

__global__ void k_to_bf(const float *P, bf16_t *Pbf, long long total)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if(i < total) Pbf[i] = f2bf(P[i]);
}

// Batch gather from the resident training set: idx[b] -> x[b][16], y[b][n_out].
__global__ void k_gather(const float *X, const float *Y, const int32_t *idx, int n, int n_out, float *x, float *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = D_IN + n_out;
    if(i >= n * row) return;
    int b = i / row, d = i % row;
    long long s = idx[b];
    if(d < D_IN) x[b * D_IN + d] = X[s * D_IN + d];
    else y[(long long)b * n_out + (d - D_IN)] = Y[s * n_out + (d - D_IN)];
}

} // namespace

struct bhc_mlp {
    int device;
    hipStream_t stream;
    int members, max_batch;
    int n_out, front;          // outputs per point; feature front (0 Fourier, 1 inputs + Fourier)
    long long np;              // parameters per member
    float *P, *Mo, *Vo, *G;    // [members][np]
    bf16_t *Pbf;
    float *Bf;                 // [N_FREQ][D_IN]
    float *x, *y, *out, *dout, *loss; // y [max_batch][n_out], out and dout [members][max_batch][n_out]
    bf16_t *F;                 // [max_batch][D_FF]
    float *Z[N_HID];           // [members][max_batch][WIDTH]
    bf16_t *H[N_HID];
    float *dH, *dH2;           // [members][max_batch][WIDTH]
    float *X, *Y;              // resident training set: X [nset][16], Y [nset][n_out]
    int32_t *idx;
    long long nset;
    long long step;
    hipEvent_t e0, e1;
    double last_ms;
};

namespace {

template<typename TA, typename TB, int EPI>
void gemm(hipStream_t s, int members, const void *A, long long sam, long long sak, long long zA, const void *B,
          long long sbk, long long sbn, long long zB, float *C, long long ldc, long long zC, int M, int N, int K,
          const float *bias = nullptr, long long zbias = 0, bf16_t *H = nullptr, int ksplit = 1)
{
    GemmArgs g;
    g.A = A;
    g.B = B;
    g.C = C;
    g.sam = sam;
    g.sak = sak;
    g.sbk = sbk;
    g.sbn = sbn;
    g.ldc = ldc;
    g.zA = zA;
    g.zB = zB;
    g.zC = zC;
    g.M = M;
    g.N = N;
    g.K = K;
    g.bias = bias;
    g.zbias = zbias;
    g.H = H;
    g.ksplit = ksplit;
    dim3 grid((N + TILE - 1) / TILE, (M + TILE - 1) / TILE, members * (EPI == EPI_ATOMIC ? ksplit : 1));
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, g);
}

// forward of a batch already in m->x (n points): fills Z, H and out
void forward(bhc_mlp *m, int n)
{
    hipStream_t s = m->stream;
    const int E = m->members;
    const long long np = m->np;
    const long long zH = (long long)m->max_batch * WIDTH, zO = (long long)m->max_batch * m->n_out;
    hipLaunchKernelGGL(k_features, dim3((n * N_FREQ + 255) / 256), dim3(256), 0, s, m->x, n, m->Bf, m->front, m->F);
    for(int l = 0; l < N_HID; ++l) {
        const int K = l == 0 ? D_FF : WIDTH;
        const void *A = l == 0 ? (const void *)m->F : (const void *)m->H[l - 1];
        const long long zA = l == 0 ? 0 : zH;
        gemm<bf16_t, bf16_t, EPI_BIAS_GELU>(s, E, A, K, 1, zA, m->Pbf + off_W(l), WIDTH, 1, np, m->Z[l], WIDTH, zH, n,
                                            WIDTH, K, m->P + off_b(l), np, m->H[l]);
    }
    if(m->n_out == 1) {
        hipLaunchKernelGGL(k_out_fwd, dim3((n + 255) / 256, E), dim3(256), 0, s, m->H[N_HID - 1], m->Pbf, m->P, n, zH,
                           zO, np, m->out);
    } else { // out = H5 Wout + bout: M = n, N = n_out, K = WIDTH
        gemm<bf16_t, bf16_t, EPI_BIAS>(s, E, m->H[N_HID - 1], WIDTH, 1, zH, m->Pbf + OFF_WOUT, m->n_out, 1, np, m->out,
                                       m->n_out, zO, n, m->n_out, WIDTH, m->P + off_bout(m->n_out), np);
    }
}

// backward of the batch after forward(): gradients into m->G, loss into m->loss
void backward(bhc_mlp *m, int n)
{
    hipStream_t s = m->stream;
    const int E = m->members, K_out = m->n_out;
    const long long np = m->np;
    const long long zH = (long long)m->max_batch * WIDTH, zO = (long long)m->max_batch * K_out;
    const long long npt = (long long)E * np;
    const int nchunk = (n + RCHUNK - 1) / RCHUNK;
    const int ksplit = std::max(1, std::min(64, n / 512));
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, s, m->G, npt);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, m->loss, (long long)E);
    if(K_out == 1) {
        hipLaunchKernelGGL(k_out_bwd, dim3(n, E), dim3(256), 0, s, m->out, m->y, n, m->Pbf, m->dH, zH, zO, np, m->dout,
                           m->loss);
        hipLaunchKernelGGL(k_out_wgrad, dim3(nchunk, E), dim3(WIDTH + 64), 0, s, m->H[N_HID - 1], m->dout, n, zH, zO,
                           np, m->G);
    } else {
        const int total = n * K_out;
        hipLaunchKernelGGL(k_out_loss, dim3((total + 255) / 256, E), dim3(256), 0, s, m->out, m->y, total, zO,
                           m->dout, m->loss);
        // dWout = H5^T dOut : M = WIDTH, N = n_out, K = n
        gemm<bf16_t, float, EPI_ATOMIC>(s, E, m->H[N_HID - 1], 1, WIDTH, zH, m->dout, K_out, 1, zO, m->G + OFF_WOUT,
                                        K_out, np, WIDTH, K_out, n, nullptr, 0, nullptr, ksplit);
        hipLaunchKernelGGL(k_bias_grad, dim3(nchunk, E), dim3(WIDTH), 0, s, m->dout, n, K_out, zO, np, m->G,
                           off_bout(K_out));
        // dH5 = dOut Wout^T : M = n, N = WIDTH, K = n_out
        gemm<float, bf16_t, EPI_STORE>(s, E, m->dout, K_out, 1, zO, m->Pbf + OFF_WOUT, 1, K_out, np, m->dH, WIDTH, zH,
                                       n, WIDTH, K_out);
    }
    float *dH = m->dH, *dHn = m->dH2;
    for(int l = N_HID - 1; l >= 0; --l) {
        const long long tot = (long long)E * zH;
        hipLaunchKernelGGL(k_gelu_bwd, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, dH, m->Z[l], tot);
        hipLaunchKernelGGL(k_bias_grad, dim3(nchunk, E), dim3(WIDTH), 0, s, dH, n, WIDTH, zH, np, m->G, off_b(l));
        const int K_in = l == 0 ? D_FF : WIDTH;
        // dW = Hprev^T dZ : M = K_in, N = WIDTH, K = n, split over K so that
        // the 16 (or 4) output tiles per member still fill the GPU
        const void *Hp = l == 0 ? (const void *)m->F : (const void *)m->H[l - 1];
        const long long zHp = l == 0 ? 0 : zH;
        gemm<bf16_t, float, EPI_ATOMIC>(s, E, Hp, 1, K_in, zHp, dH, WIDTH, 1, zH, m->G + off_W(l), WIDTH, np, K_in,
                                        WIDTH, n, nullptr, 0, nullptr, ksplit);
        if(l > 0) {
            // dHprev = dZ W^T : M = n, N = WIDTH (K_in), K = WIDTH
            gemm<float, bf16_t, EPI_STORE>(s, E, dH, WIDTH, 1, zH, m->Pbf + off_W(l), 1, WIDTH, np, dHn, WIDTH, zH, n,
                                           K_in, WIDTH);
            float *t = dH;
            dH = dHn;
            dHn = t;
        }
    }
}

} // namespace

extern "C" const char *bhc_mlp_last_error(void) { return g_mlp_error.c_str(); }

extern "C" int bhc_mlp_num_params(void) { return (int)n_param(1); }

extern "C" long long bhc_mlp_param_count(bhc_mlp *m) { return m ? m->np : 0; }

extern "C" void bhc_mlp_destroy(bhc_mlp *m)
{
    if(!m) return;
    (void)hipSetDevice(m->device);
    (void)hipStreamSynchronize(m->stream);
    void *ptrs[] = {m->P, m->Mo, m->Vo, m->G, m->Pbf, m->Bf, m->x, m->y, m->out, m->dout, m->loss, m->F, m->dH,
                    m->dH2, m->X, m->Y, m->idx};
    for(void *p : ptrs)
        if(p) (void)hipFree(p);
    for(int l = 0; l < N_HID; ++l) {
        if(m->Z[l]) (void)hipFree(m->Z[l]);
        if(m->H[l]) (void)hipFree(m->H[l]);
    }
    (void)hipEventDestroy(m->e0);
    (void)hipEventDestroy(m->e1);
    (void)hipStreamDestroy(m->stream);
    delete m;
}

// Create an ensemble of `members` networks with n_out outputs (1, or a
// multiple of 32 up to 256) and feature front `front` (0 or 1), for batches of
// up to max_batch points (a multiple of 32). Weights: He-uniform from `seed`,
// biases and the output layer 0; the projection matrix has N(0, sigma^2)
// entries from the same seed.
extern "C" int bhc_mlp_create_ex(int device, int members, int max_batch, int n_out, int front,
                                 unsigned long long seed, float sigma, bhc_mlp **out)
{
    *out = nullptr;
    if(members < 1 || members > 16 || max_batch < 32 || max_batch % 32) return mfail("bhc_mlp_create: bad sizes");
    if(!(n_out == 1 || (n_out % 32 == 0 && n_out <= WIDTH))) return mfail("bhc_mlp_create: n_out must be 1 or a multiple of 32 <= 256");
    if(front != 0 && front != 1) return mfail("bhc_mlp_create: front must be 0 or 1");
    MCHECK(hipSetDevice(device));
    bhc_mlp *m = new bhc_mlp();
    memset(m, 0, sizeof(*m));
    m->device = device;
    m->members = members;
    m->max_batch = max_batch;
    m->n_out = n_out;
    m->front = front;
    m->np = (long long)n_param(n_out);
    MCHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    MCHECK(hipEventCreate(&m->e0));
    MCHECK(hipEventCreate(&m->e1));
    size_t np = (size_t)members * m->np;
    size_t act = (size_t)members * max_batch * WIDTH;
    size_t outs = (size_t)max_batch * n_out;
    bool bad = hipMalloc((void **)&m->P, np * 4) || hipMalloc((void **)&m->Mo, np * 4)
        || hipMalloc((void **)&m->Vo, np * 4) || hipMalloc((void **)&m->G, np * 4)
        || hipMalloc((void **)&m->Pbf, np * 2) || hipMalloc((void **)&m->Bf, N_FREQ * D_IN * 4)
        || hipMalloc((void **)&m->x, (size_t)max_batch * D_IN * 4) || hipMalloc((void **)&m->y, outs * 4)
        || hipMalloc((void **)&m->out, (size_t)members * outs * 4)
        || hipMalloc((void **)&m->dout, (size_t)members * outs * 4) || hipMalloc((void **)&m->loss, 16 * 4)
        || hipMalloc((void **)&m->F, (size_t)max_batch * D_FF * 2) || hipMalloc((void **)&m->dH, act * 4)
        || hipMalloc((void **)&m->dH2, act * 4) || hipMalloc((void **)&m->idx, (size_t)max_batch * 4);
    for(int l = 0; l < N_HID && !bad; ++l)
        bad = hipMalloc((void **)&m->Z[l], act * 4) || hipMalloc((void **)&m->H[l], act * 2);
    if(bad) {
        bhc_mlp_destroy(m);
        return mfail("bhc_mlp_create: out of device memory");
    }
    std::mt19937_64 rng(seed);
    std::vector<float> P(np, 0.0f), Bf(N_FREQ * D_IN);
    std::normal_distribution<double> nd(0.0, 1.0);
    for(auto &b : Bf) b = (float)(sigma * nd(rng));
    // hidden layers: He-uniform; the output layer starts at zero, so that a
    // network trained on residuals starts from the reference it corrects
    for(int e = 0; e < members; ++e) {
        for(int l = 0; l < N_HID; ++l) {
            int fan_in = l == 0 ? D_FF : WIDTH;
            double a = std::sqrt(6.0 / fan_in);
            std::uniform_real_distribution<double> ud(-a, a);
            size_t o = (size_t)e * m->np + off_W(l);
            for(int i = 0; i < fan_in * WIDTH; ++i) P[o + i] = (float)ud(rng);
        }
    }
    MCHECK(hipMemcpy(m->P, P.data(), np * 4, hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(m->Bf, Bf.data(), Bf.size() * 4, hipMemcpyHostToDevice));
    MCHECK(hipMemset(m->Mo, 0, np * 4));
    MCHECK(hipMemset(m->Vo, 0, np * 4));
    hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, m->stream, m->P, m->Pbf,
                       (long long)np);
    MCHECK(hipStreamSynchronize(m->stream));
    *out = m;
    return 0;
}

// A coordinate network (one output, Fourier front).
extern "C" int bhc_mlp_create(int device, int members, int max_batch, unsigned long long seed, float sigma,
                              bhc_mlp **out)
{
    return bhc_mlp_create_ex(device, members, max_batch, 1, 0, seed, sigma, out);
}

extern "C" int bhc_mlp_get_params(bhc_mlp *m, float *params, float *fourier)
{
    MCHECK(hipSetDevice(m->device));
    MCHECK(hipStreamSynchronize(m->stream));
    if(params) MCHECK(hipMemcpy(params, m->P, (size_t)m->members * m->np * 4, hipMemcpyDeviceToHost));
    if(fourier) MCHECK(hipMemcpy(fourier, m->Bf, N_FREQ * D_IN * 4, hipMemcpyDeviceToHost));
    return 0;
}

// Replace the Fourier projection matrix B [32][16] (features sin / cos of
// 2 pi B x), e.g. with per-input frequency scales.
extern "C" int bhc_mlp_set_fourier(bhc_mlp *m, const float *B)
{
    MCHECK(hipSetDevice(m->device));
    MCHECK(hipMemcpy(m->Bf, B, N_FREQ * D_IN * 4, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int bhc_mlp_set_params(bhc_mlp *m, const float *params)
{
    MCHECK(hipSetDevice(m->device));
    size_t np = (size_t)m->members * m->np;
    MCHECK(hipMemcpy(m->P, params, np * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, m->stream, m->P, m->Pbf,
                       (long long)np);
    MCHECK(hipStreamSynchronize(m->stream));
    return 0;
}

// Predictions of every member for n points x[n][16] (host); out[members][n][n_out].
extern "C" int bhc_mlp_predict(bhc_mlp *m, const float *x, long long n, float *out)
{
    MCHECK(hipSetDevice(m->device));
    const int K = m->n_out;
    const long long zO = (long long)m->max_batch * K;
    std::vector<float> tmp((size_t)m->members * zO);
    for(long long b0 = 0; b0 < n; b0 += m->max_batch) {
        int nb = (int)std::min<long long>(m->max_batch, n - b0);
        MCHECK(hipMemcpyAsync(m->x, x + b0 * D_IN, (size_t)nb * D_IN * 4, hipMemcpyHostToDevice, m->stream));
        forward(m, nb);
        MCHECK(hipMemcpyAsync(tmp.data(), m->out, (size_t)m->members * zO * 4, hipMemcpyDeviceToHost, m->stream));
        MCHECK(hipStreamSynchronize(m->stream));
        for(int e = 0; e < m->members; ++e)
            memcpy(out + ((size_t)e * n + b0) * K, tmp.data() + (size_t)e * zO, (size_t)nb * K * 4);
    }
    return 0;
}

// Gradients (FP32, [members][params]) and per-member loss for one batch
// (n <= max_batch points, x [n][16], y [n][n_out], host arrays), without
// updating the weights.
extern "C" int bhc_mlp_gradients(bhc_mlp *m, const float *x, const float *y, int n, float *grads, float *loss)
{
    if(n < 1 || n > m->max_batch || n % 32) return mfail("bhc_mlp_gradients: n must be a multiple of 32 <= max_batch");
    MCHECK(hipSetDevice(m->device));
    MCHECK(hipMemcpyAsync(m->x, x, (size_t)n * D_IN * 4, hipMemcpyHostToDevice, m->stream));
    MCHECK(hipMemcpyAsync(m->y, y, (size_t)n * m->n_out * 4, hipMemcpyHostToDevice, m->stream));
    forward(m, n);
    backward(m, n);
    MCHECK(hipGetLastError());
    if(grads)
        MCHECK(hipMemcpyAsync(grads, m->G, (size_t)m->members * m->np * 4, hipMemcpyDeviceToHost, m->stream));
    if(loss) MCHECK(hipMemcpyAsync(loss, m->loss, (size_t)m->members * 4, hipMemcpyDeviceToHost, m->stream));
    MCHECK(hipStreamSynchronize(m->stream));
    return 0;
}

// Resident training set: X[n][16] normalised inputs, Y[n][n_out] normalised targets.
extern "C" int bhc_mlp_set_dataset(bhc_mlp *m, const float *X, const float *Y, long long n)
{
    MCHECK(hipSetDevice(m->device));
    if(m->X) (void)hipFree(m->X);
    if(m->Y) (void)hipFree(m->Y);
    m->X = m->Y = nullptr;
    MCHECK(hipMalloc((void **)&m->X, (size_t)n * D_IN * 4));
    MCHECK(hipMalloc((void **)&m->Y, (size_t)n * m->n_out * 4));
    MCHECK(hipMemcpy(m->X, X, (size_t)n * D_IN * 4, hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(m->Y, Y, (size_t)n * m->n_out * 4, hipMemcpyHostToDevice));
    m->nset = n;
    return 0;
}

// `steps` Adam steps on batches of `batch` points drawn (with replacement,
// from `seed`) from the resident training set. loss_out[members]: mean loss
// of the last step. The steps run back to back on the GPU; the kernel time
// is kept for bhc_mlp_last_ms.
extern "C" int bhc_mlp_train(bhc_mlp *m, int steps, int batch, float lr, unsigned long long seed, float *loss_out)
{
    if(!m->X || batch < 32 || batch > m->max_batch || batch % 32) return mfail("bhc_mlp_train: bad batch or no dataset");
    MCHECK(hipSetDevice(m->device));
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<long long> ud(0, m->nset - 1);
    std::vector<int32_t> idx((size_t)steps * batch);
    for(auto &v : idx) v = (int32_t)ud(rng);
    int32_t *d_idx = nullptr;
    MCHECK(hipMalloc((void **)&d_idx, idx.size() * 4));
    MCHECK(hipMemcpy(d_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const long long np = (long long)m->members * m->np;
    MCHECK(hipEventRecord(m->e0, m->stream));
    for(int s = 0; s < steps; ++s) {
        hipLaunchKernelGGL(k_gather, dim3((batch * (D_IN + m->n_out) + 255) / 256), dim3(256), 0, m->stream, m->X,
                           m->Y, d_idx + (size_t)s * batch, batch, m->n_out, m->x, m->y);
        forward(m, batch);
        backward(m, batch);
        ++m->step;
        float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, m->stream, m->P, m->Mo, m->Vo,
                           m->G, m->Pbf, np, lr, b1, b2, eps, c1, c2);
    }
    MCHECK(hipEventRecord(m->e1, m->stream));
    MCHECK(hipGetLastError());
    if(loss_out) MCHECK(hipMemcpyAsync(loss_out, m->loss, (size_t)m->members * 4, hipMemcpyDeviceToHost, m->stream));
    MCHECK(hipStreamSynchronize(m->stream));
    float ms = 0.f;
    MCHECK(hipEventElapsedTime(&ms, m->e0, m->e1));
    m->last_ms = ms;
    (void)hipFree(d_idx);
    return 0;
}

extern "C" double bhc_mlp_last_ms(bhc_mlp *m) { return m ? m->last_ms : 0.0; }

// Matrix-core FLOPs of one training step on `batch` points (forward +
// backward products of the hidden layers and, with several outputs, of the
// output layer; all members).
extern "C" double bhc_mlp_step_flops(bhc_mlp *m, int batch)
{
    double f = 0.0;
    for(int l = 0; l < N_HID; ++l) {
        double k = l == 0 ? D_FF : WIDTH;
        f += 2.0 * batch * k * WIDTH * (l == 0 ? 2.0 : 3.0); // forward, dW, (dH for l > 0)
    }
    if(m->n_out > 1) f += 2.0 * batch * WIDTH * m->n_out * 3.0;
    return f * m->members;
}
