// cs_mlp.hip -- the calibration emulator: a multilayer perceptron
// n_in -> n_hidden -> n_hidden -> n_out (GELU between layers, linear output)
// trained and evaluated on the GPU with hand-written BF16 MFMA kernels.
//
// Numerics: weights and activations are BF16 (round to nearest even);
// every product is accumulated in FP32 by v_mfma_f32_16x16x32_bf16; the
// pre-activations, the loss and the gradients of the weights are FP32; the
// optimiser (Adam) keeps FP32 master weights and moments and refreshes the
// BF16 copies after every step.
//
// GEMM: C[M x N] = A[M x K] . B[K x N] with arbitrary element strides for A
// and B, so one kernel serves the forward product (X . W^T), the input
// gradient (dZ . W) and the weight gradient (dZ^T . X).  A 256-thread block
// computes a 64 x 64 tile with four waves of 2 x 2 MFMA 16x16x32 tiles (the
// gfx950 BF16 shape: eight K values per lane); the K loop stages 64 x 32
// slices of A and B through LDS, with one 16-byte load per thread and
// operand where the operand is contiguous along K (X and W^T of the forward
// product, dZ of the input gradient) and element loads elsewhere.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/covidsim_hip.h"

// Device memory of the allocation-free forward (cs_mlp_forward_noalloc): the
// BF16 input and the two BF16 hidden activations of one batch.  Every GPU the
// library runs on holds its own copy for the life of the process, and every
// call on one GPU uses that GPU's copy (calls that overlap on one GPU race).
// Host code never passes the name `scratch` to a kernel: on the host it is a
// stand-in, not a device address.  cs_mlp_create looks up the device address
// once per device (hipGetSymbolAddress, with that device current) and the
// forward uses only that address.
static __device__ __align__(256) unsigned char scratch[CS_MLP_NOALLOC_SCRATCH_BYTES];

namespace {

#define CS_MAX_DEVICES 64
std::mutex g_scratch_mu;
void* g_scratch_addr[CS_MAX_DEVICES];  // device address of `scratch` on each device, once looked up

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b)
{
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// GELU(z) = z Phi(z) with Phi(z) = (1 + erf(z / sqrt 2)) / 2 = erfc(-z / sqrt 2) / 2;
// the erfc form keeps Phi accurate for large negative z, where 1 + erf
// would cancel in FP32
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * erfcf(-z * 0.70710678118654752f); }
__device__ __forceinline__ float gelu_grad(float z)
{
    return 0.5f * erfcf(-z * 0.70710678118654752f) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

struct GemmArgs {
    int M, N, K;
    const uint16_t* A;
    long long a_m, a_k;
    const uint16_t* B;
    long long b_k, b_n;
    const float* bias;      // [N] added to the product, or null
    int act;                // 1: the BF16 output is GELU(product + bias)
    const float* gz;        // backward: the BF16 output is product * GELU'(gz[m][n])
    long long gz_ld;
    float* Cf;              // FP32 output (pre-activation or weight gradient), or null
    long long cf_ld;
    uint16_t* Cb;           // BF16 output, or null
    long long cb_ld;
};

#define TM 64
#define TN 64
#define TK 32
#define LDK (TK + 8)  // padded row of a staged slice: 80 bytes, so every 16-byte fragment read stays aligned

// one thread's 8 K values of a staged row: a 16-byte load where the operand
// is contiguous along K and the 8 values lie inside it, else element loads
// (zero outside the matrix)
__device__ __forceinline__ uint4 stage8(const uint16_t* p, long long ld_row, long long ld_k, int row, int rows, int k, int K,
                                        bool vec)
{
    if (vec && row < rows && k + 8 <= K) return *reinterpret_cast<const uint4*>(p + (long long)row * ld_row + k);
    uint16_t v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = (row < rows && k + i < K) ? p[(long long)row * ld_row + (long long)(k + i) * ld_k] : (uint16_t)0;
    uint4 u;
    memcpy(&u, v, 16);
    return u;
}

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g)
{
    __shared__ __align__(16) uint16_t As[TM][LDK];   // As[m][k]
    __shared__ __align__(16) uint16_t Bt[TN][LDK];   // Bt[n][k]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int bm = blockIdx.y * TM, bn = blockIdx.x * TN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    // an operand takes 16-byte loads when it is contiguous along K with rows
    // starting on 16-byte boundaries (k0 and lk are multiples of 8)
    const bool a_vec = g.a_k == 1 && g.a_m % 8 == 0 && ((uintptr_t)g.A & 15) == 0;
    const bool b_vec = g.b_k == 1 && g.b_n % 8 == 0 && ((uintptr_t)g.B & 15) == 0;
    const int lr = tid >> 2, lk = (tid & 3) * 8;  // this thread stages row/column lr, k = lk..lk+7
    for (int k0 = 0; k0 < g.K; k0 += TK) {
        *reinterpret_cast<uint4*>(&As[lr][lk]) = stage8(g.A, g.a_m, g.a_k, bm + lr, g.M, k0 + lk, g.K, a_vec);
        *reinterpret_cast<uint4*>(&Bt[lr][lk]) = stage8(g.B, g.b_n, g.b_k, bn + lr, g.N, k0 + lk, g.K, b_vec);
        __syncthreads();
        // fragments: lane l holds row/column (l % 16) and k = 8 * (l / 16) + 0..7
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 af[2], bfr[2];
        for (int a = 0; a < 2; a++) af[a] = *reinterpret_cast<const bf16x8*>(&As[wm + a * 16 + fr][fk]);
        for (int b = 0; b < 2; b++) bfr[b] = *reinterpret_cast<const bf16x8*>(&Bt[wn + b * 16 + fr][fk]);
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // result: lane l holds column (l % 16), rows 4 * (l / 16) + 0..3 of each 16 x 16 tile
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const int n = bn + wn + b * 16 + (lane & 15);
            if (n >= g.N) continue;
            for (int i = 0; i < 4; i++) {
                const int m = bm + wm + a * 16 + (lane >> 4) * 4 + i;
                if (m >= g.M) continue;
                float v = acc[a][b][i];
                if (g.bias) v += g.bias[n];
                if (g.Cf) g.Cf[(long long)m * g.cf_ld + n] = v;
                if (g.Cb) {
                    float o = v;
                    if (g.act) o = gelu(v);
                    if (g.gz) o = v * gelu_grad(g.gz[(long long)m * g.gz_ld + n]);
                    g.Cb[(long long)m * g.cb_ld + n] = f2bf(o);
                }
            }
        }
}

__global__ void k_to_bf16(const float* in, uint16_t* out, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = f2bf(in[i]);
}

// batch rows idx[0..b) of the resident training set
__global__ void k_gather(const uint16_t* X, const float* T, const int32_t* idx, int b, int n_in, int n_out,
                         uint16_t* xb, float* tb)
{
    const long long tot = (long long)b * (n_in + n_out);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(i / (n_in + n_out)), c = (int)(i % (n_in + n_out));
        const long long src = idx[r];
        if (c < n_in) xb[(long long)r * n_in + c] = X[src * n_in + c];
        else tb[(long long)r * n_out + (c - n_in)] = T[src * n_out + (c - n_in)];
    }
}

// mean squared error over b x n outputs: dY = 2 (y - t) / (b n) in BF16, and
// per-block partial sums of the squared errors
__global__ void __launch_bounds__(256) k_mse(const float* y, const float* t, long long n_all, uint16_t* dy, double* part)
{
    __shared__ double red[256];
    double s = 0;
    const float sc = 2.0f / (float)n_all;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_all; i += (long long)gridDim.x * blockDim.x) {
        const float e = y[i] - t[i];
        s += (double)e * (double)e;
        dy[i] = f2bf(sc * e);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// bias gradient: column sums of the BF16 output gradient dz[b][n]
__global__ void k_colsum(const uint16_t* dz, int b, int n, float* out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float s = 0;
    for (int r = 0; r < b; r++) s += bf2f(dz[(long long)r * n + c]);
    out[c] = s;
}

// Adam (Kingma and Ba 2015) on the FP32 masters; refreshes the BF16 weights
// (the bias vectors stay FP32: bf_out is null for them)
__global__ void k_adam(float* p, const float* gr, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, uint16_t* bf_out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float g = gr[i];
        const float mi = b1 * m[i] + (1.0f - b1) * g;
        const float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        const float pi = p[i] - upd;
        p[i] = pi;
        if (bf_out) bf_out[i] = f2bf(pi);
    }
}

unsigned blocks_for(long long n, int bs = 256) { return (unsigned)std::max<long long>(1, std::min<long long>((n + bs - 1) / bs, 8192)); }

}  // namespace

struct cs_mlp {
    char err[512];
    int device = 0;
    hipStream_t stream = nullptr;
    int n_in = 0, n_h = 0, n_out = 0;
    // parameters: FP32 masters (one block: W1 b1 W2 b2 W3 b3), Adam moments, gradients
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    long long n_par = 0, off[6] = {0};
    uint16_t *w1b = nullptr, *w2b = nullptr, *w3b = nullptr;
    long long adam_t = 0;
    // batch buffers (capacity cap rows)
    int cap = 0;
    uint16_t *xb = nullptr, *h1 = nullptr, *h2 = nullptr, *dy = nullptr, *dz2 = nullptr, *dz1 = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *tb = nullptr;
    double* part = nullptr;
    int32_t* idx = nullptr;
    // resident training set
    uint16_t* X = nullptr;
    float* T = nullptr;
    int n_data = 0;
    // this device's address of the allocation-free forward's scratch array
    unsigned char* scratch = nullptr;
};

namespace {

int mset(cs_mlp* m, int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(m->err, sizeof(m->err), fmt, ap);
    va_end(ap);
    return code;
}

#define MCHK(m, x)                                                                      \
    do {                                                                                \
        hipError_t _e = (x);                                                            \
        if (_e != hipSuccess) return mset(m, CS_EHIP, "%s: %s", #x, hipGetErrorString(_e)); \
    } while (0)

template <class T>
int malloc_n(cs_mlp* m, T** p, long long n)
{
    void* q = nullptr;
    if (hipMalloc(&q, (size_t)std::max<long long>(n, 1) * sizeof(T)) != hipSuccess)
        return mset(m, CS_ENOMEM, "hipMalloc of %lld elements failed", n);
    *p = (T*)q;
    return 0;
}

void gemm(cs_mlp* m, const GemmArgs& a)
{
    dim3 grid((unsigned)((a.N + TN - 1) / TN), (unsigned)((a.M + TM - 1) / TM));
    k_gemm_bf16<<<grid, 256, 0, m->stream>>>(a);
}

GemmArgs ga(int M, int N, int K, const uint16_t* A, long long am, long long ak, const uint16_t* B, long long bk, long long bn)
{
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.a_m = am; g.a_k = ak;
    g.B = B; g.b_k = bk; g.b_n = bn;
    return g;
}

int ensure_batch(cs_mlp* m, int b)
{
    if (b <= m->cap) return 0;
    void* old[] = {m->xb, m->h1, m->h2, m->dy, m->dz2, m->dz1, m->z1, m->z2, m->y, m->tb, m->idx};
    for (void* q : old)
        if (q) (void)hipFree(q);
    int rc;
    const long long B = b;
    if ((rc = malloc_n(m, &m->xb, B * m->n_in)) || (rc = malloc_n(m, &m->h1, B * m->n_h)) || (rc = malloc_n(m, &m->h2, B * m->n_h)) ||
        (rc = malloc_n(m, &m->dy, B * m->n_out)) || (rc = malloc_n(m, &m->dz2, B * m->n_h)) || (rc = malloc_n(m, &m->dz1, B * m->n_h)) ||
        (rc = malloc_n(m, &m->z1, B * m->n_h)) || (rc = malloc_n(m, &m->z2, B * m->n_h)) || (rc = malloc_n(m, &m->y, B * m->n_out)) ||
        (rc = malloc_n(m, &m->tb, B * m->n_out)) || (rc = malloc_n(m, &m->idx, B)))
        return rc;
    m->cap = b;
    return 0;
}

// forward of a BF16 batch xb (b rows) into the given buffers: the FP32
// pre-activations z1, z2 (either may be null: not kept), the BF16
// activations h1, h2 and the FP32 output y
void forward_into(cs_mlp* m, int b, const uint16_t* xb, float* z1, uint16_t* h1, float* z2, uint16_t* h2, float* y)
{
    const int I = m->n_in, H = m->n_h, O = m->n_out;
    GemmArgs a = ga(b, H, I, xb, I, 1, m->w1b, 1, I);
    a.bias = m->p + m->off[1]; a.act = 1; a.Cf = z1; a.cf_ld = H; a.Cb = h1; a.cb_ld = H;
    gemm(m, a);
    a = ga(b, H, H, h1, H, 1, m->w2b, 1, H);
    a.bias = m->p + m->off[3]; a.act = 1; a.Cf = z2; a.cf_ld = H; a.Cb = h2; a.cb_ld = H;
    gemm(m, a);
    a = ga(b, O, H, h2, H, 1, m->w3b, 1, H);
    a.bias = m->p + m->off[5]; a.Cf = y; a.cf_ld = O;
    gemm(m, a);
}

// forward of the staged batch xb (b rows): z1, h1, z2, h2, y
void forward(cs_mlp* m, int b) { forward_into(m, b, m->xb, m->z1, m->h1, m->z2, m->h2, m->y); }

// bytes of scratch the allocation-free forward needs for b rows: the BF16
// input, then the two BF16 hidden activations, each 256-byte aligned
size_t noalloc_bytes(const cs_mlp* m, long long b)
{
    return align256((size_t)b * m->n_in * 2) + 2 * align256((size_t)b * m->n_h * 2);
}

// this device's address of `scratch`, looked up once per device with the
// device current (the caller has made it current)
int scratch_addr(cs_mlp* m)
{
    if (m->device >= CS_MAX_DEVICES) return mset(m, CS_EINVAL, "device %d beyond the %d the library tracks", m->device, CS_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[m->device]) {
        void* p = nullptr;
        const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess || !p) return mset(m, CS_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(st));
        g_scratch_addr[m->device] = p;
    }
    m->scratch = (unsigned char*)g_scratch_addr[m->device];
    return 0;
}

// loss and gradients of the staged batch (targets in tb) into m->g; returns the loss
int backward(cs_mlp* m, int b, double* loss)
{
    const int I = m->n_in, H = m->n_h, O = m->n_out;
    const long long nall = (long long)b * O;
    const unsigned nb = blocks_for(nall);
    k_mse<<<nb, 256, 0, m->stream>>>(m->y, m->tb, nall, m->dy, m->part);
    float* g = m->g;
    // layer 3: dW3 = dY^T . H2, db3 = colsum dY, dZ2 = (dY . W3) * GELU'(z2)
    GemmArgs a = ga(O, H, b, m->dy, 1, O, m->h2, H, 1);
    a.Cf = g + m->off[4]; a.cf_ld = H;
    gemm(m, a);
    k_colsum<<<(O + 255) / 256, 256, 0, m->stream>>>(m->dy, b, O, g + m->off[5]);
    a = ga(b, H, O, m->dy, O, 1, m->w3b, H, 1);
    a.gz = m->z2; a.gz_ld = H; a.Cb = m->dz2; a.cb_ld = H;
    gemm(m, a);
    // layer 2
    a = ga(H, H, b, m->dz2, 1, H, m->h1, H, 1);
    a.Cf = g + m->off[2]; a.cf_ld = H;
    gemm(m, a);
    k_colsum<<<(H + 255) / 256, 256, 0, m->stream>>>(m->dz2, b, H, g + m->off[3]);
    a = ga(b, H, H, m->dz2, H, 1, m->w2b, H, 1);
    a.gz = m->z1; a.gz_ld = H; a.Cb = m->dz1; a.cb_ld = H;
    gemm(m, a);
    // layer 1
    a = ga(H, I, b, m->dz1, 1, H, m->xb, I, 1);
    a.Cf = g + m->off[0]; a.cf_ld = I;
    gemm(m, a);
    k_colsum<<<(H + 255) / 256, 256, 0, m->stream>>>(m->dz1, b, H, g + m->off[1]);
    MCHK(m, hipGetLastError());
    if (loss) {
        std::vector<double> part(nb);
        MCHK(m, hipMemcpyAsync(part.data(), m->part, nb * sizeof(double), hipMemcpyDeviceToHost, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
        double s = 0;
        for (double x : part) s += x;
        *loss = s / (double)nall;
    }
    return 0;
}

int refresh_bf16(cs_mlp* m)
{
    const long long n1 = (long long)m->n_h * m->n_in, n2 = (long long)m->n_h * m->n_h, n3 = (long long)m->n_out * m->n_h;
    k_to_bf16<<<blocks_for(n1), 256, 0, m->stream>>>(m->p + m->off[0], m->w1b, n1);
    k_to_bf16<<<blocks_for(n2), 256, 0, m->stream>>>(m->p + m->off[2], m->w2b, n2);
    k_to_bf16<<<blocks_for(n3), 256, 0, m->stream>>>(m->p + m->off[4], m->w3b, n3);
    MCHK(m, hipGetLastError());
    return 0;
}

int adam_step(cs_mlp* m, float lr)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m->adam_t++;
    const float c1 = 1.0f - powf(b1, (float)m->adam_t), c2 = 1.0f - powf(b2, (float)m->adam_t);
    uint16_t* bfo[6] = {m->w1b, nullptr, m->w2b, nullptr, m->w3b, nullptr};
    for (int k = 0; k < 6; k++) {
        const long long n = (k < 5 ? m->off[k + 1] : m->n_par) - m->off[k];
        k_adam<<<blocks_for(n), 256, 0, m->stream>>>(m->p + m->off[k], m->g + m->off[k], m->m + m->off[k], m->v + m->off[k], n, lr,
                                                      b1, b2, eps, c1, c2, bfo[k]);
    }
    MCHK(m, hipGetLastError());
    return 0;
}

// x host [b][n_in] -> staged BF16 batch
int stage_x(cs_mlp* m, const float* x, int b)
{
    std::vector<uint16_t> h((size_t)b * m->n_in);
    for (size_t i = 0; i < h.size(); i++) h[i] = f2bf(x[i]);
    MCHK(m, hipMemcpyAsync(m->xb, h.data(), h.size() * 2, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

extern "C" {

const char* cs_mlp_last_error(const cs_mlp* m) { return m ? m->err : "null emulator"; }

int cs_mlp_create(cs_mlp** out, int device, int n_in, int n_hidden, int n_out, uint64_t seed)
{
    if (!out || n_in <= 0 || n_hidden <= 0 || n_out <= 0) return CS_EINVAL;
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return CS_EHIP;
    cs_mlp* m = new cs_mlp();
    m->err[0] = 0;
    m->device = device;
    m->n_in = n_in; m->n_h = n_hidden; m->n_out = n_out;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&m->stream) != hipSuccess) { delete m; return CS_EHIP; }
    int rc;
    if ((rc = scratch_addr(m))) {  // set-up of the allocation-free forward on this device
        *out = m;
        return rc;
    }
    const long long sz[6] = {(long long)n_hidden * n_in, n_hidden, (long long)n_hidden * n_hidden, n_hidden,
                             (long long)n_out * n_hidden, n_out};
    long long o = 0;
    for (int k = 0; k < 6; k++) { m->off[k] = o; o += sz[k]; }
    m->n_par = o;
    if ((rc = malloc_n(m, &m->p, o)) || (rc = malloc_n(m, &m->m, o)) || (rc = malloc_n(m, &m->v, o)) || (rc = malloc_n(m, &m->g, o)) ||
        (rc = malloc_n(m, &m->w1b, sz[0])) || (rc = malloc_n(m, &m->w2b, sz[2])) || (rc = malloc_n(m, &m->w3b, sz[4])) ||
        (rc = malloc_n(m, &m->part, 8192))) {
        *out = m;
        return rc;
    }
    // He-style uniform initialisation from a counter-based hash of (seed, index); zero biases
    std::vector<float> h((size_t)o, 0.0f);
    const int fan[3] = {n_in, n_hidden, n_hidden};
    for (int L = 0; L < 3; L++) {
        const float lim = sqrtf(6.0f / (float)fan[L]);
        for (long long i = 0; i < sz[2 * L]; i++) {
            const double u = (double)(mix64(seed ^ mix64((uint64_t)(m->off[2 * L] + i))) >> 11) * (1.0 / 9007199254740992.0);
            h[(size_t)(m->off[2 * L] + i)] = (float)((2.0 * u - 1.0) * lim) * (L == 2 ? 0.5f : 1.0f);
        }
    }
    *out = m;
    MCHK(m, hipMemcpy(m->p, h.data(), (size_t)o * 4, hipMemcpyHostToDevice));
    MCHK(m, hipMemset(m->m, 0, (size_t)o * 4));
    MCHK(m, hipMemset(m->v, 0, (size_t)o * 4));
    return refresh_bf16(m);
}

void cs_mlp_destroy(cs_mlp* m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    void* ptrs[] = {m->p, m->m, m->v, m->g, m->w1b, m->w2b, m->w3b, m->xb, m->h1, m->h2, m->dy, m->dz2, m->dz1,
                    m->z1, m->z2, m->y, m->tb, m->part, m->idx, m->X, m->T};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int64_t cs_mlp_n_params(const cs_mlp* m) { return m ? m->n_par : 0; }

int cs_mlp_get_params(cs_mlp* m, float* out)
{
    if (!m || !out) return CS_EINVAL;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipStreamSynchronize(m->stream));
    MCHK(m, hipMemcpy(out, m->p, (size_t)m->n_par * 4, hipMemcpyDeviceToHost));
    return 0;
}

int cs_mlp_set_params(cs_mlp* m, const float* in)
{
    if (!m || !in) return CS_EINVAL;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipStreamSynchronize(m->stream));
    MCHK(m, hipMemcpy(m->p, in, (size_t)m->n_par * 4, hipMemcpyHostToDevice));
    MCHK(m, hipMemset(m->m, 0, (size_t)m->n_par * 4));
    MCHK(m, hipMemset(m->v, 0, (size_t)m->n_par * 4));
    m->adam_t = 0;
    return refresh_bf16(m);
}

int cs_mlp_forward(cs_mlp* m, const float* x, int batch, float* y)
{
    if (!m || !x || !y || batch <= 0) return CS_EINVAL;
    MCHK(m, hipSetDevice(m->device));
    int rc;
    if ((rc = ensure_batch(m, batch)) || (rc = stage_x(m, x, batch))) return rc;
    forward(m, batch);
    MCHK(m, hipGetLastError());
    MCHK(m, hipMemcpyAsync(y, m->y, (size_t)batch * m->n_out * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

int cs_mlp_forward_noalloc(cs_mlp* m, const float* x, int batch, float* y)
{
    if (!m || !x || !y || batch <= 0) return CS_EINVAL;
    if (!m->scratch) return mset(m, CS_ESTATE, "the allocation-free forward was not set up on device %d", m->device);
    const size_t need = noalloc_bytes(m, batch);
    if (need > (size_t)CS_MLP_NOALLOC_SCRATCH_BYTES) {
        long long most = 0;
        while (noalloc_bytes(m, most + 1) <= (size_t)CS_MLP_NOALLOC_SCRATCH_BYTES) most++;
        return mset(m, CS_EINVAL, "batch %d needs %zu bytes of scratch; the allocation-free forward holds %d (at most %lld rows at these layer widths)",
                    batch, need, CS_MLP_NOALLOC_SCRATCH_BYTES, most);
    }
    MCHK(m, hipSetDevice(m->device));
    const long long nx = (long long)batch * m->n_in;
    unsigned char* s = m->scratch;
    uint16_t* xb = (uint16_t*)s;
    uint16_t* h1 = (uint16_t*)(s + align256((size_t)nx * 2));
    uint16_t* h2 = (uint16_t*)((unsigned char*)h1 + align256((size_t)batch * m->n_h * 2));
    k_to_bf16<<<blocks_for(nx), 256, 0, m->stream>>>(x, xb, nx);  // the rounding stage_x applies on the host
    forward_into(m, batch, xb, nullptr, h1, nullptr, h2, y);
    MCHK(m, hipGetLastError());
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

int cs_mlp_forward_trace(cs_mlp* m, const float* x, int batch, float* z1, float* z2, uint16_t* h1, uint16_t* h2, float* y)
{
    if (!m || !x || batch <= 0) return CS_EINVAL;
    MCHK(m, hipSetDevice(m->device));
    int rc;
    if ((rc = ensure_batch(m, batch)) || (rc = stage_x(m, x, batch))) return rc;
    forward(m, batch);
    MCHK(m, hipGetLastError());
    const size_t bh = (size_t)batch * m->n_h;
    if (z1) MCHK(m, hipMemcpyAsync(z1, m->z1, bh * 4, hipMemcpyDeviceToHost, m->stream));
    if (z2) MCHK(m, hipMemcpyAsync(z2, m->z2, bh * 4, hipMemcpyDeviceToHost, m->stream));
    if (h1) MCHK(m, hipMemcpyAsync(h1, m->h1, bh * 2, hipMemcpyDeviceToHost, m->stream));
    if (h2) MCHK(m, hipMemcpyAsync(h2, m->h2, bh * 2, hipMemcpyDeviceToHost, m->stream));
    if (y) MCHK(m, hipMemcpyAsync(y, m->y, (size_t)batch * m->n_out * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

int cs_mlp_loss_grad(cs_mlp* m, const float* x, const float* t, int batch, double* loss, float* grad)
{
    if (!m || !x || !t || batch <= 0) return CS_EINVAL;
    MCHK(m, hipSetDevice(m->device));
    int rc;
    if ((rc = ensure_batch(m, batch)) || (rc = stage_x(m, x, batch))) return rc;
    MCHK(m, hipMemcpyAsync(m->tb, t, (size_t)batch * m->n_out * 4, hipMemcpyHostToDevice, m->stream));
    forward(m, batch);
    if ((rc = backward(m, batch, loss))) return rc;
    if (grad) {
        MCHK(m, hipMemcpyAsync(grad, m->g, (size_t)m->n_par * 4, hipMemcpyDeviceToHost, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
    }
    return 0;
}

int cs_mlp_set_data(cs_mlp* m, const float* x, const float* t, int n)
{
    if (!m || !x || !t || n <= 0) return CS_EINVAL;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipStreamSynchronize(m->stream));
    if (m->X) (void)hipFree(m->X);
    if (m->T) (void)hipFree(m->T);
    m->X = nullptr; m->T = nullptr;
    int rc;
    if ((rc = malloc_n(m, &m->X, (long long)n * m->n_in)) || (rc = malloc_n(m, &m->T, (long long)n * m->n_out))) return rc;
    std::vector<uint16_t> h((size_t)n * m->n_in);
    for (size_t i = 0; i < h.size(); i++) h[i] = f2bf(x[i]);
    MCHK(m, hipMemcpy(m->X, h.data(), h.size() * 2, hipMemcpyHostToDevice));
    MCHK(m, hipMemcpy(m->T, t, (size_t)n * m->n_out * 4, hipMemcpyHostToDevice));
    m->n_data = n;
    return 0;
}

int cs_mlp_train(cs_mlp* m, int epochs, int batch, float lr, uint64_t seed, double* loss_per_epoch)
{
    if (!m || epochs < 0 || batch <= 0) return CS_EINVAL;
    if (!m->X) return mset(m, CS_ESTATE, "no training data (cs_mlp_set_data)");
    MCHK(m, hipSetDevice(m->device));
    const int n = m->n_data;
    batch = std::min(batch, n);
    int rc;
    if ((rc = ensure_batch(m, batch))) return rc;
    std::vector<int32_t> perm(n);
    for (int e = 0; e < epochs; e++) {
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int i = n - 1; i > 0; i--) {  // Fisher-Yates from a counter-based hash
            const uint64_t r = mix64(seed ^ mix64(((uint64_t)e << 32) | (uint64_t)i));
            std::swap(perm[i], perm[(int)(r % (uint64_t)(i + 1))]);
        }
        double tot = 0;
        int nb = 0;
        for (int s = 0; s + batch <= n; s += batch) {
            MCHK(m, hipMemcpyAsync(m->idx, perm.data() + s, (size_t)batch * 4, hipMemcpyHostToDevice, m->stream));
            k_gather<<<blocks_for((long long)batch * (m->n_in + m->n_out)), 256, 0, m->stream>>>(m->X, m->T, m->idx, batch, m->n_in,
                                                                                                 m->n_out, m->xb, m->tb);
            forward(m, batch);
            double l = 0;
            if ((rc = backward(m, batch, &l))) return rc;
            if ((rc = adam_step(m, lr))) return rc;
            tot += l;
            nb++;
        }
        if (loss_per_epoch) loss_per_epoch[e] = nb ? tot / nb : 0.0;
    }
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

}  // extern "C"
