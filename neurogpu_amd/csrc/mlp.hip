/*
 * mlp.hip -- the surrogate network: BF16 MFMA GEMMs with FP32 accumulation,
 * forward and backward passes, Adam, for an ensemble of MLPs
 * nin -> H -> H -> H -> nout with ReLU.  No reference counterpart (the
 * reference GA has no surrogate); it pre-screens GA offspring.
 *
 * The members of the ensemble run side by side: every launch covers all of
 * them (grid z = member, each operand a fixed element stride apart), so a
 * training step is one gather, four forward and ten backward GEMMs, the
 * loss, column sums and one Adam update for the whole ensemble.
 *
 * GEMM: C[M,N] = sum_k A(m,k) B(k,n), A and B bf16 (either orientation), FP32
 * accumulation on v_mfma_f32_32x32x16_bf16 (gfx950): a 256-thread block holds
 * a 64x64 tile of C, each wave a 32x32 sub-tile; operands are read straight
 * from global memory (the matrices here are at most a few hundred wide) with
 * bounds checks, so no size needs padding, four k-steps ahead of the MFMA
 * that uses them.  The epilogue optionally adds a
 * bias per column, applies ReLU, multiplies by a ReLU mask (backward), and
 * writes FP32 and/or BF16.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/neurogpu.h"

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ static inline uint16_t f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40); /* NaN */
    u += 0x7fffu + ((u >> 16) & 1u);                                                              /* nearest even */
    return (uint16_t)(u >> 16);
}
__host__ __device__ static inline float bf2f(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct GemmArgs {
    int M, N, K, lda, ldb, ldc;
    const uint16_t *A, *B;
    const float *bias;          /* [N] or null */
    const uint16_t *mask;       /* [M, ldc] bf16: keep where > 0 (ReLU backward) or null */
    float *C32;                 /* fp32 out or null */
    uint16_t *C16;              /* bf16 out or null */
    int relu;
    /* element strides between the matrices of consecutive grid z (members) */
    long long sA, sB, sBias, sMask, sC;
    int vec;                    /* set by gemm(): bit 0 A, bit 1 B read as 16-byte vectors */
};

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32, n0 = blockIdx.x * 64 + (wave & 1) * 32;
    const long long z = blockIdx.z;
    g.A += z * g.sA;
    g.B += z * g.sB;
    if (g.bias) g.bias += z * g.sBias;
    if (g.mask) g.mask += z * g.sMask;
    if (g.C32) g.C32 += z * g.sC;
    if (g.C16) g.C16 += z * g.sC;
    const int row = lane & 31, kh = (lane >> 5) * 8;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
    const int am = m0 + row, bn = n0 + row;
    /* operands of k-steps k0 .. k0 + 16 (zero beyond K or outside the matrix):
     * a lane's eight k values are contiguous in memory when the operand is
     * k-minor (A not transposed, B transposed), and then, where the host has
     * checked 16-byte alignment and K % 8 == 0 (g.vec bit 0 for A, bit 1 for
     * B), read as one 16-byte load */
    auto load = [&](int k0, bf16x8 &a, bf16x8 &b) {
        const int kb = k0 + kh;
        if (!TA && (g.vec & 1)) {
            a = (am < g.M && kb < g.K) ? *(const bf16x8 *)(g.A + (size_t)am * g.lda + kb) : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int k = kb + e;
                uint16_t av = 0;
                if (am < g.M && k < g.K) av = TA ? g.A[(size_t)k * g.lda + am] : g.A[(size_t)am * g.lda + k];
                a[e] = (short)av;
            }
        }
        if (TB && (g.vec & 2)) {
            b = (bn < g.N && kb < g.K) ? *(const bf16x8 *)(g.B + (size_t)bn * g.ldb + kb) : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int k = kb + e;
                uint16_t bv = 0;
                if (bn < g.N && k < g.K) bv = TB ? g.B[(size_t)bn * g.ldb + k] : g.B[(size_t)k * g.ldb + bn];
                b[e] = (short)bv;
            }
        }
    };
    /* four k-steps in flight: a step's slot is refilled with the operands of
     * the step four ahead as soon as its MFMA has issued */
    constexpr int PF = 4;
    bf16x8 ra[PF], rb[PF];
#pragma unroll
    for (int i = 0; i < PF; i++) load(16 * i, ra[i], rb[i]);
    for (int k0 = 0; k0 < g.K; k0 += 16 * PF) {
#pragma unroll
        for (int i = 0; i < PF; i++) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ra[i], rb[i], acc, 0, 0, 0);
            load(k0 + 16 * (i + PF), ra[i], rb[i]);
        }
    }
    /* accumulator i of this lane: row 8*(i/4) + 4*(lane/32) + i%4, column lane%32 */
    const int col = n0 + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = m0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        if (r >= g.M || col >= g.N) continue;
        float v = acc[i];
        if (g.bias) v += g.bias[col];
        if (g.relu && v < 0.0f) v = 0.0f;
        if (g.mask && !(bf2f(g.mask[(size_t)r * g.ldc + col]) > 0.0f)) v = 0.0f;
        if (g.C32) g.C32[(size_t)r * g.ldc + col] = v;
        if (g.C16) g.C16[(size_t)r * g.ldc + col] = f2bf(v);
    }
}

/* A k-minor operand (rows of ld elements, matrices `stride` elements apart)
 * can be read eight k values at a time as one 16-byte load. */
static bool vec_ok(const uint16_t *p, int K, int ld, long long stride)
{
    return K % 8 == 0 && ld % 8 == 0 && stride % 8 == 0 && ((uintptr_t)p & 15) == 0;
}

static hipError_t gemm(hipStream_t s, bool ta, bool tb, const GemmArgs &g0, int nz)
{
    GemmArgs g = g0;
    g.vec = (!ta && vec_ok(g.A, g.K, g.lda, g.sA) ? 1 : 0) | (tb && vec_ok(g.B, g.K, g.ldb, g.sB) ? 2 : 0);
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, nz), block(256);
    if (!ta && tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, block, 0, s, g);
    else if (!ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, block, 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, block, 0, s, g);
    return hipGetLastError();
}

/* member blockIdx.y: its n values at x + blockIdx.y * sx -> bf16 at
 * y + blockIdx.y * sy */
__global__ void f32_to_bf16_kernel(const float *x, uint16_t *y, size_t n, size_t sx, size_t sy)
{
    x += blockIdx.y * sx;
    y += blockIdx.y * sy;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

/* dY = 2 (Y - T) / (rows * cols) in bf16 and fp32; loss partial sums per
 * block; member blockIdx.y: its n elements at offset blockIdx.y * n, its loss
 * in loss[blockIdx.y] */
__global__ void mse_grad_kernel(const float *Y, const float *T, int n, float scale, float *dY32, uint16_t *dY16,
                                float *loss)
{
    __shared__ float part[256];
    const size_t o = (size_t)blockIdx.y * n;
    Y += o; T += o; dY32 += o; dY16 += o; loss += blockIdx.y;
    float s = 0.0f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        s += d * d;
        const float gval = 2.0f * d * scale;
        dY32[i] = gval;
        dY16[i] = f2bf(gval);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] * scale);
}

/* column sums of an [rows, cols] matrix (fp32 or bf16) -> out[cols]: a
 * block of 64 columns x 16 row groups (coalesced 64-wide row reads, each
 * thread summing every 16th row), then the 16 partial sums in a fixed order;
 * member blockIdx.y: its matrix at offset blockIdx.y * rows * cols, its sums
 * at out + blockIdx.y * sout */
#define COLSUM_RG 16
__global__ void __launch_bounds__(64 * COLSUM_RG) colsum_kernel(const float *x32, const uint16_t *x16, int rows, int cols,
                                                                float *out, long long sout)
{
    __shared__ float part[COLSUM_RG][64];
    const size_t o = (size_t)blockIdx.y * rows * cols;
    if (x32) x32 += o;
    if (x16) x16 += o;
    out += blockIdx.y * sout;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + tx;
    float s = 0.0f;
    if (c < cols)
        for (int r = ty; r < rows; r += COLSUM_RG) s += x32 ? x32[(size_t)r * cols + c] : bf2f(x16[(size_t)r * cols + c]);
    part[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && c < cols) {
        float t = 0.0f;
        for (int k = 0; k < COLSUM_RG; k++) t += part[k][tx];
        out[c] = t;
    }
}

/* member blockIdx.y: its n parameters at offset blockIdx.y * n, its bf16
 * copy at blockIdx.y * n16 */
__global__ void adam_kernel(float *w, const float *gr, float *m, float *v, uint16_t *w16, int n, int n16, float lr,
                            float b1, float b2, float eps, float c1, float c2)
{
    const size_t o = blockIdx.y * (size_t)n;
    w += o; gr += o; m += o; v += o;
    if (w16) w16 += blockIdx.y * (size_t)n16;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = gr[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (w16) w16[i] = f2bf(wi);
    }
}

/* rows idx[0..batch) of the device training set -> the batch buffers; member
 * blockIdx.y: its row order at idx + blockIdx.y * sidx, its batch at
 * offset blockIdx.y * batch rows */
__global__ void gather_batch_kernel(const uint16_t *xall, const float *tall, const int *idx, long long sidx, int nin,
                                    int nout, int batch, uint16_t *x16, float *t32)
{
    const int nx = batch * nin, nt = batch * nout;
    idx += blockIdx.y * sidx;
    x16 += (size_t)blockIdx.y * nx;
    t32 += (size_t)blockIdx.y * nt;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nx + nt; i += gridDim.x * blockDim.x) {
        if (i < nx) {
            const int r = i / nin, c = i - r * nin;
            x16[i] = xall[(size_t)idx[r] * nin + c];
        } else {
            const int j = i - nx, r = j / nout, c = j - r * nout;
            t32[j] = tall[(size_t)idx[r] * nout + c];
        }
    }
}

/* acc[i * stride] += x[i] for i < n */
__global__ void add_scalar_kernel(const float *x, float *acc, int n, int stride)
{
    if (blockIdx.x == 0 && (int)threadIdx.x < n) acc[(size_t)threadIdx.x * stride] += x[threadIdx.x];
}

/* ------------------------------------------------------------------ host */

struct ngpu_mlp {
    int device = 0, nin = 0, H = 0, nout = 0, members = 0;
    hipStream_t stream = nullptr;
    std::string err;
    /* per member: parameters in one fp32 buffer; offsets of W1 b1 W2 b2 W3 b3 W4 b4 */
    size_t off[8], nparam = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  /* [members * nparam] */
    size_t nparam16 = 0;                                          /* nparam rounded up to a multiple of 8 */
    uint16_t *w16 = nullptr;                                      /* [members * nparam16] bf16 copy (biases
                                                                     unused): every weight matrix 16-byte
                                                                     aligned, so rows load as vectors */
    long long step = 0;
    /* scratch for batches of up to cap rows per member, members side by side */
    int cap = 0;
    uint16_t *x16 = nullptr, *a16[3] = {nullptr, nullptr, nullptr}, *d16[3] = {nullptr, nullptr, nullptr}, *dy16 = nullptr;
    float *y32 = nullptr, *t32 = nullptr, *dy32 = nullptr, *loss = nullptr;
    /* training set on the device (bf16 inputs, fp32 targets), the epoch's row
     * order and the per-epoch loss sums */
    size_t cap_rows = 0, cap_lacc = 0;
    uint16_t *xall16 = nullptr;
    float *tall = nullptr, *lacc = nullptr;
    int *didx = nullptr;
};

static int mfail(ngpu_mlp *p, int code, const std::string &msg)
{
    if (p) p->err = msg;
    return code;
}
#define MCHK(p, call)                                                                              \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess) return mfail((p), NGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

static int ensure_cap(ngpu_mlp *p, int n)
{
    if (n <= p->cap) return 0;
    (void)hipFree(p->x16); (void)hipFree(p->dy16); (void)hipFree(p->y32); (void)hipFree(p->t32); (void)hipFree(p->dy32);
    for (int i = 0; i < 3; i++) { (void)hipFree(p->a16[i]); (void)hipFree(p->d16[i]); }
    p->x16 = p->dy16 = nullptr; p->y32 = p->t32 = p->dy32 = nullptr;
    for (int i = 0; i < 3; i++) p->a16[i] = p->d16[i] = nullptr;
    p->cap = 0;
    const size_t H = p->H, rows = (size_t)n * p->members;
    MCHK(p, hipMalloc((void **)&p->x16, rows * p->nin * 2));
    for (int i = 0; i < 3; i++) {
        MCHK(p, hipMalloc((void **)&p->a16[i], rows * H * 2));
        MCHK(p, hipMalloc((void **)&p->d16[i], rows * H * 2));
    }
    MCHK(p, hipMalloc((void **)&p->dy16, rows * p->nout * 2));
    MCHK(p, hipMalloc((void **)&p->y32, rows * p->nout * 4));
    MCHK(p, hipMalloc((void **)&p->t32, rows * p->nout * 4));
    MCHK(p, hipMalloc((void **)&p->dy32, rows * p->nout * 4));
    p->cap = n;
    return 0;
}

extern "C" int ngpu_mlp_create(int device, int nin, int hidden, int nout, int members, ngpu_mlp **out)
{
    if (!out || nin < 1 || hidden < 1 || nout < 1 || members < 1 || members > 64 || nin > 4096 || hidden > 4096 ||
        nout > 4096)
        return NGPU_E_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return NGPU_E_HIP;
    ngpu_mlp *p = new ngpu_mlp();
    p->device = device; p->nin = nin; p->H = hidden; p->nout = nout; p->members = members;
    const size_t H = hidden;
    const size_t sz[8] = {H * nin, H, H * H, H, H * H, H, (size_t)nout * H, (size_t)nout};
    size_t o = 0;
    for (int i = 0; i < 8; i++) { p->off[i] = o; o += sz[i]; }
    p->nparam = o;
    p->nparam16 = (o + 7) & ~(size_t)7;
    const size_t tot = o * members;
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&p->w, tot * 4) != hipSuccess || hipMalloc((void **)&p->g, tot * 4) != hipSuccess ||
        hipMalloc((void **)&p->m, tot * 4) != hipSuccess || hipMalloc((void **)&p->v, tot * 4) != hipSuccess ||
        hipMalloc((void **)&p->w16, p->nparam16 * members * 2) != hipSuccess ||
        hipMalloc((void **)&p->loss, members * 4) != hipSuccess) {
        delete p;
        return NGPU_E_NOMEM;
    }
    (void)hipMemset(p->m, 0, tot * 4);
    (void)hipMemset(p->v, 0, tot * 4);
    *out = p;
    return NGPU_OK;
}

extern "C" void ngpu_mlp_close(ngpu_mlp *p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (void *q : {(void *)p->w, (void *)p->g, (void *)p->m, (void *)p->v, (void *)p->w16, (void *)p->loss, (void *)p->x16,
                    (void *)p->dy16, (void *)p->y32, (void *)p->t32, (void *)p->dy32, (void *)p->xall16, (void *)p->tall,
                    (void *)p->lacc, (void *)p->didx})
        if (q) (void)hipFree(q);
    for (int i = 0; i < 3; i++) { (void)hipFree(p->a16[i]); (void)hipFree(p->d16[i]); }
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

extern "C" const char *ngpu_mlp_last_error(const ngpu_mlp *p) { return p ? p->err.c_str() : ""; }

extern "C" long long ngpu_mlp_nparams(const ngpu_mlp *p) { return p ? (long long)p->nparam : 0; }

extern "C" int ngpu_mlp_set_weights(ngpu_mlp *p, const float *w)
{
    if (!p || !w) return NGPU_E_ARG;
    MCHK(p, hipSetDevice(p->device));
    const size_t tot = p->nparam * p->members;
    MCHK(p, hipMemcpy(p->w, w, tot * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(64, p->members), dim3(256), 0, p->stream, p->w, p->w16, p->nparam, p->nparam,
                       p->nparam16);
    MCHK(p, hipMemsetAsync(p->m, 0, tot * 4, p->stream));
    MCHK(p, hipMemsetAsync(p->v, 0, tot * 4, p->stream));
    MCHK(p, hipStreamSynchronize(p->stream));
    p->step = 0;
    return NGPU_OK;
}

extern "C" int ngpu_mlp_get_weights(ngpu_mlp *p, float *w)
{
    if (!p || !w) return NGPU_E_ARG;
    MCHK(p, hipSetDevice(p->device));
    MCHK(p, hipMemcpy(w, p->w, p->nparam * p->members * 4, hipMemcpyDeviceToHost));
    return NGPU_OK;
}

/* forward of members m0 .. m0+nm-1, slot j = member m0+j: inputs x16 (slot
 * j at j * xs elements; xs = 0 shares one input) -> activations a16[0..2]
 * and outputs y32, slot j at j * n rows */
static int forward(ngpu_mlp *p, int m0, int nm, int n, long long xs)
{
    const float *W = p->w + (size_t)m0 * p->nparam;
    const uint16_t *W16 = p->w16 + (size_t)m0 * p->nparam16;
    const long long P = (long long)p->nparam, P16 = (long long)p->nparam16, H = p->H;
    GemmArgs g{};
    const uint16_t *in = p->x16;
    int K = p->nin;
    long long sin = xs;
    for (int l = 0; l < 3; l++) {
        g = GemmArgs{n, (int)H, K, K, K, (int)H, in, W16 + p->off[2 * l], W + p->off[2 * l + 1], nullptr, nullptr, p->a16[l], 1,
                     sin, P16, P, 0, n * H};
        MCHK(p, gemm(p->stream, false, true, g, nm));
        in = p->a16[l];
        K = (int)H;
        sin = n * H;
    }
    g = GemmArgs{n, p->nout, (int)H, (int)H, (int)H, p->nout, in, W16 + p->off[6], W + p->off[7], nullptr, p->y32, nullptr, 0,
                 n * H, P16, P, 0, (long long)n * p->nout};
    MCHK(p, gemm(p->stream, false, true, g, nm));
    return 0;
}

/* gradients of members m0 .. m0+nm-1 (loss = mean squared error over n x
 * nout against t32, slot j) into p->g; loss of slot j in p->loss[j] */
static int backward(ngpu_mlp *p, int m0, int nm, int n)
{
    const uint16_t *W16 = p->w16 + (size_t)m0 * p->nparam16;
    float *G = p->g + (size_t)m0 * p->nparam;
    const long long P16 = (long long)p->nparam16;
    const int H = p->H, O = p->nout;
    const long long P = (long long)p->nparam, nH = (long long)n * H, nO = (long long)n * O;
    const int tot = n * O;
    MCHK(p, hipMemsetAsync(p->loss, 0, (size_t)nm * 4, p->stream));
    hipLaunchKernelGGL(mse_grad_kernel, dim3(std::min(256, (tot + 255) / 256), nm), dim3(256), 0, p->stream, p->y32, p->t32,
                       tot, 1.0f / (float)tot, p->dy32, p->dy16, p->loss);
    /* layer 4: dW4 = dY^T A3, db4 = colsum(dY), dZ3 = (dY W4) * (A3 > 0) */
    GemmArgs g{O, H, n, O, H, H, p->dy16, p->a16[2], nullptr, nullptr, G + p->off[6], nullptr, 0, nO, nH, 0, 0, P};
    MCHK(p, gemm(p->stream, true, false, g, nm));
    hipLaunchKernelGGL(colsum_kernel, dim3((O + 63) / 64, nm), dim3(64 * COLSUM_RG), 0, p->stream, p->dy32,
                       (const uint16_t *)nullptr, n, O, G + p->off[7], P);
    g = GemmArgs{n, H, O, O, H, H, p->dy16, W16 + p->off[6], nullptr, p->a16[2], nullptr, p->d16[2], 0, nO, P16, 0, nH, nH};
    MCHK(p, gemm(p->stream, false, false, g, nm));
    for (int l = 2; l >= 0; l--) {
        const uint16_t *prev = l > 0 ? p->a16[l - 1] : p->x16;
        const int K = l > 0 ? H : p->nin;
        g = GemmArgs{H, K, n, H, K, K, p->d16[l], prev, nullptr, nullptr, G + p->off[2 * l], nullptr, 0, nH, (long long)n * K, 0, 0, P};
        MCHK(p, gemm(p->stream, true, false, g, nm));
        hipLaunchKernelGGL(colsum_kernel, dim3((H + 63) / 64, nm), dim3(64 * COLSUM_RG), 0, p->stream, (const float *)nullptr,
                           p->d16[l], n, H, G + p->off[2 * l + 1], P);
        if (l > 0) {
            g = GemmArgs{n, H, H, H, H, H, p->d16[l], W16 + p->off[2 * l], nullptr, p->a16[l - 1], nullptr, p->d16[l - 1], 0,
                         nH, P16, 0, nH, nH};
            MCHK(p, gemm(p->stream, false, false, g, nm));
        }
    }
    return 0;
}

static int upload_x(ngpu_mlp *p, const float *x, int n)
{
    std::vector<uint16_t> h((size_t)n * p->nin);
    for (size_t i = 0; i < h.size(); i++) h[i] = f2bf(x[i]);
    MCHK(p, hipMemcpyAsync(p->x16, h.data(), h.size() * 2, hipMemcpyHostToDevice, p->stream));
    MCHK(p, hipStreamSynchronize(p->stream));
    return 0;
}

extern "C" int ngpu_mlp_predict(ngpu_mlp *p, const float *x, int n, float *out)
{
    if (!p || !x || !out || n < 1) return NGPU_E_ARG;
    MCHK(p, hipSetDevice(p->device));
    int rc;
    if ((rc = ensure_cap(p, n)) || (rc = upload_x(p, x, n))) return rc;
    /* every member on the one input; y32 holds member j's outputs at j * n rows */
    if ((rc = forward(p, 0, p->members, n, 0))) return rc;
    MCHK(p, hipMemcpyAsync(out, p->y32, (size_t)p->members * n * p->nout * 4, hipMemcpyDeviceToHost, p->stream));
    MCHK(p, hipStreamSynchronize(p->stream));
    return NGPU_OK;
}

extern "C" int ngpu_mlp_gradients(ngpu_mlp *p, int mem, const float *x, const float *t, int n, float *grad, float *loss)
{
    if (!p || !x || !t || !grad || n < 1 || mem < 0 || mem >= p->members) return NGPU_E_ARG;
    MCHK(p, hipSetDevice(p->device));
    int rc;
    if ((rc = ensure_cap(p, n)) || (rc = upload_x(p, x, n))) return rc;
    MCHK(p, hipMemcpyAsync(p->t32, t, (size_t)n * p->nout * 4, hipMemcpyHostToDevice, p->stream));
    if ((rc = forward(p, mem, 1, n, 0)) || (rc = backward(p, mem, 1, n))) return rc;
    MCHK(p, hipMemcpyAsync(grad, p->g + mem * p->nparam, p->nparam * 4, hipMemcpyDeviceToHost, p->stream));
    if (loss) MCHK(p, hipMemcpyAsync(loss, p->loss, 4, hipMemcpyDeviceToHost, p->stream));
    MCHK(p, hipStreamSynchronize(p->stream));
    return NGPU_OK;
}

/* Mini-batch Adam on (x [n, nin], t [n, nout]) for every member; losses_out
 * [members * epochs] receives the mean batch loss per epoch (may be null). */
extern "C" int ngpu_mlp_train(ngpu_mlp *p, const float *x, const float *t, int n, int epochs, int batch, float lr,
                              unsigned seed, float *losses_out)
{
    if (!p || !x || !t || n < 1 || epochs < 0 || batch < 1) return NGPU_E_ARG;
    MCHK(p, hipSetDevice(p->device));
    batch = std::min(batch, n);
    int rc;
    if ((rc = ensure_cap(p, batch))) return rc;
    /* the training set stays on the device; each epoch uploads its row order
     * and each step gathers its rows there, so the steps of an epoch run
     * without a host round trip (the loss sums are read once at the end) */
    if ((size_t)n > p->cap_rows) {
        (void)hipFree(p->xall16); (void)hipFree(p->tall); (void)hipFree(p->didx);
        p->xall16 = nullptr; p->tall = nullptr; p->didx = nullptr; p->cap_rows = 0;
        MCHK(p, hipMalloc((void **)&p->xall16, (size_t)n * p->nin * 2));
        MCHK(p, hipMalloc((void **)&p->tall, (size_t)n * p->nout * 4));
        MCHK(p, hipMalloc((void **)&p->didx, (size_t)n * p->members * 4));
        p->cap_rows = n;
    }
    const size_t nl = (size_t)p->members * std::max(epochs, 1);
    if (nl > p->cap_lacc) {
        (void)hipFree(p->lacc);
        p->lacc = nullptr; p->cap_lacc = 0;
        MCHK(p, hipMalloc((void **)&p->lacc, nl * 4));
        p->cap_lacc = nl;
    }
    std::vector<uint16_t> xb((size_t)n * p->nin);
    for (size_t i = 0; i < xb.size(); i++) xb[i] = f2bf(x[i]);
    MCHK(p, hipMemcpyAsync(p->xall16, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, p->stream));
    MCHK(p, hipMemcpyAsync(p->tall, t, (size_t)n * p->nout * 4, hipMemcpyHostToDevice, p->stream));
    MCHK(p, hipMemsetAsync(p->lacc, 0, nl * 4, p->stream));
    /* each member draws its own row order per epoch (its own generator);
     * the members then take their steps side by side */
    const int M = p->members;
    std::vector<std::mt19937> rng;
    for (int mem = 0; mem < M; mem++) rng.emplace_back(seed * 1009u + (unsigned)mem);
    std::vector<int> idx((size_t)n * M);
    for (int mem = 0; mem < M; mem++)
        for (int i = 0; i < n; i++) idx[(size_t)mem * n + i] = i;
    const int nb = n / batch;
    const int gb = std::min(1024, (batch * (p->nin + p->nout) + 255) / 256);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long long step = 0;
    for (int ep = 0; ep < epochs; ep++) {
        for (int mem = 0; mem < M; mem++)
            std::shuffle(idx.begin() + (size_t)mem * n, idx.begin() + (size_t)(mem + 1) * n, rng[mem]);
        MCHK(p, hipMemcpyAsync(p->didx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, p->stream));
        for (int s0 = 0; s0 + batch <= n; s0 += batch) {
            hipLaunchKernelGGL(gather_batch_kernel, dim3(gb, M), dim3(256), 0, p->stream, p->xall16, p->tall, p->didx + s0,
                               (long long)n, p->nin, p->nout, batch, p->x16, p->t32);
            if ((rc = forward(p, 0, M, batch, (long long)batch * p->nin)) || (rc = backward(p, 0, M, batch))) return rc;
            step++;
            const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
            hipLaunchKernelGGL(adam_kernel, dim3(256, M), dim3(256), 0, p->stream, p->w, p->g, p->m, p->v, p->w16,
                               (int)p->nparam, (int)p->nparam16, lr, b1, b2, eps, c1, c2);
            hipLaunchKernelGGL(add_scalar_kernel, dim3(1), dim3(64), 0, p->stream, p->loss, p->lacc + ep, M, epochs);
        }
        /* the next epoch rewrites the host row order */
        MCHK(p, hipStreamSynchronize(p->stream));
    }
    MCHK(p, hipGetLastError());
    MCHK(p, hipStreamSynchronize(p->stream));    /* xb and t are read by the uploads */
    if (losses_out && epochs > 0) {
        std::vector<float> l(nl);
        MCHK(p, hipMemcpy(l.data(), p->lacc, nl * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nl; i++) losses_out[i] = nb ? l[i] / (float)nb : 0.0f;
    }
    return NGPU_OK;
}

/* Plain GEMM on host buffers, for tests: C[M,N] = op(A) op(B), A/B bf16 bit patterns. */
extern "C" int ngpu_gemm_bf16(int device, int ta, int tb, int M, int N, int K, const unsigned short *A,
                              const unsigned short *B, float *C)
{
    if (M < 1 || N < 1 || K < 1 || !A || !B || !C) return NGPU_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return NGPU_E_HIP;
    uint16_t *dA = nullptr, *dB = nullptr;
    float *dC = nullptr;
    int rc = NGPU_OK;
    if (hipMalloc((void **)&dA, (size_t)M * K * 2) != hipSuccess || hipMalloc((void **)&dB, (size_t)K * N * 2) != hipSuccess ||
        hipMalloc((void **)&dC, (size_t)M * N * 4) != hipSuccess) {
        rc = NGPU_E_NOMEM;
    } else {
        (void)hipMemcpy(dA, A, (size_t)M * K * 2, hipMemcpyHostToDevice);
        (void)hipMemcpy(dB, B, (size_t)K * N * 2, hipMemcpyHostToDevice);
        GemmArgs g{M, N, K, ta ? M : K, tb ? K : N, N, dA, dB, nullptr, nullptr, dC, nullptr, 0};
        if (gemm(nullptr, ta != 0, tb != 0, g, 1) != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = NGPU_E_HIP;
        else if (hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = NGPU_E_HIP;
    }
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC);
    return rc;
}
