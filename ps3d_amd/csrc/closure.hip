// closure.hip -- learned subgrid closure for the coarse solver: a 4-layer 3x3x3
// convolution net (6 -> 32 -> 32 -> 32 -> 3 channels, GELU between layers) on the
// solver grid, its backward pass, Adam, and the hooks that add its output to the
// vorticity tendency of every RK stage.  No counterpart in ps3d.
//
// Tensors are channels-last over the nx x ny x (nz+1) points, point index
// p = (ix*ny + iy)*(nz+1) + iz (x, y periodic; zero beyond the lids in z).
// Activations and weights feeding the matrix cores are bf16; accumulation, biases,
// pre-activations, gradients and the Adam state are fp32.
//
// Convolution = implicit GEMM on the MFMA units: a block owns a brick of 4 x 4
// columns x 16 levels (input halo staged in LDS), a wavefront 4 columns x all
// output channels; per tap and 16/32-channel slice it issues
// v_mfma_f32_16x16x16_bf16 (16 input channels) or v_mfma_f32_16x16x32_bf16 (32).
// Operand lanes: A row = lane%16 (point), k = (K/4)(lane/16)+v (channel);
// B column = lane%16 (output channel), same k; D row = 4(lane/16)+v, column = lane%16.
// The input gradient is the same kernel with the weights transposed and the taps
// mirrored; the weight gradient is a matrix-core reduction over point chunks
// (bf16 gradient x bf16 input, fp32 accumulation; an fp32 path is kept).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ps3d_closure.h"
#include "kernels.hpp"

namespace ps3d {
namespace {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

constexpr int NL = 4;
constexpr int CIN_[NL] = {6, 32, 32, 32};
constexpr int COUT_[NL] = {32, 32, 32, 3};
constexpr int CINP[NL] = {16, 32, 32, 32};    // padded to the MFMA K slices
constexpr int COUTP[NL] = {32, 32, 32, 16};   // padded to 16-column tiles
constexpr int WG_CHUNK = 1024;                // points per weight-gradient block
constexpr int WG_SUB = 128;                   // points staged in LDS at a time

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7fffu + ((u >> 16) & 1u);          // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.39894228040143268f;
}

enum { CM_HID = 0, CM_OUT = 1, CM_DG = 2, CM_DX = 3 };

struct ConvArgs {
    const uint16_t *in;     // [npts][CIN] bf16
    const uint16_t *w;      // [27][COUT][CIN] bf16
    const float *bias;      // [COUT] (CM_HID, CM_OUT)
    const float *pre;       // [npts][COUT] pre-activation of the previous layer (CM_DG)
    float *outf;            // [npts][COUT]: pre-activation / output / gradient
    uint16_t *outb;         // [npts][COUT] bf16: activation (CM_HID) / gradient (CM_DG)
    int nx, ny, nz1;
    long npts;
};

// Brick of 4 x 4 (x, y) columns x 16 z levels per block (256 points, one 16-row
// M tile per column); its input halo (6 x 6 x 18 points, all channels) is staged in
// LDS once and read by all 27 taps.
constexpr int BX = 4, BY = 4, BZ = 16, HX = BX + 2, HY = BY + 2, HZ = BZ + 2;

__host__ __device__ inline int conv_blocks(int nx, int ny, int nz1) {
    return ((nx + BX - 1) / BX) * ((ny + BY - 1) / BY) * ((nz1 + BZ - 1) / BZ);
}

template <int CIN, int COUT, int MODE>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
    constexpr int KK = (CIN % 32 == 0) ? 32 : 16;
    constexpr int NT = COUT / 16;
    constexpr int MT = 4;                      // columns (M tiles) per wave
    constexpr int CH = CIN * 2 / 16;           // 16-byte chunks per halo row
    __shared__ uint4 halo4[HX * HY * HZ * CH];
    const uint16_t *hal = reinterpret_cast<const uint16_t *>(halo4);
    const int nbx = (a.nx + BX - 1) / BX, nby = (a.ny + BY - 1) / BY;
    int b = blockIdx.x;
    const int bz = b / (nbx * nby);
    b -= bz * nbx * nby;
    const int x0 = (b / nby) * BX, y0 = (b % nby) * BY, z0 = bz * BZ;
    for (int e = threadIdx.x; e < HX * HY * HZ * CH; e += 256) {
        const int ch = e % CH, r = e / CH;
        const int hz = r % HZ, hy = (r / HZ) % HY, hx = r / (HZ * HY);
        const int z = z0 + hz - 1;
        int x = (x0 + hx - 1) % a.nx, y = (y0 + hy - 1) % a.ny;
        x = x < 0 ? x + a.nx : x;
        y = y < 0 ? y + a.ny : y;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (z >= 0 && z < a.nz1)
            v = *reinterpret_cast<const uint4 *>(a.in + (((size_t)x * a.ny + y) * a.nz1 + z) * CIN + ch * 8);
        halo4[e] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    f4v acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 3
    for (int t = 0; t < 27; ++t) {
        const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
        const int c0 = kq * (KK / 4);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int c = wave * MT + mt;                 // brick column
            const int hr = (((c / BY) + 1 + dx) * HY + ((c % BY) + 1 + dy)) * HZ + (row + 1 + dz);
            if constexpr (KK == 16) {
                const s4v af = *reinterpret_cast<const s4v *>(hal + (size_t)hr * CIN + c0);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const s4v bf =
                        *reinterpret_cast<const s4v *>(a.w + ((size_t)(t * COUT + nt * 16 + row) * CIN + c0));
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf, acc[mt][nt], 0, 0, 0);
                }
            } else {
                const b8v af = *reinterpret_cast<const b8v *>(hal + (size_t)hr * CIN + c0);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const b8v bf =
                        *reinterpret_cast<const b8v *>(a.w + ((size_t)(t * COUT + nt * 16 + row) * CIN + c0));
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf, acc[mt][nt], 0, 0, 0);
                }
            }
        }
    }
    // D: row 4*(lane/16)+v (z in the brick), column lane%16 (channel)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int c = wave * MT + mt;
        const int x = x0 + c / BY, y = y0 + c % BY;
        if (x >= a.nx || y >= a.ny) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int z = z0 + 4 * kq + v;
                if (z >= a.nz1) continue;
                const long p = ((long)x * a.ny + y) * a.nz1 + z;
                const int co = nt * 16 + row;
                const size_t o = (size_t)p * COUT + co;
                float val = acc[mt][nt][v];
                if (MODE == CM_HID) {
                    val += a.bias[co];
                    a.outf[o] = val;
                    a.outb[o] = f2bf(gelu(val));
                } else if (MODE == CM_OUT) {
                    a.outf[o] = val + a.bias[co];
                } else if (MODE == CM_DG) {
                    val *= gelu_grad(a.pre[o]);
                    a.outf[o] = val;
                    a.outb[o] = f2bf(val);
                } else {
                    a.outf[o] = val;
                }
            }
    }
}

// Weight (and bias) gradient partials of one layer over a chunk of points and one
// tap: part[(chunk*27 + t)][co][ci] = sum_p g[p][co] in[p + d_t][ci].
struct WgArgs {
    const float *g;          // [npts][COUT] fp32
    const uint16_t *in;      // [npts][CIN] bf16
    float *part;             // [nchunk][27][COUT*CIN]
    float *partb;            // [nchunk][COUT]
    int nx, ny, nz1;
    long npts;
};

template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad(WgArgs a) {
    constexpr int NPAIR = COUT * CIN;
    constexpr int PPT = (NPAIR + 255) / 256;
    __shared__ float gs[WG_SUB][COUT];
    __shared__ float as[WG_SUB][CIN];
    const int t = blockIdx.y;
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    const int tid = threadIdx.x;
    float acc[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) acc[q] = 0.0f;
    float bacc = 0.0f;
    const long base0 = (long)blockIdx.x * WG_CHUNK;
    for (int sub = 0; sub < WG_CHUNK; sub += WG_SUB) {
        const long base = base0 + sub;
        for (int e = tid; e < WG_SUB * COUT; e += 256) {
            const int i = e / COUT, c = e % COUT;
            const long p = base + i;
            gs[i][c] = p < a.npts ? a.g[(size_t)p * COUT + c] : 0.0f;
        }
        for (int e = tid; e < WG_SUB * CIN; e += 256) {
            const int i = e / CIN, c = e % CIN;
            const long p = base + i;
            float v = 0.0f;
            if (p < a.npts) {
                const int z = (int)(p % a.nz1) + dz;
                const long r = p / a.nz1;
                int y = (int)(r % a.ny) + dy, x = (int)(r / a.ny) + dx;
                x = x < 0 ? x + a.nx : (x >= a.nx ? x - a.nx : x);
                y = y < 0 ? y + a.ny : (y >= a.ny ? y - a.ny : y);
                if (z >= 0 && z < a.nz1) v = bf2f(a.in[(((size_t)x * a.ny + y) * a.nz1 + z) * CIN + c]);
            }
            as[i][c] = v;
        }
        __syncthreads();
#pragma unroll 4
        for (int i = 0; i < WG_SUB; ++i) {
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int pr = tid + q * 256;
                if (pr < NPAIR) acc[q] += gs[i][pr / CIN] * as[i][pr % CIN];
            }
            if (t == 13 && tid < COUT) bacc += gs[i][tid];
        }
        __syncthreads();
    }
    float *out = a.part + ((size_t)blockIdx.x * 27 + t) * NPAIR;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int pr = tid + q * 256;
        if (pr < NPAIR) out[pr] = acc[q];
    }
    if (t == 13 && tid < COUT) a.partb[(size_t)blockIdx.x * COUT + tid] = bacc;
}

// The same partials on the matrix cores: per block one chunk of WG_CHUNK points
// and one tap, bf16 operands (the layer's bf16 output gradient and input), fp32
// accumulation.  dW[co][ci] = sum_p g[p][co] in[p + d_t][ci] is a 16 x 16 tile
// product per wave with K = points: A (16 co x 32 p) and B (32 p x 16 ci) are read
// from LDS copies transposed to [channel][point], so that each lane's 8 consecutive
// points of one channel are one 16-byte read.  The bias gradient (tap 13) stays an
// fp32 sum of the fp32 gradient.
constexpr int WM_SUB = 256;                   // points staged per sub-chunk
constexpr int WM_LD = WM_SUB + 8;             // LDS row (bf16), 16-byte aligned rows

struct WgmArgs {
    const uint16_t *g;       // [npts][COUT] bf16
    const float *gf;         // [npts][COUT] fp32 (bias)
    const uint16_t *in;      // [npts][CIN] bf16
    float *part;             // [nchunk][27][COUT*CIN]
    float *partb;            // [nchunk][COUT]
    int nx, ny, nz1;
    long npts;
};

template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad_mfma(WgmArgs a) {
    constexpr int TCO = COUT / 16, TCI = CIN / 16, NT = TCO * TCI;   // 16 x 16 tiles (<= 4)
    __shared__ __attribute__((aligned(16))) uint16_t gT[COUT][WM_LD];
    __shared__ __attribute__((aligned(16))) uint16_t iT[CIN][WM_LD];
    const int t = blockIdx.y;
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tco = (wave % NT) / TCI, tci = (wave % NT) % TCI;
    f4v acc = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    float bacc = 0.0f;
    const long base0 = (long)blockIdx.x * WG_CHUNK;
    for (int sub = 0; sub < WG_CHUNK; sub += WM_SUB) {
        const long base = base0 + sub;
        for (int e = tid; e < WM_SUB * COUT; e += 256) {
            const int i = e / COUT, c = e % COUT;
            const long p = base + i;
            gT[c][i] = p < a.npts ? a.g[(size_t)p * COUT + c] : (uint16_t)0;
        }
        for (int e = tid; e < WM_SUB * CIN; e += 256) {
            const int i = e / CIN, c = e % CIN;
            const long p = base + i;
            uint16_t v = 0;
            if (p < a.npts) {
                const int z = (int)(p % a.nz1) + dz;
                const long r = p / a.nz1;
                int y = (int)(r % a.ny) + dy, x = (int)(r / a.ny) + dx;
                x = x < 0 ? x + a.nx : (x >= a.nx ? x - a.nx : x);
                y = y < 0 ? y + a.ny : (y >= a.ny ? y - a.ny : y);
                if (z >= 0 && z < a.nz1) v = a.in[(((size_t)x * a.ny + y) * a.nz1 + z) * CIN + c];
            }
            iT[c][i] = v;
        }
        if (t == 13 && tid < COUT)
            for (int i = 0; i < WM_SUB; ++i) {
                const long p = base + i;
                if (p < a.npts) bacc += a.gf[(size_t)p * COUT + tid];
            }
        __syncthreads();
        if (wave < NT) {
            const int co = tco * 16 + (lane & 15), ci = tci * 16 + (lane & 15), k8 = 8 * (lane >> 4);
#pragma unroll 4
            for (int pk = 0; pk < WM_SUB; pk += 32) {
                const b8v af = *reinterpret_cast<const b8v *>(&gT[co][pk + k8]);
                const b8v bf = *reinterpret_cast<const b8v *>(&iT[ci][pk + k8]);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D row 4*(lane/16)+v = output channel, column lane%16 = input channel
    if (wave < NT) {
        float *out = a.part + ((size_t)blockIdx.x * 27 + t) * (COUT * CIN);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int co = tco * 16 + 4 * (lane >> 4) + v, ci = tci * 16 + (lane & 15);
            out[co * CIN + ci] = acc[v];
        }
    }
    if (t == 13 && tid < COUT) a.partb[(size_t)blockIdx.x * COUT + tid] = bacc;
}

// grad[w] = sum over chunks (fixed order: deterministic)
__global__ void __launch_bounds__(256) k_wgrad_reduce(const float *part, const float *partb, int nchunk, int nw,
                                                      int cout, float *gw, float *gb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nw) {
        float s = 0.0f;
        for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * nw + i];
        gw[i] = s;
    } else if (i < nw + cout) {
        const int co = i - nw;
        float s = 0.0f;
        for (int c = 0; c < nchunk; ++c) s += partb[(size_t)c * cout + co];
        gb[co] = s;
    }
}

// Mean-square loss over the 3 output channels and its gradient (fp32 + bf16).
__global__ void __launch_bounds__(256) k_mse(const float *y, const float *target, long npts, float *gy, uint16_t *gyb,
                                             float *lpart) {
    __shared__ float red[256];
    const float sc = 2.0f / (3.0f * (float)npts);   // d loss / d y
    float s = 0.0f;
    for (long p = blockIdx.x * 256L + threadIdx.x; p < npts; p += (long)gridDim.x * 256) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            float g = 0.0f;
            if (c < 3) {
                const float d = y[p * 16 + c] - target[p * 3 + c];
                s += d * d;
                g = sc * d;
            }
            gy[p * 16 + c] = g;
            gyb[p * 16 + c] = f2bf(g);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

__global__ void k_loss_reduce(const float *lpart, int n, float inv, float *loss, float *loss_acc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.0f;
    for (int i = 0; i < n; ++i) s += lpart[i];
    s *= inv;
    loss[0] = s;
    loss_acc[0] += s;
}

__global__ void __launch_bounds__(256) k_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of one layer: wf[t][co][ci] and the mirrored transpose wt[26-t][ci][co]
__global__ void __launch_bounds__(256) k_refresh(const float *w, int coutp, int cinp, uint16_t *wf, uint16_t *wt) {
    const int n = 27 * coutp * cinp;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int t = i / (coutp * cinp), r = i % (coutp * cinp), co = r / cinp, ci = r % cinp;
    const uint16_t b = f2bf(w[i]);
    wf[i] = b;
    wt[((size_t)(26 - t) * cinp + ci) * coutp + co] = b;
}

// Closure output (fp32 [p][16], channels 0..2) scaled into three solver-layout volumes.
__global__ void __launch_bounds__(256) k_unpack(const float *y, long npts, int ny, int nz1, int zs, float s0, float s1,
                                                float s2, double *o0, double *o1, double *o2) {
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= npts) return;
    const int iz = (int)(p % nz1);
    const long col = p / nz1;
    const size_t i = (size_t)col * zs + iz;
    o0[i] = (double)(y[p * 16 + 0] * s0);
    o1[i] = (double)(y[p * 16 + 1] * s1);
    o2[i] = (double)(y[p * 16 + 2] * s2);
    (void)ny;
}

thread_local std::string c_err;
int cfail(int code, const std::string &m) {
    c_err = m;
    return code;
}
#define CHK(expr)                                                                                             \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess) return cfail(PS3D_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

}  // namespace
}  // namespace ps3d

using namespace ps3d;

struct ps3d_cnn {
    int nx = 0, ny = 0, nz1 = 0;
    long npts = 0;
    hipStream_t own = nullptr;
    size_t woff[NL]{}, boff[NL]{}, nparam = 0;     // padded master layout
    float *w = nullptr, *gw = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wf[NL]{}, *wt[NL]{};
    uint16_t *a0 = nullptr;                         // input [npts][16]
    float *pre[3]{};
    uint16_t *act[3]{};
    float *y = nullptr, *gy = nullptr, *gpA = nullptr, *gpB = nullptr, *gx = nullptr;
    uint16_t *gyb = nullptr, *gbA = nullptr, *gbB = nullptr;
    float *part = nullptr, *partb = nullptr, *lpart = nullptr, *loss = nullptr, *loss_acc = nullptr;
    int nchunk = 0, nlblk = 0;
    long adam_k = 0;
    bool wg_mfma = true;                            // weight gradient on the matrix cores
    // device dataset
    uint16_t *dx = nullptr;
    float *dt = nullptr;
    int ns = 0;
    float *tgt = nullptr;                           // host-target staging [npts][3]
};

namespace {

unsigned blocks_for(long n, int per) { return (unsigned)((n + per - 1) / per); }

int free_cnn(ps3d_cnn *c) {
    if (!c) return 0;
    if (c->own) (void)hipStreamSynchronize(c->own);
    void *ptrs[] = {c->w, c->gw, c->m, c->v, c->a0, c->y, c->gy, c->gpA, c->gpB, c->gx, c->gyb, c->gbA, c->gbB,
                    c->part, c->partb, c->lpart, c->loss, c->loss_acc, c->dx, c->dt, c->tgt};
    for (void *p : ptrs) (void)hipFree(p);
    for (int l = 0; l < NL; ++l) {
        (void)hipFree(c->wf[l]);
        (void)hipFree(c->wt[l]);
    }
    for (int l = 0; l < 3; ++l) {
        (void)hipFree(c->pre[l]);
        (void)hipFree(c->act[l]);
    }
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return 0;
}

template <int CIN, int COUT, int MODE>
void conv(const ps3d_cnn *c, const uint16_t *in, const uint16_t *w, const float *bias, const float *pre, float *outf,
          uint16_t *outb, hipStream_t s) {
    ConvArgs a;
    a.in = in;
    a.w = w;
    a.bias = bias;
    a.pre = pre;
    a.outf = outf;
    a.outb = outb;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nz1 = c->nz1;
    a.npts = c->npts;
    hipLaunchKernelGGL((k_conv<CIN, COUT, MODE>), dim3(conv_blocks(c->nx, c->ny, c->nz1)), dim3(256), 0, s, a);
}

// weight and bias gradient of one layer: g the fp32 output gradient, gb its bf16
// copy (the matrix-core path; nullptr: the fp32 path)
template <int CIN, int COUT>
void wgrad(const ps3d_cnn *c, const float *g, const uint16_t *gb, const uint16_t *in, int layer, hipStream_t s) {
    if (gb && c->wg_mfma) {
        WgmArgs m;
        m.g = gb;
        m.gf = g;
        m.in = in;
        m.part = c->part;
        m.partb = c->partb;
        m.nx = c->nx;
        m.ny = c->ny;
        m.nz1 = c->nz1;
        m.npts = c->npts;
        hipLaunchKernelGGL((k_wgrad_mfma<CIN, COUT>), dim3(c->nchunk, 27), dim3(256), 0, s, m);
    } else {
        WgArgs a;
        a.g = g;
        a.in = in;
        a.part = c->part;
        a.partb = c->partb;
        a.nx = c->nx;
        a.ny = c->ny;
        a.nz1 = c->nz1;
        a.npts = c->npts;
        hipLaunchKernelGGL((k_wgrad<CIN, COUT>), dim3(c->nchunk, 27), dim3(256), 0, s, a);
    }
    const int nw = 27 * COUT * CIN;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(nw + COUT, 256)), dim3(256), 0, s, (const float *)c->part,
                       (const float *)c->partb, c->nchunk, nw, COUT, c->gw + c->woff[layer], c->gw + c->boff[layer]);
}

void forward(ps3d_cnn *c, const uint16_t *x, hipStream_t s) {
    conv<16, 32, CM_HID>(c, x, c->wf[0], c->w + c->boff[0], nullptr, c->pre[0], c->act[0], s);
    conv<32, 32, CM_HID>(c, c->act[0], c->wf[1], c->w + c->boff[1], nullptr, c->pre[1], c->act[1], s);
    conv<32, 32, CM_HID>(c, c->act[1], c->wf[2], c->w + c->boff[2], nullptr, c->pre[2], c->act[2], s);
    conv<32, 16, CM_OUT>(c, c->act[2], c->wf[3], c->w + c->boff[3], nullptr, c->y, nullptr, s);
}

// loss + gradients into gw (x: the input of the forward pass just run)
void backward(ps3d_cnn *c, const uint16_t *x, const float *target, bool want_gx, hipStream_t s) {
    hipLaunchKernelGGL(k_mse, dim3(c->nlblk), dim3(256), 0, s, (const float *)c->y, target, c->npts, c->gy, c->gyb,
                       c->lpart);
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(1), 0, s, (const float *)c->lpart, c->nlblk,
                       1.0f / (3.0f * (float)c->npts), c->loss, c->loss_acc);
    wgrad<32, 16>(c, c->gy, c->gyb, c->act[2], 3, s);
    conv<16, 32, CM_DG>(c, c->gyb, c->wt[3], nullptr, c->pre[2], c->gpA, c->gbA, s);
    wgrad<32, 32>(c, c->gpA, c->gbA, c->act[1], 2, s);
    conv<32, 32, CM_DG>(c, c->gbA, c->wt[2], nullptr, c->pre[1], c->gpB, c->gbB, s);
    wgrad<32, 32>(c, c->gpB, c->gbB, c->act[0], 1, s);
    conv<32, 32, CM_DG>(c, c->gbB, c->wt[1], nullptr, c->pre[0], c->gpA, c->gbA, s);
    wgrad<16, 32>(c, c->gpA, c->gbA, x, 0, s);
    if (want_gx) conv<32, 16, CM_DX>(c, c->gbA, c->wt[0], nullptr, nullptr, c->gx, nullptr, s);
}

void refresh(ps3d_cnn *c, hipStream_t s) {
    for (int l = 0; l < NL; ++l)
        hipLaunchKernelGGL(k_refresh, dim3(blocks_for(27L * COUTP[l] * CINP[l], 256)), dim3(256), 0, s,
                           (const float *)(c->w + c->woff[l]), COUTP[l], CINP[l], c->wf[l], c->wt[l]);
}

void adam(ps3d_cnn *c, float lr, hipStream_t s) {
    c->adam_k += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)c->adam_k), c2 = 1.0f - std::pow(b2, (float)c->adam_k);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for((long)c->nparam, 256)), dim3(256), 0, s, c->w, (const float *)c->gw,
                       c->m, c->v, (long)c->nparam, lr, b1, b2, eps, c1, c2);
    refresh(c, s);
}

// packed (reference order, unpadded) <-> padded master layout
void pack_to_padded(const ps3d_cnn *c, const float *packed, std::vector<float> &pad) {
    pad.assign(c->nparam, 0.0f);
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        for (int t = 0; t < 27; ++t)
            for (int co = 0; co < COUT_[l]; ++co)
                for (int ci = 0; ci < CIN_[l]; ++ci)
                    pad[c->woff[l] + ((size_t)t * COUTP[l] + co) * CINP[l] + ci] = packed[o++];
        for (int co = 0; co < COUT_[l]; ++co) pad[c->boff[l] + co] = packed[o++];
    }
}
void padded_to_pack(const ps3d_cnn *c, const std::vector<float> &pad, float *packed) {
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        for (int t = 0; t < 27; ++t)
            for (int co = 0; co < COUT_[l]; ++co)
                for (int ci = 0; ci < CIN_[l]; ++ci)
                    packed[o++] = pad[c->woff[l] + ((size_t)t * COUTP[l] + co) * CINP[l] + ci];
        for (int co = 0; co < COUT_[l]; ++co) packed[o++] = pad[c->boff[l] + co];
    }
}

size_t packed_count() {
    size_t n = 0;
    for (int l = 0; l < NL; ++l) n += 27 * (size_t)COUT_[l] * CIN_[l] + COUT_[l];
    return n;
}

// host [npts][6] fp32 -> device [npts][16] bf16
int upload_input(const ps3d_cnn *c, const float *x, uint16_t *dst, hipStream_t s) {
    std::vector<uint16_t> hb((size_t)c->npts * 16, 0);
    for (long p = 0; p < c->npts; ++p)
        for (int k = 0; k < 6; ++k) hb[p * 16 + k] = f2bf(x[p * 6 + k]);
    CHK(hipMemcpyAsync(dst, hb.data(), hb.size() * 2, hipMemcpyHostToDevice, s));
    CHK(hipStreamSynchronize(s));
    return 0;
}

}  // namespace

extern "C" {

const char *ps3d_cnn_last_error(void) { return ps3d::c_err.c_str(); }

int ps3d_cnn_nparams(int *n) {
    if (!n) return cfail(PS3D_E_ARG, "null argument");
    *n = (int)packed_count();
    return 0;
}

int ps3d_cnn_create(int nx, int ny, int nz, ps3d_cnn **out) {
    if (!out || nx < 3 || ny < 3 || nz < 2) return cfail(PS3D_E_ARG, "invalid grid");
    *out = nullptr;
    ps3d_cnn *c = new ps3d_cnn();
    c->nx = nx;
    c->ny = ny;
    c->nz1 = nz + 1;
    c->npts = (long)nx * ny * (nz + 1);
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        c->woff[l] = o;
        o += 27 * (size_t)COUTP[l] * CINP[l];
        c->boff[l] = o;
        o += COUTP[l];
    }
    c->nparam = o;
    c->nchunk = (int)((c->npts + WG_CHUNK - 1) / WG_CHUNK);
    // PS3D_CNN_WGRAD=fp32 keeps the fp32 weight-gradient reduction (for comparison)
    if (const char *ev = std::getenv("PS3D_CNN_WGRAD")) c->wg_mfma = std::string(ev) != "fp32";
    c->nlblk = (int)std::min<long>((c->npts + 255) / 256, 1024);
    const size_t P = (size_t)c->npts;
    auto fl = [&](float **p, size_t n) { return hipMalloc((void **)p, n * sizeof(float)); };
    auto bf = [&](uint16_t **p, size_t n) { return hipMalloc((void **)p, n * sizeof(uint16_t)); };
    hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = fl(&c->w, o);
    if (e == hipSuccess) e = fl(&c->gw, o);
    if (e == hipSuccess) e = fl(&c->m, o);
    if (e == hipSuccess) e = fl(&c->v, o);
    for (int l = 0; l < NL && e == hipSuccess; ++l) {
        e = bf(&c->wf[l], 27 * (size_t)COUTP[l] * CINP[l]);
        if (e == hipSuccess) e = bf(&c->wt[l], 27 * (size_t)COUTP[l] * CINP[l]);
    }
    if (e == hipSuccess) e = bf(&c->a0, P * 16);
    for (int l = 0; l < 3 && e == hipSuccess; ++l) {
        e = fl(&c->pre[l], P * 32);
        if (e == hipSuccess) e = bf(&c->act[l], P * 32);
    }
    if (e == hipSuccess) e = fl(&c->y, P * 16);
    if (e == hipSuccess) e = fl(&c->gy, P * 16);
    if (e == hipSuccess) e = bf(&c->gyb, P * 16);
    if (e == hipSuccess) e = fl(&c->gpA, P * 32);
    if (e == hipSuccess) e = fl(&c->gpB, P * 32);
    if (e == hipSuccess) e = bf(&c->gbA, P * 32);
    if (e == hipSuccess) e = bf(&c->gbB, P * 32);
    if (e == hipSuccess) e = fl(&c->gx, P * 16);
    if (e == hipSuccess) e = fl(&c->part, (size_t)c->nchunk * 27 * 1024);
    if (e == hipSuccess) e = fl(&c->partb, (size_t)c->nchunk * 32);
    if (e == hipSuccess) e = fl(&c->lpart, (size_t)c->nlblk);
    if (e == hipSuccess) e = fl(&c->loss, 1);
    if (e == hipSuccess) e = fl(&c->loss_acc, 1);
    if (e == hipSuccess) e = fl(&c->tgt, P * 3);
    if (e == hipSuccess) e = hipMemset(c->a0, 0, P * 16 * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMemset(c->w, 0, o * sizeof(float));
    if (e == hipSuccess) e = hipMemset(c->m, 0, o * sizeof(float));
    if (e == hipSuccess) e = hipMemset(c->v, 0, o * sizeof(float));
    if (e == hipSuccess) e = hipMemset(c->loss_acc, 0, sizeof(float));
    // the clears ran on the null stream; the net's kernels run on non-blocking streams
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_cnn(c);
        return cfail(PS3D_E_NOMEM, std::string("closure allocation: ") + hipGetErrorString(e));
    }
    refresh(c, c->own);
    *out = c;
    return 0;
}

int ps3d_cnn_destroy(ps3d_cnn *c) { return free_cnn(c); }

int ps3d_cnn_set_weights(ps3d_cnn *c, const float *packed) {
    if (!c || !packed) return cfail(PS3D_E_ARG, "null argument");
    std::vector<float> pad;
    pack_to_padded(c, packed, pad);
    CHK(hipMemcpyAsync(c->w, pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice, c->own));
    CHK(hipMemsetAsync(c->m, 0, c->nparam * sizeof(float), c->own));
    CHK(hipMemsetAsync(c->v, 0, c->nparam * sizeof(float), c->own));
    c->adam_k = 0;
    refresh(c, c->own);
    CHK(hipStreamSynchronize(c->own));
    return 0;
}

int ps3d_cnn_get_weights(ps3d_cnn *c, float *packed) {
    if (!c || !packed) return cfail(PS3D_E_ARG, "null argument");
    std::vector<float> pad(c->nparam);
    CHK(hipMemcpyAsync(pad.data(), c->w, pad.size() * sizeof(float), hipMemcpyDeviceToHost, c->own));
    CHK(hipStreamSynchronize(c->own));
    padded_to_pack(c, pad, packed);
    return 0;
}

int ps3d_cnn_forward(ps3d_cnn *c, const float *x, float *y) {
    if (!c || !x || !y) return cfail(PS3D_E_ARG, "null argument");
    int st;
    if ((st = upload_input(c, x, c->a0, c->own))) return st;
    forward(c, c->a0, c->own);
    CHK(hipGetLastError());
    std::vector<float> hy((size_t)c->npts * 16);
    CHK(hipMemcpyAsync(hy.data(), c->y, hy.size() * sizeof(float), hipMemcpyDeviceToHost, c->own));
    CHK(hipStreamSynchronize(c->own));
    for (long p = 0; p < c->npts; ++p)
        for (int k = 0; k < 3; ++k) y[p * 3 + k] = hy[p * 16 + k];
    return 0;
}

int ps3d_cnn_grad(ps3d_cnn *c, const float *x, const float *target, float *loss, float *grad, float *gx) {
    if (!c || !x || !target) return cfail(PS3D_E_ARG, "null argument");
    int st;
    if ((st = upload_input(c, x, c->a0, c->own))) return st;
    CHK(hipMemcpyAsync(c->tgt, target, (size_t)c->npts * 3 * sizeof(float), hipMemcpyHostToDevice, c->own));
    forward(c, c->a0, c->own);
    backward(c, c->a0, c->tgt, gx != nullptr, c->own);
    CHK(hipGetLastError());
    if (loss) CHK(hipMemcpyAsync(loss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->own));
    std::vector<float> pad(c->nparam), hx;
    CHK(hipMemcpyAsync(pad.data(), c->gw, pad.size() * sizeof(float), hipMemcpyDeviceToHost, c->own));
    if (gx) {
        hx.resize((size_t)c->npts * 16);
        CHK(hipMemcpyAsync(hx.data(), c->gx, hx.size() * sizeof(float), hipMemcpyDeviceToHost, c->own));
    }
    CHK(hipStreamSynchronize(c->own));
    if (grad) padded_to_pack(c, pad, grad);
    if (gx)
        for (long p = 0; p < c->npts; ++p)
            for (int k = 0; k < 6; ++k) gx[p * 6 + k] = hx[p * 16 + k];
    return 0;
}

int ps3d_cnn_train_step(ps3d_cnn *c, const float *x, const float *target, float lr, float *loss) {
    if (!c || !x || !target) return cfail(PS3D_E_ARG, "null argument");
    int st;
    if ((st = upload_input(c, x, c->a0, c->own))) return st;
    CHK(hipMemcpyAsync(c->tgt, target, (size_t)c->npts * 3 * sizeof(float), hipMemcpyHostToDevice, c->own));
    forward(c, c->a0, c->own);
    backward(c, c->a0, c->tgt, false, c->own);
    adam(c, lr, c->own);
    CHK(hipGetLastError());
    if (loss) CHK(hipMemcpyAsync(loss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->own));
    CHK(hipStreamSynchronize(c->own));
    return 0;
}

int ps3d_cnn_dataset_reserve(ps3d_cnn *c, int nsamples) {
    if (!c || nsamples < 1) return cfail(PS3D_E_ARG, "invalid argument");
    (void)hipFree(c->dx);
    (void)hipFree(c->dt);
    c->dx = nullptr;
    c->dt = nullptr;
    c->ns = 0;
    CHK(hipMalloc((void **)&c->dx, (size_t)nsamples * c->npts * 16 * sizeof(uint16_t)));
    CHK(hipMalloc((void **)&c->dt, (size_t)nsamples * c->npts * 3 * sizeof(float)));
    c->ns = nsamples;
    return 0;
}

int ps3d_cnn_dataset_set(ps3d_cnn *c, int i, const float *x, const float *target) {
    if (!c || !x || !target || i < 0 || i >= c->ns) return cfail(PS3D_E_ARG, "invalid argument");
    int st;
    if ((st = upload_input(c, x, c->dx + (size_t)i * c->npts * 16, c->own))) return st;
    CHK(hipMemcpyAsync(c->dt + (size_t)i * c->npts * 3, target, (size_t)c->npts * 3 * sizeof(float),
                       hipMemcpyHostToDevice, c->own));
    CHK(hipStreamSynchronize(c->own));
    return 0;
}

int ps3d_cnn_train_epoch(ps3d_cnn *c, const int *order, int n, float lr, float *mean_loss) {
    if (!c || !order || n < 1) return cfail(PS3D_E_ARG, "invalid argument");
    for (int k = 0; k < n; ++k)
        if (order[k] < 0 || order[k] >= c->ns) return cfail(PS3D_E_ARG, "sample index out of range");
    CHK(hipMemsetAsync(c->loss_acc, 0, sizeof(float), c->own));
    for (int k = 0; k < n; ++k) {
        const uint16_t *x = c->dx + (size_t)order[k] * c->npts * 16;
        forward(c, x, c->own);
        backward(c, x, c->dt + (size_t)order[k] * c->npts * 3, false, c->own);
        adam(c, lr, c->own);
    }
    CHK(hipGetLastError());
    float acc = 0.0f;
    CHK(hipMemcpyAsync(&acc, c->loss_acc, sizeof(float), hipMemcpyDeviceToHost, c->own));
    CHK(hipStreamSynchronize(c->own));
    if (mean_loss) *mean_loss = acc / (float)n;
    return 0;
}

int ps3d_cnn_timing(ps3d_cnn *c, int reps, float *fwd_ms, float *train_ms) {
    if (!c || reps < 1) return cfail(PS3D_E_ARG, "invalid argument");
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    forward(c, c->a0, c->own);
    CHK(hipEventRecord(e0, c->own));
    for (int r = 0; r < reps; ++r) forward(c, c->a0, c->own);
    CHK(hipEventRecord(e1, c->own));
    CHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    if (fwd_ms) *fwd_ms = ms / reps;
    CHK(hipMemsetAsync(c->tgt, 0, (size_t)c->npts * 3 * sizeof(float), c->own));
    CHK(hipEventRecord(e0, c->own));
    for (int r = 0; r < reps; ++r) {
        forward(c, c->a0, c->own);
        backward(c, c->a0, c->tgt, false, c->own);
    }
    CHK(hipEventRecord(e1, c->own));
    CHK(hipEventSynchronize(e1));
    CHK(hipEventElapsedTime(&ms, e0, e1));
    if (train_ms) *train_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// solver hooks (ps3d_gpu.hip): the net's input buffer and the stage forward
// ---------------------------------------------------------------------------
namespace ps3d {

bool cnn_matches(const ps3d_cnn *c, int nx, int ny, int nz) { return c->nx == nx && c->ny == ny && c->nz1 == nz + 1; }

uint16_t *cnn_input(ps3d_cnn *c) { return c->a0; }

// forward on the stage input, output scaled into three solver-layout volumes
hipError_t cnn_apply(ps3d_cnn *c, const float *out_scale, int zs, double *o0, double *o1, double *o2,
                     hipStream_t s) {
    forward(c, c->a0, s);
    hipLaunchKernelGGL(k_unpack, dim3(blocks_for(c->npts, 256)), dim3(256), 0, s, (const float *)c->y, c->npts, c->ny,
                       c->nz1, zs, out_scale[0], out_scale[1], out_scale[2], o0, o1, o2);
    return hipGetLastError();
}

}  // namespace ps3d
