// Learned candidate screener: a 24-128-128-1 ReLU perceptron with BF16
// weights and FP32 accumulation, trained with Adam on FP32 master weights.
// Forward (inference and training), backward and the optimizer step are
// hand-written kernels; the host loop in nw_mlp_train keeps a whole training
// run on the device.  No reference counterpart (noci-rs screens candidates
// by exact evaluation only).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "nw_internal.h"

using namespace nw;
using namespace nwi;

namespace {

constexpr int NH = 128;  // hidden width (= workgroup size)
constexpr int MAXIN = 32;
constexpr int TSF = 16;  // samples per forward tile
constexpr int TSB = 32;  // samples per backward tile

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h)
{
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

struct Net {  // device pointers into one parameter block
    const uint16_t *W1, *W2, *W3;  // bf16 copies
    const float *b1, *b2, *b3;
    int nin;
};

// Forward over rows idx[r] (or r when idx is null) of X (n x nin).  One
// workgroup = one hidden unit per thread, TSF samples per tile, W2 held
// transposed in LDS as BF16 so thread j reads W2[j][k] conflict-free.
__global__ __launch_bounds__(NH) void k_mlp_fwd(const float *__restrict__ X, const int *__restrict__ idx, long long n,
                                                 Net net, float *__restrict__ y, float *__restrict__ H1,
                                                 float *__restrict__ H2)
{
    __shared__ uint16_t w2t[NH * NH];
    __shared__ float xs[TSF][MAXIN + 1];
    __shared__ float h1s[TSF][NH + 1];
    __shared__ float part[TSF][NH + 1];
    const int j = threadIdx.x, nin = net.nin;
    for (int e = j; e < NH * NH; e += NH) {
        const int r = e / NH, k = e % NH;  // W2[r][k] -> w2t[k][r]
        w2t[k * NH + r] = net.W2[e];
    }
    float w1[MAXIN];
#pragma unroll
    for (int i = 0; i < MAXIN; ++i) w1[i] = i < nin ? bf2f(net.W1[j * nin + i]) : 0.f;
    const float b1 = net.b1[j], b2 = net.b2[j], w3 = bf2f(net.W3[j]), b3 = net.b3[0];
    const long long ntile = (n + TSF - 1) / TSF;
    for (long long t = blockIdx.x; t < ntile; t += gridDim.x) {
        const long long r0 = t * TSF;
        __syncthreads();
        for (int e = j; e < TSF * MAXIN; e += NH) {
            const int s = e / MAXIN, i = e % MAXIN;
            const long long r = r0 + s;
            float v = 0.f;
            if (r < n && i < nin) v = X[(size_t)(idx ? idx[r] : r) * nin + i];
            xs[s][i] = v;
        }
        __syncthreads();
#pragma unroll 4
        for (int s = 0; s < TSF; ++s) {
            float z = b1;
#pragma unroll
            for (int i = 0; i < MAXIN; ++i) z += w1[i] * xs[s][i];
            const float h = z > 0.f ? z : 0.f;
            h1s[s][j] = h;
            if (H1 && r0 + s < n) H1[(size_t)(r0 + s) * NH + j] = h;
        }
        __syncthreads();
        float z2[TSF];
#pragma unroll
        for (int s = 0; s < TSF; ++s) z2[s] = b2;
        for (int k = 0; k < NH; ++k) {
            const float w = bf2f(w2t[k * NH + j]);
#pragma unroll
            for (int s = 0; s < TSF; ++s) z2[s] += w * h1s[s][k];
        }
#pragma unroll
        for (int s = 0; s < TSF; ++s) {
            const float h = z2[s] > 0.f ? z2[s] : 0.f;
            if (H2 && r0 + s < n) H2[(size_t)(r0 + s) * NH + j] = h;
            part[s][j] = w3 * h;
        }
        __syncthreads();
        if (j < TSF && r0 + j < n) {
            float o = b3;
            for (int k = 0; k < NH; ++k) o += part[j][k];
            y[r0 + j] = o;
        }
    }
}

typedef __bf16 mlp_bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short mlp_u8 __attribute__((ext_vector_type(8)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// v = h + m + l in BF16, each round to nearest (24 significant bits in all).
__device__ inline void mlp_split3(float v, __bf16 &h, __bf16 &m, __bf16 &l)
{
    h = (__bf16)v;
    const float r = v - (float)h;
    m = (__bf16)r;
    l = (__bf16)(r - (float)m);
}

constexpr int IW = 4;  // waves per inference workgroup, 16 samples per wave tile

// Inference y = W3 relu(W2 relu(W1 x + b1) + b2) + b3 (the forward of
// k_mlp_fwd without stored activations) on BF16 MFMA,
// v_mfma_f32_16x16x32_bf16 with D[sample][unit] += A[sample][k] B[k][unit]:
// the BF16 weights are exact operands, x and the layer-1 activations are
// split into three BF16 parts (hi + mid + lo), so every product carries
// FP32 accuracy and sums run in FP32.  Lane l holds A[l % 16][8 (l / 16) +
// 0..7], B[8 (l / 16) + 0..7][l % 16] and D[4 (l / 16) + 0..3][l % 16];
// layer-1 results pass to layer 2 through LDS, W2 stays in LDS.  All waves
// of a workgroup run the same number of tile rounds (inactive ones idle),
// so the barriers are uniform.
__global__ __launch_bounds__(64 * IW) void k_mlp_infer(const float *__restrict__ X, long long n, Net net,
                                                       float *__restrict__ y)
{
    constexpr int LW = NH + 8;       // bf16 per LDS row of W2 (16-byte pad)
    constexpr int LH = NH + 4;       // floats per LDS row of the activations
    __shared__ uint16_t w2s[NH * LW];  // W2[out][in]
    __shared__ float hs[IW][16 * LH];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int nin = net.nin;
    for (int e = tid; e < NH * NH; e += 64 * IW) w2s[(e / NH) * LW + e % NH] = net.W2[e];
    mlp_bf8 b1f[NH / 16];
    float bias1[NH / 16], bias2[NH / 16], w3v[NH / 16];
#pragma unroll
    for (int ob = 0; ob < NH / 16; ++ob) {
        const int j = 16 * ob + lr;
        mlp_u8 u;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = 8 * lk + q;
            u[q] = k < nin ? net.W1[j * nin + k] : (unsigned short)0;
        }
        b1f[ob] = __builtin_bit_cast(mlp_bf8, u);
        bias1[ob] = net.b1[j];
        bias2[ob] = net.b2[j];
        w3v[ob] = bf2f(net.W3[j]);
    }
    const float b3 = net.b3[0];
    float *hw = hs[wave];
    const long long ntile = (n + 15) / 16;
    for (long long tb = blockIdx.x; tb * IW < ntile; tb += gridDim.x) {
        const long long t = tb * IW + wave;
        const bool active = t < ntile;
        const long long r0 = t * 16;
        __syncthreads();  // W2 in place (first round); this wave's previous tile read
        if (active) {
            // layer 1
            mlp_bf8 xh, xm, xl;
            const long long r = r0 + lr;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int k = 8 * lk + q;
                const float v = (r < n && k < nin) ? X[(size_t)r * nin + k] : 0.f;
                __bf16 h, m, l;
                mlp_split3(v, h, m, l);
                xh[q] = h;
                xm[q] = m;
                xl[q] = l;
            }
#pragma unroll
            for (int ob = 0; ob < NH / 16; ++ob) {
                mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, b1f[ob], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xm, b1f[ob], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, b1f[ob], acc, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float z = acc[i] + bias1[ob];
                    hw[(4 * lk + i) * LH + 16 * ob + lr] = z > 0.f ? z : 0.f;
                }
            }
        }
        __syncthreads();
        if (active) {
            // layer 2
            mlp_f4 acc2[NH / 16];
#pragma unroll
            for (int ob = 0; ob < NH / 16; ++ob) acc2[ob] = mlp_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kc = 0; kc < NH / 32; ++kc) {
                mlp_bf8 ah, am, al;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    __bf16 h, m, l;
                    mlp_split3(hw[lr * LH + 32 * kc + 8 * lk + q], h, m, l);
                    ah[q] = h;
                    am[q] = m;
                    al[q] = l;
                }
#pragma unroll
                for (int ob = 0; ob < NH / 16; ++ob) {
                    const mlp_bf8 bw = *reinterpret_cast<const mlp_bf8 *>(&w2s[(16 * ob + lr) * LW + 32 * kc + 8 * lk]);
                    acc2[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bw, acc2[ob], 0, 0, 0);
                    acc2[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bw, acc2[ob], 0, 0, 0);
                    acc2[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bw, acc2[ob], 0, 0, 0);
                }
            }
            // layer 3: each lane's 8 units for its 4 samples, then the 16 lanes of a sample group
            float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ob = 0; ob < NH / 16; ++ob)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float z = acc2[ob][i] + bias2[ob];
                    part[i] += w3v[ob] * (z > 0.f ? z : 0.f);
                }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) part[i] += __shfl_xor(part[i], o, 16);
            if (lr == 0) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const long long row = r0 + 4 * lk + i;
                    if (row < n) y[row] = b3 + part[i];
                }
            }
        }
    }
}

// Backward of the weighted mean squared error L = sum_r w_r (y_r - t_r)^2 / (2 n)
// for the rows of one forward pass (H1, H2, y stored by k_mlp_fwd).  Each
// workgroup sums the gradients of its tiles in a fixed order into its own
// partial vector (gpart + blockIdx.x * stride; the loss in the last slot);
// every parameter belongs to one thread, so no atomics are needed and
// k_mlp_reduce adds the partials in workgroup order: the result is the same
// bit for bit on every run.
__global__ __launch_bounds__(NH) void k_mlp_bwd(const float *__restrict__ X, const int *__restrict__ idx, long long n,
                                                 double inv_n, Net net, const float *__restrict__ H1,
                                                 const float *__restrict__ H2, const float *__restrict__ y,
                                                 const float *__restrict__ tgt, const float *__restrict__ wt,
                                                 float *__restrict__ gpart, long long stride)
{
    float *g = gpart + (size_t)blockIdx.x * stride;
    float *loss = g + stride - 1;
    bool first = true;  // the block's first tile stores, later tiles add
    auto acc = [&](float *p, float v) { *p = first ? v : *p + v; };
    __shared__ float xs[TSB][MAXIN + 1];
    __shared__ float h1s[TSB][NH + 1];
    __shared__ float dh2s[TSB][NH + 1];
    __shared__ float dys[TSB];
    const int j = threadIdx.x, nin = net.nin;
    float *gW1 = g, *gb1 = gW1 + NH * nin, *gW2 = gb1 + NH, *gb2 = gW2 + NH * NH, *gW3 = gb2 + NH, *gb3 = gW3 + NH;
    const float w3j = bf2f(net.W3[j]);
    const long long ntile = (n + TSB - 1) / TSB;
    for (long long t = blockIdx.x; t < ntile; t += gridDim.x) {
        const long long r0 = t * TSB;
        __syncthreads();
        if (j < TSB) {
            const long long r = r0 + j;
            float d = 0.f, l = 0.f;
            if (r < n) {
                const long long src = idx ? idx[r] : r;
                const float e = y[r] - tgt[src], w = wt ? wt[src] : 1.f;
                d = (float)(w * e * inv_n);
                l = (float)(0.5 * w * e * e * inv_n);
            }
            dys[j] = d;
            // tile loss: reduce within the first wave
            float ls = l;
#pragma unroll
            for (int o = 16; o >= 1; o >>= 1) ls += __shfl_xor(ls, o, 32);
            if (j == 0) acc(loss, ls);
        }
        for (int e = j; e < TSB * MAXIN; e += NH) {
            const int s = e / MAXIN, i = e % MAXIN;
            const long long r = r0 + s;
            xs[s][i] = (r < n && i < nin) ? X[(size_t)(idx ? idx[r] : r) * nin + i] : 0.f;
        }
        for (int s = 0; s < TSB; ++s) {
            const long long r = r0 + s;
            h1s[s][j] = r < n ? H1[(size_t)r * NH + j] : 0.f;
        }
        __syncthreads();
        // layer 3 and the hidden-2 error
        float gw3 = 0.f, gb3s = 0.f, gb2j = 0.f;
        for (int s = 0; s < TSB; ++s) {
            const long long r = r0 + s;
            const float h2 = r < n ? H2[(size_t)r * NH + j] : 0.f;
            gw3 += dys[s] * h2;
            gb3s += dys[s];
            const float d2 = h2 > 0.f ? dys[s] * w3j : 0.f;
            dh2s[s][j] = d2;
            gb2j += d2;
        }
        acc(gW3 + j, gw3);
        acc(gb2 + j, gb2j);
        if (j == 0) acc(gb3, gb3s);
        __syncthreads();
        // gW2[r][k] for column k = j (coalesced over threads per row r)
        for (int r = 0; r < NH; ++r) {
            float a = 0.f;
#pragma unroll 8
            for (int s = 0; s < TSB; ++s) a += dh2s[s][r] * h1s[s][j];
            acc(gW2 + r * NH + j, a);
        }
        // hidden-1 error for unit k = j, then gW1 row j and gb1
        float d1[TSB];
#pragma unroll
        for (int s = 0; s < TSB; ++s) d1[s] = 0.f;
        for (int r = 0; r < NH; ++r) {
            const float w = bf2f(net.W2[r * NH + j]);
#pragma unroll
            for (int s = 0; s < TSB; ++s) d1[s] += w * dh2s[s][r];
        }
        float gb1j = 0.f;
#pragma unroll
        for (int s = 0; s < TSB; ++s) {
            d1[s] = h1s[s][j] > 0.f ? d1[s] : 0.f;
            gb1j += d1[s];
        }
        acc(gb1 + j, gb1j);
        for (int i = 0; i < nin; ++i) {
            float a = 0.f;
#pragma unroll
            for (int s = 0; s < TSB; ++s) a += d1[s] * xs[s][i];
            acc(gW1 + j * nin + i, a);
        }
        first = false;
    }
}

// g[p] = sum over workgroups b (in order) of gpart[b][p]; the loss slot is
// added to *loss.
__global__ void k_mlp_reduce(const float *__restrict__ gpart, long long stride, int nblk, float *__restrict__ g,
                             float *__restrict__ loss)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= stride) return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += gpart[(size_t)b * stride + p];
    if (p < stride - 1) g[p] = s;
    else *loss += s;
}

__global__ void k_fill_f32(float *p, long long n, float v)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        p[e] = v;
}

// Adam on the FP32 master weights; refreshes the BF16 copies of the weight
// matrices (biases stay FP32).  nw = number of leading weight-matrix entries
// of each segment is encoded by the mask array: bfidx[e] >= 0 gives the BF16
// slot of master entry e.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, const int *__restrict__ bfidx, uint16_t *__restrict__ wb,
                       long long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const float gg = g[e];
        const float mm = b1 * m[e] + (1.f - b1) * gg;
        const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
        m[e] = mm;
        v[e] = vv;
        const float mh = mm / c1, vh = vv / c2;
        const float nw = w[e] - lr * mh / (sqrtf(vh) + eps);
        w[e] = nw;
        if (bfidx[e] >= 0) wb[bfidx[e]] = f2bf(nw);
    }
}

__global__ void k_to_bf16(const float *__restrict__ w, const int *__restrict__ bfidx, uint16_t *__restrict__ wb, long long n)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        if (bfidx[e] >= 0) wb[bfidx[e]] = f2bf(w[e]);
}

}  // namespace

struct nw_mlp {
    int device = 0, nin = 24;
    hipStream_t st = nullptr;
    std::mutex mu;
    std::string err;
    long long np = 0, nwb = 0;  // master parameters, bf16 weight entries
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *loss = nullptr;
    uint16_t *wb = nullptr;
    int *bfidx = nullptr;
    long long adam_t = 0;
    float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    DBuf X, T, Wt, H1, H2, Y, perm, gpart;
    ~nw_mlp()
    {
        (void)hipSetDevice(device);
        if (st) (void)hipStreamSynchronize(st);
        for (void *p : {(void *)w, (void *)m, (void *)v, (void *)g, (void *)loss, (void *)wb, (void *)bfidx})
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    Net net() const
    {
        Net n{};
        n.W1 = wb;
        n.W2 = wb + (size_t)NH * nin;
        n.W3 = wb + (size_t)NH * nin + NH * NH;
        n.b1 = w + (size_t)NH * nin;
        n.b2 = w + (size_t)NH * nin + NH + NH * NH;
        n.b3 = w + (size_t)NH * nin + NH + NH * NH + NH + NH;
        n.nin = nin;
        return n;
    }
};

namespace {

int mfail(nw_mlp *p, int code, const std::string &s)
{
    if (p) p->err = s;
    return code;
}

template <class F>
int mguard(nw_mlp *p, F &&f)
{
    if (!p) return NW_E_ARG;
    std::lock_guard<std::mutex> lk(p->mu);
    try {
        hipError_t e = hipSetDevice(p->device);
        if (e != hipSuccess) return mfail(p, NW_E_HIP, hipGetErrorString(e));
        return f();
    } catch (const NwError &e) {
        (void)hipStreamSynchronize(p->st);
        return mfail(p, e.code, e.what());
    } catch (const std::exception &e) {
        return mfail(p, NW_E_ARG, e.what());
    }
}

const float *dev_f32(nw_mlp *p, DBuf &b, const float *src, size_t count)
{
    if (!src) return nullptr;
    if (is_device_ptr(src)) return src;
    float *d = b.get<float>(count);
    HIPCHK(hipMemcpyAsync(d, src, count * sizeof(float), hipMemcpyHostToDevice, p->st));
    return d;
}

int grid_for(long long ntile) { return (int)std::max<long long>(1, std::min<long long>(ntile, 2048)); }

void forward(nw_mlp *p, const float *dX, const int *didx, long long n, float *dy, float *dH1, float *dH2)
{
    if (!didx && !dH1 && !dH2 && p->nin <= 32) {  // inference: BF16 MFMA, no stored activations
        const long long tiles = (n + 15) / 16;
        hipLaunchKernelGGL(k_mlp_infer, dim3(grid_for((tiles + IW - 1) / IW)), dim3(64 * IW), 0, p->st, dX, n,
                           p->net(), dy);
        return;
    }
    hipLaunchKernelGGL(k_mlp_fwd, dim3(grid_for((n + TSF - 1) / TSF)), dim3(NH), 0, p->st, dX, didx, n, p->net(), dy,
                       dH1, dH2);
}

// One forward + backward over rows (didx or all), gradient into p->g, loss into *dloss.
void fwd_bwd(nw_mlp *p, const float *dX, const int *didx, long long n, const float *dT, const float *dW, float *dloss)
{
    float *H1 = p->H1.get<float>((size_t)n * NH), *H2 = p->H2.get<float>((size_t)n * NH), *Y = p->Y.get<float>(n);
    forward(p, dX, didx, n, Y, H1, H2);
    const int nblk = (int)std::max<long long>(1, std::min<long long>((n + TSB - 1) / TSB, 256));
    const long long stride = p->np + 1;
    float *gp = p->gpart.get<float>((size_t)nblk * stride);
    hipLaunchKernelGGL(k_mlp_bwd, dim3(nblk), dim3(NH), 0, p->st, dX, didx, n, 1.0 / (double)n, p->net(), H1, H2, Y, dT,
                       dW, gp, stride);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((unsigned)((stride + 255) / 256)), dim3(256), 0, p->st, gp, stride, nblk, p->g,
                       dloss);
}

void adam(nw_mlp *p, float lr)
{
    p->adam_t++;
    const float c1 = 1.f - std::pow(p->beta1, (float)p->adam_t), c2 = 1.f - std::pow(p->beta2, (float)p->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(80), dim3(256), 0, p->st, p->w, p->m, p->v, p->g, p->bfidx, p->wb, p->np, lr,
                       p->beta1, p->beta2, p->eps, c1, c2);
}

}  // namespace

extern "C" {

int nw_mlp_create(nw_mlp **out, int device, int n_in, uint64_t seed)
{
    if (!out) return NW_E_ARG;
    *out = nullptr;
    if (n_in < 1 || n_in > MAXIN) return NW_E_SHAPE;
    nw_mlp *p = new nw_mlp();
    p->device = device;
    p->nin = n_in;
    int rc = mguard(p, [&]() -> int {
        HIPCHK(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
        const long long n1 = (long long)NH * n_in, n2 = (long long)NH * NH;
        p->np = n1 + NH + n2 + NH + NH + 1;
        p->nwb = n1 + n2 + NH;
        p->w = dalloc<float>(p->np);
        p->m = dalloc<float>(p->np);
        p->v = dalloc<float>(p->np);
        p->g = dalloc<float>(p->np);
        p->loss = dalloc<float>(1024);
        p->wb = dalloc<uint16_t>(p->nwb);
        p->bfidx = dalloc<int>(p->np);
        // layout: W1 [NH x nin], b1 [NH], W2 [NH x NH], b2 [NH], W3 [NH], b3 [1]
        std::vector<int> bi(p->np, -1);
        long long o = 0, q = 0;
        for (long long e = 0; e < n1; ++e) bi[o++] = (int)q++;
        o += NH;
        for (long long e = 0; e < n2; ++e) bi[o++] = (int)q++;
        o += NH;
        for (long long e = 0; e < NH; ++e) bi[o++] = (int)q++;
        HIPCHK(hipMemcpy(p->bfidx, bi.data(), bi.size() * sizeof(int), hipMemcpyHostToDevice));
        // He-uniform initialisation, zero biases
        std::mt19937_64 rng(seed);
        std::vector<float> w(p->np, 0.f);
        auto fill = [&](long long off, long long count, int fan_in) {
            std::uniform_real_distribution<float> U(-std::sqrt(6.f / fan_in), std::sqrt(6.f / fan_in));
            for (long long e = 0; e < count; ++e) w[off + e] = U(rng);
        };
        fill(0, n1, n_in);
        fill(n1 + NH, n2, NH);
        fill(n1 + NH + n2 + NH, NH, NH);
        HIPCHK(hipMemcpy(p->w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(p->m, 0, p->np * 4));
        HIPCHK(hipMemset(p->v, 0, p->np * 4));
        hipLaunchKernelGGL(k_to_bf16, dim3(80), dim3(256), 0, p->st, p->w, p->bfidx, p->wb, p->np);
        HIPCHK(hipStreamSynchronize(p->st));
        return NW_OK;
    });
    if (rc != NW_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return NW_OK;
}

int nw_mlp_num_params(nw_mlp *p, long long *n)
{
    return mguard(p, [&]() -> int {
        if (n) *n = p->np;
        return NW_OK;
    });
}

int nw_mlp_set_params(nw_mlp *p, const float *params, long long n)
{
    return mguard(p, [&]() -> int {
        if (!params || n != p->np) return mfail(p, NW_E_SHAPE, "parameter count mismatch");
        HIPCHK(hipMemcpyAsync(p->w, params, n * 4, hipMemcpyDefault, p->st));
        HIPCHK(hipMemsetAsync(p->m, 0, n * 4, p->st));
        HIPCHK(hipMemsetAsync(p->v, 0, n * 4, p->st));
        p->adam_t = 0;
        hipLaunchKernelGGL(k_to_bf16, dim3(80), dim3(256), 0, p->st, p->w, p->bfidx, p->wb, p->np);
        HIPCHK(hipStreamSynchronize(p->st));
        return NW_OK;
    });
}

int nw_mlp_get_params(nw_mlp *p, float *params, long long n)
{
    return mguard(p, [&]() -> int {
        if (!params || n != p->np) return mfail(p, NW_E_SHAPE, "parameter count mismatch");
        HIPCHK(hipMemcpyAsync(params, p->w, n * 4, hipMemcpyDefault, p->st));
        HIPCHK(hipStreamSynchronize(p->st));
        return NW_OK;
    });
}

int nw_mlp_forward(nw_mlp *p, const float *X, long long n, float *y)
{
    return mguard(p, [&]() -> int {
        if (!X || !y || n <= 0) return mfail(p, NW_E_ARG, "bad arguments");
        const float *dX = dev_f32(p, p->X, X, (size_t)n * p->nin);
        const bool ydev = is_device_ptr(y);
        float *dy = ydev ? y : p->Y.get<float>(n);
        forward(p, dX, nullptr, n, dy, nullptr, nullptr);
        HIPCHK(hipGetLastError());
        if (!ydev) HIPCHK(hipMemcpyAsync(y, dy, n * 4, hipMemcpyDeviceToHost, p->st));
        HIPCHK(hipStreamSynchronize(p->st));
        return NW_OK;
    });
}

int nw_mlp_grad(nw_mlp *p, const float *X, const float *t, const float *w, long long n, float *loss, float *grad)
{
    return mguard(p, [&]() -> int {
        if (!X || !t || !loss || !grad || n <= 0) return mfail(p, NW_E_ARG, "bad arguments");
        const float *dX = dev_f32(p, p->X, X, (size_t)n * p->nin);
        const float *dT = dev_f32(p, p->T, t, n);
        const float *dW = dev_f32(p, p->Wt, w, n);
        hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, p->st, p->loss, 1LL, 0.f);
        fwd_bwd(p, dX, nullptr, n, dT, dW, p->loss);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(loss, p->loss, 4, hipMemcpyDeviceToHost, p->st));
        HIPCHK(hipMemcpyAsync(grad, p->g, p->np * 4, hipMemcpyDefault, p->st));
        HIPCHK(hipStreamSynchronize(p->st));
        return NW_OK;
    });
}

int nw_mlp_step(nw_mlp *p, const float *X, const float *t, const float *w, long long n, float lr, float *loss)
{
    return mguard(p, [&]() -> int {
        if (!X || !t || n <= 0) return mfail(p, NW_E_ARG, "bad arguments");
        const float *dX = dev_f32(p, p->X, X, (size_t)n * p->nin);
        const float *dT = dev_f32(p, p->T, t, n);
        const float *dW = dev_f32(p, p->Wt, w, n);
        hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, p->st, p->loss, 1LL, 0.f);
        fwd_bwd(p, dX, nullptr, n, dT, dW, p->loss);
        adam(p, lr);
        HIPCHK(hipGetLastError());
        if (loss) HIPCHK(hipMemcpyAsync(loss, p->loss, 4, hipMemcpyDeviceToHost, p->st));
        HIPCHK(hipStreamSynchronize(p->st));
        return NW_OK;
    });
}

int nw_mlp_train(nw_mlp *p, const float *X, const float *t, const float *w, long long n, int epochs, int batch,
                 float lr, uint64_t seed, float *epoch_loss)
{
    return mguard(p, [&]() -> int {
        if (!X || !t || n <= 0 || epochs <= 0 || epochs > 1024 || batch <= 0) return mfail(p, NW_E_ARG, "bad arguments");
        const float *dX = dev_f32(p, p->X, X, (size_t)n * p->nin);
        const float *dT = dev_f32(p, p->T, t, n);
        const float *dW = dev_f32(p, p->Wt, w, n);
        hipLaunchKernelGGL(k_fill_f32, dim3(4), dim3(256), 0, p->st, p->loss, 1024LL, 0.f);
        std::mt19937_64 rng(seed);
        std::vector<int> perm(n);
        int *dperm = p->perm.get<int>(n);
        std::vector<float> tmp(1024);
        const long long nb = (n + batch - 1) / batch;
        for (int ep = 0; ep < epochs; ++ep) {
            for (long long k = 0; k < n; ++k) perm[k] = (int)k;
            std::shuffle(perm.begin(), perm.end(), rng);
            HIPCHK(hipMemcpyAsync(dperm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, p->st));
            for (long long b = 0; b < nb; ++b) {
                const long long r0 = b * batch, m = std::min<long long>(batch, n - r0);
                fwd_bwd(p, dX, dperm + r0, m, dT, dW, p->loss + ep);
                adam(p, lr);
            }
            HIPCHK(hipStreamSynchronize(p->st));  // perm is reused next epoch
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(tmp.data(), p->loss, epochs * 4, hipMemcpyDeviceToHost));
        if (epoch_loss)
            for (int ep = 0; ep < epochs; ++ep) epoch_loss[ep] = tmp[ep] / (float)nb;
        return NW_OK;
    });
}

const char *nw_mlp_last_error(nw_mlp *p)
{
    return p ? p->err.c_str() : "null network";
}

void nw_mlp_destroy(nw_mlp *p)
{
    delete p;
}

}  // extern "C"
