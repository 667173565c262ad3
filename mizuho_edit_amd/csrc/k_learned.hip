// k_learned.hip -- the pressure-correction network: a 3-D convolutional
// encoder-decoder (4 levels, 16/32/48/64 channels, 3x3x3 kernels), BF16
// weights and activations with FP32 accumulation on MFMA
// (v_mfma_f32_16x16x32_bf16), and its training kernels.
//
// Architecture (inputs pooled by 4 per axis, level 1 = N/4):
//   x0 = pool4(log k, p, r) -> 8 channels (3 used)      L1
//   s1 = relu(conv 8->16)            L1      s2 = relu(conv pool(s1) 16->32)  L2
//   s3 = relu(conv pool(s2) 32->48)  L3      b4 = relu(conv pool(s3) 48->64)  L4
//   d3 = relu(conv [up(b4), s3] 112->48)  L3
//   d2 = relu(conv [up(d3), s2]  80->32)  L2
//   d1 = relu(conv [up(d2), s1]  48->16)  L1
//   y  =      conv d1 16->1               L1, FP32
// The upsampling and concatenation of the decoder are never materialised:
// the convolution gathers channels [0, ca) from the coarse tensor at
// (x/2, y/2, z/2) and [ca, ca+cb) from the skip tensor.
//
// Forward convolution = implicit GEMM, M = voxels (16 per wavefront),
// N = output channels (16 per MFMA), K = 27 * Cin in chunks of 32; each lane
// loads its 8 K-values as one 16-byte vector (Cin is a multiple of 8).
// Training (offline, not on the timed path) runs the data and weight
// gradients as implicit GEMMs on the same MFMA instruction (BF16 operands,
// FP32 accumulation), with a scale-invariant loss and Adam.
#include "wpf_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace wpf {

struct NetLayer {
    int cin, cout;    // cin includes both concatenated sources
    int ca, cb;       // channels from the first (maybe coarse) and second source
    int coarse_a;     // first source is at half resolution (decoder)
    int level;        // resolution level (1..4) of the output
    int kpad, coutpad;
    int64_t woff;     // offset of W in the flat parameter vector (b follows W)
};

static const int kNL = 8;
static const int kBiasBlocks = 256;  // voxel chunks of the bias gradient
static const int kCIN0 = 8;
static const int kPool = 4;

static void layer_table(NetLayer* L)
{
    // name: e1 e2 e3 e4 d3 d2 d1 out
    const int cfg[kNL][6] = {
        {8, 16, 8, 0, 0, 1},   {16, 32, 16, 0, 0, 2},  {32, 48, 32, 0, 0, 3},
        {48, 64, 48, 0, 0, 4}, {112, 48, 64, 48, 1, 3}, {80, 32, 48, 32, 1, 2},
        {48, 16, 32, 16, 1, 1}, {16, 1, 16, 0, 0, 1}};
    int64_t off = 0;
    for (int i = 0; i < kNL; i++) {
        L[i].cin = cfg[i][0];
        L[i].cout = cfg[i][1];
        L[i].ca = cfg[i][2];
        L[i].cb = cfg[i][3];
        L[i].coarse_a = cfg[i][4];
        L[i].level = cfg[i][5];
        L[i].kpad = ((27 * L[i].cin + 31) / 32) * 32;
        L[i].coutpad = ((L[i].cout + 15) / 16) * 16;
        L[i].woff = off;
        off += (int64_t)L[i].cout * 27 * L[i].cin + L[i].cout;
    }
}

static int64_t count_params()
{
    NetLayer L[kNL];
    layer_table(L);
    return L[kNL - 1].woff + (int64_t)L[kNL - 1].cout * 27 * L[kNL - 1].cin + L[kNL - 1].cout;
}

struct Geo {
    int nx, ny, nz;
    int64_t nv;
};
static Geo geo(int nx, int ny, int nz) { return Geo{nx, ny, nz, (int64_t)nx * ny * nz}; }

struct Src {
    const bf16* a;
    int ca;
    int a_coarse;
    const bf16* b;
    int cb;
};

// ---------------------------------------------------------------------------
// forward convolution on MFMA
// ---------------------------------------------------------------------------
template <int NT, bool RELU, bool OUTF32>
__global__ void __launch_bounds__(256) k_conv_fwd(Geo g, Src s, const bf16* __restrict__ wt,
                                                  const float* __restrict__ bias, int kpad, int cout,
                                                  void* __restrict__ out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t tile = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (tile >= g.nv) return;  // whole wavefront leaves together; no barriers below
    const int cin = s.ca + s.cb;
    const int ktot = 27 * cin;
    const int64_t v = tile + (lane & 15);
    const bool vin = v < g.nv;
    int x = 0, y = 0, z = 0;
    if (vin) {
        x = (int)(v % g.nx);
        y = (int)((v / g.nx) % g.ny);
        z = (int)(v / ((int64_t)g.nx * g.ny));
    }
    const int kq = 8 * (lane >> 4);
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < kpad; k0 += 32) {
        const int kk = k0 + kq;
        bf16x8 a;
#pragma unroll
        for (int e = 0; e < 8; e++) a[e] = (bf16)0.0f;
        if (vin && kk < ktot) {
            const int tap = kk / cin, ci = kk - tap * cin;
            const int xx = x + tap % 3 - 1, yy = y + (tap / 3) % 3 - 1, zz = z + tap / 9 - 1;
            if (xx >= 0 && xx < g.nx && yy >= 0 && yy < g.ny && zz >= 0 && zz < g.nz) {
                if (ci < s.ca) {
                    int64_t va;
                    if (s.a_coarse)
                        va = (xx >> 1) + (int64_t)(g.nx >> 1) * ((yy >> 1) + (int64_t)(g.ny >> 1) * (zz >> 1));
                    else
                        va = xx + (int64_t)g.nx * (yy + (int64_t)g.ny * zz);
                    a = *reinterpret_cast<const bf16x8*>(s.a + va * s.ca + ci);
                } else {
                    const int64_t vb = xx + (int64_t)g.nx * (yy + (int64_t)g.ny * zz);
                    a = *reinterpret_cast<const bf16x8*>(s.b + vb * s.cb + (ci - s.ca));
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int n = t * 16 + (lane & 15);
            const bf16x8 bw = *reinterpret_cast<const bf16x8*>(wt + (int64_t)n * kpad + kk);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int n = t * 16 + (lane & 15);
        if (n >= cout) continue;
        const float bn = bias[n];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int64_t vo = tile + 4 * (lane >> 4) + e;
            if (vo >= g.nv) continue;
            float val = acc[t][e] + bn;
            if (RELU) val = fmaxf(val, 0.0f);
            if (OUTF32)
                reinterpret_cast<float*>(out)[vo * cout + n] = val;
            else
                reinterpret_cast<bf16*>(out)[vo * cout + n] = (bf16)val;
        }
    }
}

// packed BF16 weights Wt[n][k] (k = tap*cin + ci), zero-padded
__global__ void k_pack_w(const float* __restrict__ w, int cout, int cin, int coutpad, int kpad,
                         bf16* __restrict__ wt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)coutpad * kpad) return;
    const int n = (int)(i / kpad), k = (int)(i % kpad);
    const float v = (n < cout && k < 27 * cin) ? w[(int64_t)n * 27 * cin + k] : 0.0f;
    wt[i] = (bf16)v;
}

// average pool 2x2x2 of a channels-last BF16 tensor
__global__ void k_pool2(Geo cg, int C, const bf16* __restrict__ in, bf16* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cg.nv * C) return;
    const int c = (int)(i % C);
    const int64_t V = i / C;
    const int X = (int)(V % cg.nx), Y = (int)((V / cg.nx) % cg.ny), Z = (int)(V / ((int64_t)cg.nx * cg.ny));
    const int fx = 2 * cg.nx, fy = 2 * cg.ny;
    float s = 0.0f;
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
                const int64_t u = (2 * X + dx) + (int64_t)fx * ((2 * Y + dy) + (int64_t)fy * (2 * Z + dz));
                s += (float)in[u * C + c];
            }
    out[i] = (bf16)(0.125f * s);
}

// network input: planar FP32 fields [3][n] at full resolution -> level-1
// BF16 tensor with kCIN0 channels (mean over the kPool^3 children)
__global__ void k_input_pool(Geo fg, Geo cg, const float* __restrict__ fields, bf16* __restrict__ x0)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cg.nv * kCIN0) return;
    const int c = (int)(i % kCIN0);
    const int64_t V = i / kCIN0;
    if (c >= 3) {
        x0[i] = (bf16)0.0f;
        return;
    }
    const int X = (int)(V % cg.nx), Y = (int)((V / cg.nx) % cg.ny), Z = (int)(V / ((int64_t)cg.nx * cg.ny));
    const float* f = fields + (int64_t)c * fg.nv;
    float s = 0.0f;
    for (int dz = 0; dz < kPool; dz++)
        for (int dy = 0; dy < kPool; dy++)
            for (int dx = 0; dx < kPool; dx++)
                s += f[(kPool * X + dx) + (int64_t)fg.nx * ((kPool * Y + dy) + (int64_t)fg.ny * (kPool * Z + dz))];
    x0[i] = (bf16)(s / (float)(kPool * kPool * kPool));
}

__global__ void k_target_pool(Geo fg, Geo cg, const float* __restrict__ t, float* __restrict__ out)
{
    const int64_t V = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (V >= cg.nv) return;
    const int X = (int)(V % cg.nx), Y = (int)((V / cg.nx) % cg.ny), Z = (int)(V / ((int64_t)cg.nx * cg.ny));
    float s = 0.0f;
    for (int dz = 0; dz < kPool; dz++)
        for (int dy = 0; dy < kPool; dy++)
            for (int dx = 0; dx < kPool; dx++)
                s += t[(kPool * X + dx) + (int64_t)fg.nx * ((kPool * Y + dy) + (int64_t)fg.ny * (kPool * Z + dz))];
    out[V] = s / (float)(kPool * kPool * kPool);
}

// ---------------------------------------------------------------------------
// training kernels (FP32)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float src_at(const Src& s, const Geo& g, int xx, int yy, int zz, int ci)
{
    if (xx < 0 || xx >= g.nx || yy < 0 || yy >= g.ny || zz < 0 || zz >= g.nz) return 0.0f;
    if (ci < s.ca) {
        int64_t va;
        if (s.a_coarse)
            va = (xx >> 1) + (int64_t)(g.nx >> 1) * ((yy >> 1) + (int64_t)(g.ny >> 1) * (zz >> 1));
        else
            va = xx + (int64_t)g.nx * (yy + (int64_t)g.ny * zz);
        return (float)s.a[va * s.ca + ci];
    }
    const int64_t vb = xx + (int64_t)g.nx * (yy + (int64_t)g.ny * zz);
    return (float)s.b[vb * s.cb + (ci - s.ca)];
}

// The backward convolutions are implicit GEMMs on MFMA too, with the
// incoming gradient gz rounded to BF16 (weights and activations already
// are) and FP32 accumulation.
//
// Data gradient: gin[u][ci] = sum_tap sum_co W[co][tap][ci] gz[u - off(tap)][co]
// M = voxels u (16 per wavefront), N = input channels ci, K = (tap, co) with
// co padded to copad (a multiple of 8, so the 8 K-values a lane loads lie in
// one tap and are one 16-byte vector of the padded BF16 gz).
__global__ void k_gz_bf16(int64_t nv, int cout, int copad, const float* __restrict__ gz, bf16* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv * copad) return;
    const int c = (int)(i % copad);
    const int64_t v = i / copad;
    out[i] = (bf16)(c < cout ? gz[v * cout + c] : 0.0f);
}

// transposed packed weights WT[ci][tap * copad + co], zero padded
__global__ void k_pack_wT(const float* __restrict__ w, int cout, int cin, int copad, int cinpad, int kpadT,
                          bf16* __restrict__ wt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)cinpad * kpadT) return;
    const int ci = (int)(i / kpadT), k = (int)(i % kpadT);
    const int tap = k / copad, co = k % copad;
    const float v = (ci < cin && tap < 27 && co < cout) ? w[((int64_t)co * 27 + tap) * cin + ci] : 0.0f;
    wt[i] = (bf16)v;
}

template <int NT>
__global__ void __launch_bounds__(256) k_conv_bwd_data(Geo g, const bf16* __restrict__ gz16, int copad,
                                                       const bf16* __restrict__ wt, int kpadT, int cin,
                                                       float* __restrict__ gin)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t tile = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (tile >= g.nv) return;  // whole wavefront leaves together; no barriers below
    const int64_t v = tile + (lane & 15);
    const bool vin = v < g.nv;
    int x = 0, y = 0, z = 0;
    if (vin) {
        x = (int)(v % g.nx);
        y = (int)((v / g.nx) % g.ny);
        z = (int)(v / ((int64_t)g.nx * g.ny));
    }
    const int kq = 8 * (lane >> 4);
    const int ktot = 27 * copad;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < kpadT; k0 += 32) {
        const int kk = k0 + kq;
        bf16x8 a;
#pragma unroll
        for (int e = 0; e < 8; e++) a[e] = (bf16)0.0f;
        if (vin && kk < ktot) {
            const int tap = kk / copad, co = kk - tap * copad;
            // the output voxel that reads u through this tap
            const int xx = x - (tap % 3 - 1), yy = y - ((tap / 3) % 3 - 1), zz = z - (tap / 9 - 1);
            if (xx >= 0 && xx < g.nx && yy >= 0 && yy < g.ny && zz >= 0 && zz < g.nz)
                a = *reinterpret_cast<const bf16x8*>(gz16 + (xx + (int64_t)g.nx * (yy + (int64_t)g.ny * zz)) * copad + co);
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int n = t * 16 + (lane & 15);
            const bf16x8 bw = *reinterpret_cast<const bf16x8*>(wt + (int64_t)n * kpadT + kk);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int n = t * 16 + (lane & 15);
        if (n >= cin) continue;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int64_t vo = tile + 4 * (lane >> 4) + e;
            if (vo < g.nv) gin[vo * cin + n] = acc[t][e];
        }
    }
}

// Weight gradient: dW[co][tap][ci] = sum_v gz[v][co] in(v + off(tap))[ci]
// M = co (16 per tile, blockIdx.y), N = tap * cin + ci (16 per wavefront, 4
// wavefronts per block along blockIdx.x), K = voxels, split over blockIdx.z
// into chunks of vchunk (a multiple of 32); each chunk's partial dW goes to
// part[chunk][co][n] and k_sum_chunks adds the chunks in a fixed order.
// gzT is gz transposed to [co][v] in BF16 (zero rows beyond cout, zero
// voxels beyond nv up to nvpad), so a lane's 8 K-values are one vector.
__global__ void k_gz_bf16_t(int64_t nv, int64_t nvpad, int cout, int coutpad, const float* __restrict__ gz,
                            bf16* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)coutpad * nvpad) return;
    const int co = (int)(i / nvpad);
    const int64_t v = i % nvpad;
    out[i] = (bf16)((co < cout && v < nv) ? gz[v * cout + co] : 0.0f);
}

__global__ void __launch_bounds__(256) k_conv_bwd_w(Geo g, Src s, const bf16* __restrict__ gzT, int64_t nvpad,
                                                    int cout, int64_t vchunk, float* __restrict__ part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cin = s.ca + s.cb, ncol = 27 * cin;
    const int ncol0 = (blockIdx.x * 4 + wave) * 16;
    if (ncol0 >= ncol) return;  // whole wavefront leaves together; no barriers below
    const int n = ncol0 + (lane & 15);
    const bool nvalid = n < ncol;
    const int tap = nvalid ? n / cin : 0, ci = nvalid ? n - tap * cin : 0;
    const int ox = tap % 3 - 1, oy = (tap / 3) % 3 - 1, oz = tap / 9 - 1;
    const int co = blockIdx.y * 16 + (lane & 15);
    const bf16* arow = gzT + (int64_t)co * nvpad;
    const int64_t v0 = (int64_t)blockIdx.z * vchunk;
    const int64_t v1 = min(v0 + vchunk, g.nv);
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    // (the network's grids hold far fewer than 2^31 voxels: 32-bit indices)
    const int nv = (int)g.nv;
    for (int64_t vb = v0; vb < v1; vb += 32) {
        const int kv = (int)vb + 8 * (lane >> 4);
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(arow + kv);
        bf16x8 b;
        // (x, y, z) of the lane's first voxel, then stepped through the 8
        int x = kv % g.nx, y = (kv / g.nx) % g.ny, z = kv / (g.nx * g.ny);
#pragma unroll
        for (int e = 0; e < 8; e++) {
            float val = 0.0f;
            if (nvalid && kv + e < nv) val = src_at(s, g, x + ox, y + oy, z + oz, ci);
            b[e] = (bf16)val;
            if (++x == g.nx) {
                x = 0;
                if (++y == g.ny) {
                    y = 0;
                    ++z;
                }
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    if (!nvalid) return;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int r = blockIdx.y * 16 + 4 * (lane >> 4) + e;
        if (r < cout) part[((int64_t)blockIdx.z * cout + r) * ncol + n] = acc[e];
    }
}

// out[i] = sum over chunks c of part[c][i], c in order
__global__ void k_sum_chunks(int64_t count, int nchunk, const float* __restrict__ part, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int c = 0; c < nchunk; c++) s += part[(int64_t)c * count + i];
    out[i] = s;
}

// Bias gradient, first half: block b sums gz over its chunk of voxels for
// every channel into part[b][co].  Thread t reads channel t % cout of voxel
// v0 + t / cout + k R (R = 256 / cout rows), so a wavefront's loads are
// contiguous; the rows are added in a fixed order, and k_sum_chunks then
// adds the blocks in a fixed order.  cout <= 256.
__global__ void __launch_bounds__(256) k_bias_part(Geo g, int cout, int64_t chunk, const float* __restrict__ gz,
                                                   float* __restrict__ part)
{
    __shared__ float sh[256];
    const int R = 256 / cout;
    const int co = (int)threadIdx.x % cout, row = (int)threadIdx.x / cout;
    const int64_t v0 = (int64_t)blockIdx.x * chunk, v1 = min(v0 + chunk, g.nv);
    float s = 0.0f;
    if (row < R)
        for (int64_t v = v0 + row; v < v1; v += R) s += gz[v * cout + co];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (row == 0) {
        float t = 0.0f;
        for (int r = 0; r < R; r++) t += sh[r * cout + co];
        part[(int64_t)blockIdx.x * cout + co] = t;
    }
}

__global__ void k_relu_mask(int64_t n, float* __restrict__ g, const bf16* __restrict__ act)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !((float)act[i] > 0.0f)) g[i] = 0.0f;
}

// gradient of the concatenated decoder input [up(A) | B] split back:
// gA (coarse, ca channels) = sum over the 8 children; gB += second part
__global__ void k_upcat_bwd_a(Geo cg, int ca, int cb, const float* __restrict__ gin, float* __restrict__ ga)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cg.nv * ca) return;
    const int c = (int)(i % ca);
    const int64_t V = i / ca;
    const int X = (int)(V % cg.nx), Y = (int)((V / cg.nx) % cg.ny), Z = (int)(V / ((int64_t)cg.nx * cg.ny));
    const int fx = 2 * cg.nx, fy = 2 * cg.ny, cin = ca + cb;
    float s = 0.0f;
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++)
                s += gin[((2 * X + dx) + (int64_t)fx * ((2 * Y + dy) + (int64_t)fy * (2 * Z + dz))) * cin + c];
    ga[i] = s;
}
__global__ void k_upcat_bwd_b(Geo g, int ca, int cb, const float* __restrict__ gin, float* __restrict__ gb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nv * cb) return;
    const int c = (int)(i % cb);
    const int64_t u = i / cb;
    gb[i] = gin[u * (ca + cb) + ca + c];
}
// gfine[u][c] += gcoarse[parent(u)][c] / 8  (average-pool backward)
__global__ void k_pool_bwd_add(Geo fg, int C, const float* __restrict__ gc, float* __restrict__ gf)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fg.nv * C) return;
    const int c = (int)(i % C);
    const int64_t u = i / C;
    const int x = (int)(u % fg.nx), y = (int)((u / fg.nx) % fg.ny), z = (int)(u / ((int64_t)fg.nx * fg.ny));
    const int64_t V = (x >> 1) + (int64_t)(fg.nx >> 1) * ((y >> 1) + (int64_t)(fg.ny >> 1) * (z >> 1));
    gf[i] += 0.125f * gc[V * C + c];
}

// y.t, y.y, t.t partials (one value per block per quantity)
__global__ void __launch_bounds__(256) k_loss_dots(int64_t n, const float* __restrict__ y,
                                                   const float* __restrict__ t,
                                                   double* __restrict__ partials)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double a = y[i], b = t[i];
        v[0] += a * b;
        v[1] += a * a;
        v[2] += b * b;
    }
    block_reduce_store<3>(v, partials, nullptr, nullptr, gridDim.x, blockIdx.x);
}
// dL/dy of L = 1 - (y.t)^2/((y.y)(t.t)); dots in d[0..2]; loss into d[3]
__global__ void k_loss_grad(int64_t n, double* __restrict__ d, const float* __restrict__ y,
                            const float* __restrict__ t, float* __restrict__ gy)
{
    const double yt = d[0], yy = d[1], tt = d[2];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // a zero output or target carries no direction: loss 1, no gradient
    const bool ok = yy > 1e-300 && tt > 1e-300;
    if (i == 0) d[3] = ok ? 1.0 - yt * yt / (yy * tt) : 1.0;
    if (i < n) {
        const double a = ok ? -2.0 * yt / (yy * tt) : 0.0;
        const double b = ok ? 2.0 * yt * yt / (yy * yy * tt) : 0.0;
        gy[i] = (float)(a * t[i] + b * y[i]);
    }
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// full-resolution correction from the level-1 output: cell-centred trilinear
// interpolation (clamped at the domain faces), 0 on fixed cells
template <typename R>
__global__ void k_upsample_out(Geo fg, Geo cg, const float* __restrict__ y, const uint8_t* __restrict__ mask,
                               R* __restrict__ out)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= fg.nv) return;
    if (mask[u] & WPF_MASK_FIXED) {
        out[u] = (R)0;
        return;
    }
    const int x = (int)(u % fg.nx), yy = (int)((u / fg.nx) % fg.ny), z = (int)(u / ((int64_t)fg.nx * fg.ny));
    auto coord = [](int i, int n) {
        float c = ((float)i + 0.5f) / (float)kPool - 0.5f;
        c = fminf(fmaxf(c, 0.0f), (float)(n - 1));
        return c;
    };
    const float cx = coord(x, cg.nx), cy = coord(yy, cg.ny), cz = coord(z, cg.nz);
    const int x0 = min((int)cx, cg.nx - 1), y0 = min((int)cy, cg.ny - 1), z0 = min((int)cz, cg.nz - 1);
    const int x1 = min(x0 + 1, cg.nx - 1), y1 = min(y0 + 1, cg.ny - 1), z1 = min(z0 + 1, cg.nz - 1);
    const float fx = cx - x0, fy = cy - y0, fz = cz - z0;
    auto Y = [&](int a, int b, int c) { return y[a + (int64_t)cg.nx * (b + (int64_t)cg.ny * c)]; };
    const float v = (1 - fz) * ((1 - fy) * ((1 - fx) * Y(x0, y0, z0) + fx * Y(x1, y0, z0)) +
                                fy * ((1 - fx) * Y(x0, y1, z0) + fx * Y(x1, y1, z0))) +
                    fz * ((1 - fy) * ((1 - fx) * Y(x0, y0, z1) + fx * Y(x1, y0, z1)) +
                          fy * ((1 - fx) * Y(x0, y1, z1) + fx * Y(x1, y1, z1)));
    out[u] = (R)v;
}

// ---------------------------------------------------------------------------
// energy-norm loss
//
// For a correction u (the network's output upsampled to full resolution, 0
// at Dirichlet cells) and the error t it should approximate (r = A t), the
// exact line search along u removes the share
//     (u, r)^2 / ((u, A u) (t, r))
// of the A-norm of the error: the loss is 1 minus that share.  It depends
// on neither the scale of u nor that of t.
// ---------------------------------------------------------------------------
// out = A v on free cells (7-point: the sum of the cell's faces times v
// minus each face times the neighbour), 0 on Dirichlet cells; faces
// [3][n] hold the +x/+y/+z face of each cell, 0 on the global edge.
__global__ void k_energy_apply(Geo g, const float* __restrict__ faces, const uint8_t* __restrict__ mask,
                               const float* __restrict__ v, float* __restrict__ out)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.nv) return;
    if (mask[c] & WPF_MASK_FIXED) {
        out[c] = 0.0f;
        return;
    }
    const int64_t P = (int64_t)g.nx * g.ny;
    const int x = (int)(c % g.nx), y = (int)((c / g.nx) % g.ny), z = (int)(c / P);
    const float* tx = faces;
    const float* ty = faces + g.nv;
    const float* tz = faces + 2 * g.nv;
    float d = tx[c] + ty[c] + tz[c], s = 0.0f;
    if (x < g.nx - 1) s += tx[c] * v[c + 1];
    if (y < g.ny - 1) s += ty[c] * v[c + g.nx];
    if (z < g.nz - 1) s += tz[c] * v[c + P];
    if (x > 0) { d += tx[c - 1]; s += tx[c - 1] * v[c - 1]; }
    if (y > 0) { d += ty[c - g.nx]; s += ty[c - g.nx] * v[c - g.nx]; }
    if (z > 0) { d += tz[c - P]; s += tz[c - P] * v[c - P]; }
    out[c] = d * v[c] - s;
}

// (u, At), (u, Au), (t, At) partials
__global__ void __launch_bounds__(256) k_energy_dots(int64_t n, const float* __restrict__ u,
                                                     const float* __restrict__ au, const float* __restrict__ t,
                                                     const float* __restrict__ at, double* __restrict__ partials)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ui = u[i], ati = at[i];
        v[0] += ui * ati;
        v[1] += ui * (double)au[i];
        v[2] += (double)t[i] * ati;
    }
    block_reduce_store<3>(v, partials, nullptr, nullptr, gridDim.x, blockIdx.x);
}

// loss into d[3]; dL/du = -2a/(bc) At + 2a^2/(b^2 c) Au (a, b, c the dots)
__global__ void k_energy_grad(int64_t n, double* __restrict__ d, const float* __restrict__ at,
                              const float* __restrict__ au, float* __restrict__ gu)
{
    const double a = d[0], b = d[1], c = d[2];
    const bool ok = b > 1e-300 && c > 1e-300;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) d[3] = ok ? 1.0 - a * a / (b * c) : 1.0;
    if (i < n) {
        const double ka = ok ? -2.0 * a / (b * c) : 0.0;
        const double kb = ok ? 2.0 * a * a / (b * b * c) : 0.0;
        gu[i] = (float)(ka * at[i] + kb * au[i]);
    }
}

// weight of coarse index X in the cell-centred linear interpolation of fine
// index x (the 1-D factor of k_upsample_out)
__device__ __forceinline__ float up_weight(int x, int X, int nc)
{
    float c = ((float)x + 0.5f) / (float)kPool - 0.5f;
    c = fminf(fmaxf(c, 0.0f), (float)(nc - 1));
    const int x0 = min((int)c, nc - 1), x1 = min(x0 + 1, nc - 1);
    const float f = c - x0;
    return (x0 == X ? 1.0f - f : 0.0f) + (x1 == X ? f : 0.0f);
}

// gy = U^T gu: the adjoint of k_upsample_out (Dirichlet cells carry gu = 0)
__global__ void k_upsample_adj(Geo fg, Geo cg, const float* __restrict__ gu, float* __restrict__ gy)
{
    const int64_t V = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (V >= cg.nv) return;
    const int X = (int)(V % cg.nx), Y = (int)((V / cg.nx) % cg.ny), Z = (int)(V / ((int64_t)cg.nx * cg.ny));
    // fine cells whose interpolation reads coarse cell X lie within
    // [kPool (X - 1), kPool (X + 2)) on each axis
    const int xa = max(kPool * (X - 1), 0), xb = min(kPool * (X + 2), fg.nx);
    const int ya = max(kPool * (Y - 1), 0), yb = min(kPool * (Y + 2), fg.ny);
    const int za = max(kPool * (Z - 1), 0), zb = min(kPool * (Z + 2), fg.nz);
    float s = 0.0f;
    for (int z = za; z < zb; z++) {
        const float wz = up_weight(z, Z, cg.nz);
        if (wz == 0.0f) continue;
        for (int y = ya; y < yb; y++) {
            const float wy = up_weight(y, Y, cg.ny);
            if (wy == 0.0f) continue;
            const float* row = gu + (int64_t)fg.nx * (y + (int64_t)fg.ny * z);
            float sx = 0.0f;
            for (int x = xa; x < xb; x++) {
                const float wx = up_weight(x, X, cg.nx);
                if (wx != 0.0f) sx += wx * row[x];
            }
            s += wz * wy * sx;
        }
    }
    gy[V] = s;
}

// the solver's FP64 faces as FP32 times a scale (the loss ignores the
// operator's scale; this keeps the values near 1)
__global__ void k_faces32(int64_t n, double scale, const double* __restrict__ tx, const double* __restrict__ ty,
                          const double* __restrict__ tz, float* __restrict__ out)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    out[c] = (float)(tx[c] * scale);
    out[n + c] = (float)(ty[c] * scale);
    out[2 * n + c] = (float)(tz[c] * scale);
}

}  // namespace wpf

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
using namespace wpf;

struct wpf_net {
    int dev = 0;
    hipStream_t stream = nullptr;
    NetLayer L[kNL];
    int64_t np = 0;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    bf16* wt[kNL] = {};
    int adam_t = 0;
    bool packed = false;
    // activations for the current full-resolution grid
    int fx = 0, fy = 0, fz = 0;
    Geo lv[5];
    bf16 *x0 = nullptr, *s1 = nullptr, *p1 = nullptr, *s2 = nullptr, *p2 = nullptr, *s3 = nullptr,
         *p3 = nullptr, *b4 = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr;
    float* y = nullptr;
    // training buffers
    float *fields = nullptr, *target = nullptr, *tpool = nullptr, *gy = nullptr;
    float *g_d1 = nullptr, *gcat1 = nullptr, *g_d2 = nullptr, *g_s1 = nullptr, *gcat2 = nullptr,
          *g_d3 = nullptr, *g_s2 = nullptr, *gcat3 = nullptr, *g_b4 = nullptr, *g_s3 = nullptr,
          *g_p3 = nullptr, *g_p2 = nullptr, *g_p1 = nullptr;
    // backward on MFMA: transposed packed weights per layer, the BF16 copies
    // of gz (padded channels; transposed) and the weight gradient's chunks
    bf16* wT[kNL] = {};
    bf16 *gz16 = nullptr, *gzT = nullptr;
    float* wpart = nullptr;
    double* dots = nullptr;
    double* partials = nullptr;
    // energy-norm loss at full resolution: u = U y, A u, A t, dL/du
    float *eu = nullptr, *eau = nullptr, *eat = nullptr, *egu = nullptr;
    float* efaces = nullptr;       // [3][n] operator of a host-given sample
    uint8_t* emask = nullptr;      // [n]
    // training set on the device: captured pairs and their operators, all
    // on one grid (dnx, dny, dnz)
    int dnx = 0, dny = 0, dnz = 0;
    int dcap_s = 0, dcap_o = 0, dn_s = 0, dn_o = 0;
    float *dfields = nullptr, *dtargets = nullptr, *dfaces = nullptr;
    uint8_t* dmask = nullptr;
    std::vector<int> dop;          // operator of each pair
};

namespace {

void fr(void* p)
{
    if (p) (void)hipFree(p);
}

void free_acts(wpf_net* n)
{
    void* ps[] = {n->x0, n->s1, n->p1, n->s2, n->p2, n->s3, n->p3, n->b4, n->d3, n->d2, n->d1, n->y,
                  n->fields, n->target, n->tpool, n->gy, n->g_d1, n->gcat1, n->g_d2, n->g_s1,
                  n->gcat2, n->g_d3, n->g_s2, n->gcat3, n->g_b4, n->g_s3, n->g_p3, n->g_p2, n->g_p1,
                  n->gz16, n->gzT, n->wpart, n->eu, n->eau, n->eat, n->egu, n->efaces, n->emask};
    for (void* p : ps) fr(p);
    n->gz16 = n->gzT = nullptr;
    n->wpart = nullptr;
    n->eu = n->eau = n->eat = n->egu = n->efaces = nullptr;
    n->emask = nullptr;
    n->x0 = n->s1 = n->p1 = n->s2 = n->p2 = n->s3 = n->p3 = n->b4 = n->d3 = n->d2 = n->d1 = nullptr;
    n->y = n->fields = n->target = n->tpool = n->gy = nullptr;
    n->g_d1 = n->gcat1 = n->g_d2 = n->g_s1 = n->gcat2 = n->g_d3 = n->g_s2 = n->gcat3 = nullptr;
    n->g_b4 = n->g_s3 = n->g_p3 = n->g_p2 = n->g_p1 = nullptr;
    n->fx = n->fy = n->fz = 0;
}

template <typename T>
int alloc(T** p, int64_t count)
{
    WPF_HIP(hipMalloc((void**)p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
    WPF_HIP(hipMemset(*p, 0, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
    return WPF_OK;
}

int ensure_grid(wpf_net* n, int nx, int ny, int nz)
{
    WPF_CHECK(nx % 32 == 0 && ny % 32 == 0 && nz % 32 == 0, WPF_ERR_ARG,
              "network grid %dx%dx%d: every size must be a multiple of 32", nx, ny, nz);
    if (n->fx == nx && n->fy == ny && n->fz == nz) return WPF_OK;
    free_acts(n);
    n->fx = nx; n->fy = ny; n->fz = nz;
    n->lv[0] = geo(nx, ny, nz);
    for (int l = 1; l <= 4; l++) {
        const int f = kPool << (l - 1);
        n->lv[l] = geo(nx / f, ny / f, nz / f);
    }
    const int64_t v1 = n->lv[1].nv, v2 = n->lv[2].nv, v3 = n->lv[3].nv, v4 = n->lv[4].nv;
    int rc;
    if ((rc = alloc(&n->x0, v1 * kCIN0)) || (rc = alloc(&n->s1, v1 * 16)) ||
        (rc = alloc(&n->p1, v2 * 16)) || (rc = alloc(&n->s2, v2 * 32)) ||
        (rc = alloc(&n->p2, v3 * 32)) || (rc = alloc(&n->s3, v3 * 48)) ||
        (rc = alloc(&n->p3, v4 * 48)) || (rc = alloc(&n->b4, v4 * 64)) ||
        (rc = alloc(&n->d3, v3 * 48)) || (rc = alloc(&n->d2, v2 * 32)) ||
        (rc = alloc(&n->d1, v1 * 16)) || (rc = alloc(&n->y, v1)))
        return rc;
    return WPF_OK;
}

// backward shapes of layer li (output level g): channel paddings, the
// transposed weights' K, and the voxel split of the weight gradient
struct BwdShape {
    int copad, cinpad, kpadT, coutpad16;
    int64_t nvpad, vchunk;
    int nchunk;
};
BwdShape bwd_shape(const NetLayer& L, const Geo& g)
{
    BwdShape b;
    b.copad = ((L.cout + 7) / 8) * 8;
    b.cinpad = ((L.cin + 15) / 16) * 16;
    b.kpadT = ((27 * b.copad + 31) / 32) * 32;
    b.coutpad16 = ((L.cout + 15) / 16) * 16;
    b.nvpad = ((g.nv + 31) / 32) * 32;
    const int64_t want = std::min<int64_t>(std::max<int64_t>((g.nv + 1023) / 1024, 1), 512);
    b.vchunk = (((g.nv + want - 1) / want + 31) / 32) * 32;
    b.nchunk = (int)((g.nv + b.vchunk - 1) / b.vchunk);
    return b;
}

int ensure_train(wpf_net* n)
{
    if (n->gy) return WPF_OK;
    const int64_t v0 = n->lv[0].nv, v1 = n->lv[1].nv, v2 = n->lv[2].nv, v3 = n->lv[3].nv,
                  v4 = n->lv[4].nv;
    int rc;
    if (!n->fields && (rc = alloc(&n->fields, 3 * v0))) return rc;
    if ((rc = alloc(&n->target, v0)) ||
        (rc = alloc(&n->tpool, v1)) || (rc = alloc(&n->gy, v1)) || (rc = alloc(&n->g_d1, v1 * 16)) ||
        (rc = alloc(&n->gcat1, v1 * 48)) || (rc = alloc(&n->g_d2, v2 * 32)) ||
        (rc = alloc(&n->g_s1, v1 * 16)) || (rc = alloc(&n->gcat2, v2 * 80)) ||
        (rc = alloc(&n->g_d3, v3 * 48)) || (rc = alloc(&n->g_s2, v2 * 32)) ||
        (rc = alloc(&n->gcat3, v3 * 112)) || (rc = alloc(&n->g_b4, v4 * 64)) ||
        (rc = alloc(&n->g_s3, v3 * 48)) || (rc = alloc(&n->g_p3, v4 * 48)) ||
        (rc = alloc(&n->g_p2, v3 * 32)) || (rc = alloc(&n->g_p1, v2 * 16)))
        return rc;
    // mp also holds the bias gradient's partials (kBiasBlocks x cout <= 256)
    int64_t m16 = 0, mT = 0, mp = (int64_t)kBiasBlocks * 256;
    for (int i = 0; i < kNL; i++) {
        const NetLayer& L = n->L[i];
        const BwdShape b = bwd_shape(L, n->lv[L.level]);
        m16 = std::max(m16, n->lv[L.level].nv * b.copad);
        mT = std::max(mT, (int64_t)b.coutpad16 * b.nvpad);
        mp = std::max(mp, (int64_t)b.nchunk * L.cout * 27 * L.cin);
    }
    if ((rc = alloc(&n->gz16, m16)) || (rc = alloc(&n->gzT, mT)) || (rc = alloc(&n->wpart, mp))) return rc;
    return WPF_OK;
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

int pack_weights(wpf_net* n)
{
    for (int i = 0; i < kNL; i++) {
        const NetLayer& L = n->L[i];
        hipLaunchKernelGGL(k_pack_w, dim3(nblk((int64_t)L.coutpad * L.kpad)), dim3(256), 0, n->stream,
                           (const float*)(n->params + L.woff), L.cout, L.cin, L.coutpad, L.kpad, n->wt[i]);
        WPF_HIP(hipGetLastError());
        const BwdShape b = bwd_shape(L, n->lv[L.level]);
        hipLaunchKernelGGL(k_pack_wT, dim3(nblk((int64_t)b.cinpad * b.kpadT)), dim3(256), 0, n->stream,
                           (const float*)(n->params + L.woff), L.cout, L.cin, b.copad, b.cinpad, b.kpadT, n->wT[i]);
        WPF_HIP(hipGetLastError());
    }
    n->packed = true;
    return WPF_OK;
}

int conv_fwd(wpf_net* n, int li, const Geo& g, Src s, void* out, bool relu, bool f32)
{
    const NetLayer& L = n->L[li];
    const float* bias = n->params + L.woff + (int64_t)L.cout * 27 * L.cin;
    const unsigned blocks = (unsigned)((g.nv + 63) / 64);
    const int nt = L.coutpad / 16;
#define CONVL(NT, RELU, F32)                                                                      \
    hipLaunchKernelGGL((k_conv_fwd<NT, RELU, F32>), dim3(blocks), dim3(256), 0, n->stream, g, s, \
                       (const bf16*)n->wt[li], bias, L.kpad, L.cout, out)
    if (f32) CONVL(1, false, true);
    else if (nt == 1) CONVL(1, true, false);
    else if (nt == 2) CONVL(2, true, false);
    else if (nt == 3) CONVL(3, true, false);
    else CONVL(4, true, false);
#undef CONVL
    (void)relu;
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

// forward from the planar FP32 fields already on the device (3 * nv0)
int forward(wpf_net* n, const float* fields_dev)
{
    if (!n->packed) {
        int rc = pack_weights(n);
        if (rc) return rc;
    }
    const Geo* G = n->lv;
    hipLaunchKernelGGL(k_input_pool, dim3(nblk(G[1].nv * kCIN0)), dim3(256), 0, n->stream, G[0], G[1],
                       fields_dev, n->x0);
    WPF_HIP(hipGetLastError());
    int rc;
    auto pool = [&](const Geo& cg, int C, const bf16* in, bf16* out) {
        hipLaunchKernelGGL(k_pool2, dim3(nblk(cg.nv * C)), dim3(256), 0, n->stream, cg, C, in, out);
        return hipGetLastError() == hipSuccess ? WPF_OK : WPF_ERR_HIP;
    };
    if ((rc = conv_fwd(n, 0, G[1], Src{n->x0, 8, 0, nullptr, 0}, n->s1, true, false))) return rc;
    if ((rc = pool(G[2], 16, n->s1, n->p1))) return rc;
    if ((rc = conv_fwd(n, 1, G[2], Src{n->p1, 16, 0, nullptr, 0}, n->s2, true, false))) return rc;
    if ((rc = pool(G[3], 32, n->s2, n->p2))) return rc;
    if ((rc = conv_fwd(n, 2, G[3], Src{n->p2, 32, 0, nullptr, 0}, n->s3, true, false))) return rc;
    if ((rc = pool(G[4], 48, n->s3, n->p3))) return rc;
    if ((rc = conv_fwd(n, 3, G[4], Src{n->p3, 48, 0, nullptr, 0}, n->b4, true, false))) return rc;
    if ((rc = conv_fwd(n, 4, G[3], Src{n->b4, 64, 1, n->s3, 48}, n->d3, true, false))) return rc;
    if ((rc = conv_fwd(n, 5, G[2], Src{n->d3, 48, 1, n->s2, 32}, n->d2, true, false))) return rc;
    if ((rc = conv_fwd(n, 6, G[1], Src{n->d2, 32, 1, n->s1, 16}, n->d1, true, false))) return rc;
    if ((rc = conv_fwd(n, 7, G[1], Src{n->d1, 16, 0, nullptr, 0}, n->y, false, true))) return rc;
    return WPF_OK;
}

// backward of layer li: gz (already ReLU-masked) -> dW/db, and (optionally) gin
int layer_bwd(wpf_net* n, int li, const Geo& g, Src s, const float* gz, float* gin)
{
    const NetLayer& L = n->L[li];
    float* dw = n->grads + L.woff;
    float* db = dw + (int64_t)L.cout * 27 * L.cin;
    const BwdShape b = bwd_shape(L, g);
    const int ncol = 27 * L.cin;
    // weight gradient: gz transposed to BF16, chunks of voxels, fixed-order sum
    hipLaunchKernelGGL(k_gz_bf16_t, dim3(nblk((int64_t)b.coutpad16 * b.nvpad)), dim3(256), 0, n->stream, g.nv,
                       b.nvpad, L.cout, b.coutpad16, gz, n->gzT);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_conv_bwd_w, dim3((unsigned)((ncol + 63) / 64), (unsigned)(b.coutpad16 / 16), (unsigned)b.nchunk),
                       dim3(256), 0, n->stream, g, s, (const bf16*)n->gzT, b.nvpad, L.cout, b.vchunk, n->wpart);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_chunks, dim3(nblk((int64_t)L.cout * ncol)), dim3(256), 0, n->stream,
                       (int64_t)L.cout * ncol, b.nchunk, (const float*)n->wpart, dw);
    WPF_HIP(hipGetLastError());
    {
        // the bias partials reuse the weight partials' buffer after their sum
        const int64_t bchunk = (g.nv + kBiasBlocks - 1) / kBiasBlocks;
        const int nb = (int)((g.nv + bchunk - 1) / bchunk);
        hipLaunchKernelGGL(k_bias_part, dim3((unsigned)nb), dim3(256), 0, n->stream, g, L.cout, bchunk, gz,
                           n->wpart);
        WPF_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sum_chunks, dim3(nblk(L.cout)), dim3(256), 0, n->stream, (int64_t)L.cout, nb,
                           (const float*)n->wpart, db);
        WPF_HIP(hipGetLastError());
    }
    if (gin) {
        if (!n->packed) {
            int rc = pack_weights(n);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_gz_bf16, dim3(nblk(g.nv * b.copad)), dim3(256), 0, n->stream, g.nv, L.cout, b.copad,
                           gz, n->gz16);
        WPF_HIP(hipGetLastError());
        const unsigned blocks = (unsigned)((g.nv + 63) / 64);
        const int nt = b.cinpad / 16;
#define BWDD(NT)                                                                                          \
    hipLaunchKernelGGL((k_conv_bwd_data<NT>), dim3(blocks), dim3(256), 0, n->stream, g, (const bf16*)n->gz16, \
                       b.copad, (const bf16*)n->wT[li], b.kpadT, L.cin, gin)
        switch (nt) {
        case 1: BWDD(1); break;
        case 2: BWDD(2); break;
        case 3: BWDD(3); break;
        case 4: BWDD(4); break;
        case 5: BWDD(5); break;
        case 6: BWDD(6); break;
        case 7: BWDD(7); break;
        default:
            set_error("backward: %d input channels", L.cin);
            return WPF_ERR_ARG;
        }
#undef BWDD
        WPF_HIP(hipGetLastError());
    }
    return WPF_OK;
}

// gradient of every parameter from dL/dy in n->gy (level 1), through the
// activations of the last forward pass
int backprop(wpf_net* n)
{
    const Geo* G = n->lv;
    int rc;
    auto relu = [&](int64_t cnt, float* g, const bf16* act) {
        hipLaunchKernelGGL(k_relu_mask, dim3(nblk(cnt)), dim3(256), 0, n->stream, cnt, g, act);
    };
    auto upcat = [&](const Geo& fine, const Geo& coarse, int ca, int cb, const float* gin, float* ga,
                     float* gb) {
        hipLaunchKernelGGL(k_upcat_bwd_a, dim3(nblk(coarse.nv * ca)), dim3(256), 0, n->stream, coarse, ca,
                           cb, gin, ga);
        hipLaunchKernelGGL(k_upcat_bwd_b, dim3(nblk(fine.nv * cb)), dim3(256), 0, n->stream, fine, ca, cb,
                           gin, gb);
    };
    auto poolb = [&](const Geo& fine, int C, const float* gc, float* gf) {
        hipLaunchKernelGGL(k_pool_bwd_add, dim3(nblk(fine.nv * C)), dim3(256), 0, n->stream, fine, C, gc,
                           gf);
    };
    // out layer (linear): gy -> g_d1
    if ((rc = layer_bwd(n, 7, G[1], Src{n->d1, 16, 0, nullptr, 0}, n->gy, n->g_d1))) return rc;
    relu(G[1].nv * 16, n->g_d1, n->d1);
    if ((rc = layer_bwd(n, 6, G[1], Src{n->d2, 32, 1, n->s1, 16}, n->g_d1, n->gcat1))) return rc;
    upcat(G[1], G[2], 32, 16, n->gcat1, n->g_d2, n->g_s1);
    relu(G[2].nv * 32, n->g_d2, n->d2);
    if ((rc = layer_bwd(n, 5, G[2], Src{n->d3, 48, 1, n->s2, 32}, n->g_d2, n->gcat2))) return rc;
    upcat(G[2], G[3], 48, 32, n->gcat2, n->g_d3, n->g_s2);
    relu(G[3].nv * 48, n->g_d3, n->d3);
    if ((rc = layer_bwd(n, 4, G[3], Src{n->b4, 64, 1, n->s3, 48}, n->g_d3, n->gcat3))) return rc;
    upcat(G[3], G[4], 64, 48, n->gcat3, n->g_b4, n->g_s3);
    relu(G[4].nv * 64, n->g_b4, n->b4);
    if ((rc = layer_bwd(n, 3, G[4], Src{n->p3, 48, 0, nullptr, 0}, n->g_b4, n->g_p3))) return rc;
    poolb(G[3], 48, n->g_p3, n->g_s3);
    relu(G[3].nv * 48, n->g_s3, n->s3);
    if ((rc = layer_bwd(n, 2, G[3], Src{n->p2, 32, 0, nullptr, 0}, n->g_s3, n->g_p2))) return rc;
    poolb(G[2], 32, n->g_p2, n->g_s2);
    relu(G[2].nv * 32, n->g_s2, n->s2);
    if ((rc = layer_bwd(n, 1, G[2], Src{n->p1, 16, 0, nullptr, 0}, n->g_s2, n->g_p1))) return rc;
    poolb(G[1], 16, n->g_p1, n->g_s1);
    relu(G[1].nv * 16, n->g_s1, n->s1);
    if ((rc = layer_bwd(n, 0, G[1], Src{n->x0, 8, 0, nullptr, 0}, n->g_s1, nullptr))) return rc;
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

int read_loss(wpf_net* n, double* loss)
{
    double d[4];
    WPF_HIP(hipMemcpyAsync(d, n->dots, sizeof(d), hipMemcpyDeviceToHost, n->stream));
    WPF_HIP(hipStreamSynchronize(n->stream));
    if (loss) *loss = d[3];
    return WPF_OK;
}

// cosine loss and full gradient for the sample in n->fields / n->target
int loss_and_grad(wpf_net* n, double* loss)
{
    const Geo* G = n->lv;
    int rc = forward(n, n->fields);
    if (rc) return rc;
    hipLaunchKernelGGL(k_target_pool, dim3(nblk(G[1].nv)), dim3(256), 0, n->stream, G[0], G[1],
                       (const float*)n->target, n->tpool);
    WPF_HIP(hipGetLastError());
    const int nb = 64;
    hipLaunchKernelGGL(k_loss_dots, dim3(nb), dim3(256), 0, n->stream, G[1].nv, (const float*)n->y,
                       (const float*)n->tpool, n->partials);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, n->stream, (const double*)n->partials, nb, 3,
                       n->dots);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_loss_grad, dim3(nblk(G[1].nv)), dim3(256), 0, n->stream, G[1].nv, n->dots,
                       (const float*)n->y, (const float*)n->tpool, n->gy);
    WPF_HIP(hipGetLastError());
    if ((rc = backprop(n))) return rc;
    return read_loss(n, loss);
}

int ensure_energy(wpf_net* n)
{
    if (n->eu) return WPF_OK;
    const int64_t v0 = n->lv[0].nv;
    int rc;
    if ((rc = alloc(&n->eu, v0)) || (rc = alloc(&n->eau, v0)) || (rc = alloc(&n->eat, v0)) ||
        (rc = alloc(&n->egu, v0)) || (rc = alloc(&n->efaces, 3 * v0)) || (rc = alloc(&n->emask, v0)))
        return rc;
    return WPF_OK;
}

// energy-norm loss and full gradient for one pair on the device: network
// input fields [3][n], error t [n], operator faces [3][n] and mask [n]
int energy_loss_and_grad(wpf_net* n, const float* fields, const float* t, const float* faces,
                         const uint8_t* mask, double* loss)
{
    const Geo* G = n->lv;
    int rc;
    if ((rc = ensure_train(n)) || (rc = ensure_energy(n))) return rc;
    if ((rc = forward(n, fields))) return rc;
    const int64_t v0 = G[0].nv;
    hipLaunchKernelGGL(k_upsample_out<float>, dim3(nblk(v0)), dim3(256), 0, n->stream, G[0], G[1],
                       (const float*)n->y, mask, n->eu);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_energy_apply, dim3(nblk(v0)), dim3(256), 0, n->stream, G[0], faces, mask,
                       (const float*)n->eu, n->eau);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_energy_apply, dim3(nblk(v0)), dim3(256), 0, n->stream, G[0], faces, mask, t, n->eat);
    WPF_HIP(hipGetLastError());
    const int nb = 64;
    hipLaunchKernelGGL(k_energy_dots, dim3(nb), dim3(256), 0, n->stream, v0, (const float*)n->eu,
                       (const float*)n->eau, t, (const float*)n->eat, n->partials);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, n->stream, (const double*)n->partials, nb, 3, n->dots);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_energy_grad, dim3(nblk(v0)), dim3(256), 0, n->stream, v0, n->dots,
                       (const float*)n->eat, (const float*)n->eau, n->egu);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_upsample_adj, dim3(nblk(G[1].nv)), dim3(256), 0, n->stream, G[0], G[1],
                       (const float*)n->egu, n->gy);
    WPF_HIP(hipGetLastError());
    if ((rc = backprop(n))) return rc;
    return read_loss(n, loss);
}

void adam_step(wpf_net* n, float lr)
{
    n->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n->np)), dim3(256), 0, n->stream, n->np, n->params,
                       (const float*)n->grads, n->m, n->v, lr, b1, b2, 1e-8f, c1, c2);
    n->packed = false;
}

int upload_sample(wpf_net* n, int nx, int ny, int nz, const float* fields, const float* target)
{
    int rc = ensure_grid(n, nx, ny, nz);
    if (rc) return rc;
    if ((rc = ensure_train(n))) return rc;
    const size_t nv = (size_t)nx * ny * nz;
    WPF_HIP(hipMemcpyAsync(n->fields, fields, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, n->stream));
    if (target)
        WPF_HIP(hipMemcpyAsync(n->target, target, sizeof(float) * nv, hipMemcpyHostToDevice, n->stream));
    return WPF_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// solver-side use (called from k_solver.hip)
// ---------------------------------------------------------------------------
namespace wpf {

// Net input from the solver state: planar [log10 k scaled, p / p_scale,
// invD r / rms] at full resolution, then the forward pass; the correction d
// (0 on fixed cells) is written to out at full resolution.
// the same from the scaled FP32 residual of the mixed solver: invD32 r32 =
// invD64 r64, so the third channel is identical up to rounding
__global__ void k_solver_input32(int64_t n, double kwh, double rms_inv, double p_scale,
                                 const double* __restrict__ phi, const double* __restrict__ kr,
                                 const double* __restrict__ p, const float* __restrict__ r,
                                 const float* __restrict__ invD, float* __restrict__ fields)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double f = 0.5 * (1.0 + phi[c]);
    const double k0 = kwh * f + kr[c] * (1.0 - f);
    fields[c] = (float)((log10(fmax(k0, 1e-300)) + 13.0) * 0.5);
    fields[n + c] = (float)(p[c] / p_scale);
    fields[2 * n + c] = (float)((double)r[c] * (double)invD[c] * rms_inv);
}

__global__ void k_solver_input(int64_t n, double kwh, double rms_inv, double p_scale,
                               const double* __restrict__ phi, const double* __restrict__ kr,
                               const double* __restrict__ p, const double* __restrict__ r,
                               const double* __restrict__ invD, float* __restrict__ fields)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double f = 0.5 * (1.0 + phi[c]);
    const double k0 = kwh * f + kr[c] * (1.0 - f);
    fields[c] = (float)((log10(fmax(k0, 1e-300)) + 13.0) * 0.5);
    fields[n + c] = (float)(p[c] / p_scale);
    fields[2 * n + c] = (float)(r[c] * invD[c] * rms_inv);
}

int solver_input(wpf_ctx* c, const double* r64, double rnorm, float* fields)
{
    const Grid& g = c->g;
    const double rms_inv = rnorm > 0 ? std::sqrt((double)g.n) / rnorm : 0.0;
    hipLaunchKernelGGL(k_solver_input, dim3(nblk(g.n)), dim3(256), 0, c->stream, g.n, c->prm.permeab_hole,
                       rms_inv, 1.0e5, (const double*)c->phi[c->cur], (const double*)c->krock,
                       (const double*)c->p.p, r64, (const double*)c->invD.p, fields);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

template <typename R>
int net_correction(wpf_ctx* c, const double* r64, double rnorm, R* out, float* fields_copy)
{
    wpf_net* n = c->net;
    WPF_CHECK(n, WPF_ERR_STATE, "no network attached");
    const Grid& g = c->g;
    int rc = ensure_grid(n, g.nx, g.ny, g.nz);
    if (rc) return rc;
    if (!n->fields) {
        if ((rc = alloc(&n->fields, 3 * n->lv[0].nv))) return rc;
    }
    hipStream_t saved = n->stream;
    n->stream = c->stream;
    const double rms_inv = rnorm > 0 ? std::sqrt((double)g.n) / rnorm : 0.0;
    hipLaunchKernelGGL(k_solver_input, dim3(nblk(g.n)), dim3(256), 0, c->stream, g.n, c->prm.permeab_hole,
                       rms_inv, 1.0e5, (const double*)c->phi[c->cur], (const double*)c->krock,
                       (const double*)c->p.p, r64, (const double*)c->invD.p, n->fields);
    WPF_HIP(hipGetLastError());
    if (fields_copy)
        WPF_HIP(hipMemcpyAsync(fields_copy, n->fields, sizeof(float) * 3 * (size_t)g.n, hipMemcpyDeviceToDevice,
                               c->stream));
    rc = forward(n, n->fields);
    if (!rc) {
        hipLaunchKernelGGL(k_upsample_out<R>, dim3(nblk(g.n)), dim3(256), 0, c->stream, n->lv[0], n->lv[1],
                           (const float*)n->y, (const uint8_t*)c->mask, out);
        if (hipGetLastError() != hipSuccess) rc = WPF_ERR_HIP;
    }
    n->stream = saved;
    return rc;
}
int net_correction32(wpf_ctx* c, const float* r32, const float* invD32, double rnorm, float* out)
{
    wpf_net* n = c->net;
    WPF_CHECK(n, WPF_ERR_STATE, "no network attached");
    const Grid& g = c->g;
    int rc = ensure_grid(n, g.nx, g.ny, g.nz);
    if (rc) return rc;
    if (!n->fields) {
        if ((rc = alloc(&n->fields, 3 * n->lv[0].nv))) return rc;
    }
    hipStream_t saved = n->stream;
    n->stream = c->stream;
    const double rms_inv = rnorm > 0 ? std::sqrt((double)g.n) / rnorm : 0.0;
    hipLaunchKernelGGL(k_solver_input32, dim3(nblk(g.n)), dim3(256), 0, c->stream, g.n, c->prm.permeab_hole,
                       rms_inv, 1.0e5, (const double*)c->phi[c->cur], (const double*)c->krock,
                       (const double*)c->p.p, r32, invD32, n->fields);
    WPF_HIP(hipGetLastError());
    rc = forward(n, n->fields);
    if (!rc) {
        hipLaunchKernelGGL(k_upsample_out<float>, dim3(nblk(g.n)), dim3(256), 0, c->stream, n->lv[0], n->lv[1],
                           (const float*)n->y, (const uint8_t*)c->mask, out);
        if (hipGetLastError() != hipSuccess) rc = WPF_ERR_HIP;
    }
    n->stream = saved;
    return rc;
}

template int net_correction<double>(wpf_ctx*, const double*, double, double*, float*);
template int net_correction<float>(wpf_ctx*, const double*, double, float*, float*);

void free_learned(wpf_ctx* c) { c->net = nullptr; }  // the network is owned by its creator

}  // namespace wpf

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int64_t wpf_net_num_params(void) { return count_params(); }

int wpf_net_create(int32_t seed, wpf_net** out)
{
    WPF_CHECK(out, WPF_ERR_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("wpf_net_create: no GPU visible");
        return WPF_ERR_NODEV;
    }
    wpf_net* n = new wpf_net();
    WPF_HIP(hipGetDevice(&n->dev));
    WPF_HIP(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    layer_table(n->L);
    n->np = count_params();
    int rc;
    if ((rc = alloc(&n->params, n->np)) || (rc = alloc(&n->grads, n->np)) || (rc = alloc(&n->m, n->np)) ||
        (rc = alloc(&n->v, n->np)) || (rc = alloc(&n->dots, 8)) || (rc = alloc(&n->partials, 3 * 64)))
        return rc;
    for (int i = 0; i < kNL; i++)
        if ((rc = alloc(&n->wt[i], (int64_t)n->L[i].coutpad * n->L[i].kpad))) return rc;
    for (int i = 0; i < kNL; i++) {
        const BwdShape b = bwd_shape(n->L[i], geo(1, 1, 1));
        if ((rc = alloc(&n->wT[i], (int64_t)b.cinpad * b.kpadT))) return rc;
    }
    // He initialisation from a fixed-seed generator (host), zero biases
    std::vector<float> h(n->np, 0.0f);
    uint64_t s = 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)seed;
    auto nextu = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int i = 0; i < kNL; i++) {
        const NetLayer& L = n->L[i];
        const double std = std::sqrt(2.0 / (27.0 * L.cin));
        for (int64_t k = 0; k < (int64_t)L.cout * 27 * L.cin; k++) {
            const double u1 = nextu(), u2 = nextu();
            h[L.woff + k] = (float)(std * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    }
    WPF_HIP(hipMemcpy(n->params, h.data(), sizeof(float) * n->np, hipMemcpyHostToDevice));
    *out = n;
    return WPF_OK;
}

int wpf_net_destroy(wpf_net* n)
{
    if (!n) return WPF_OK;
    (void)hipStreamSynchronize(n->stream);
    free_acts(n);
    fr(n->params); fr(n->grads); fr(n->m); fr(n->v); fr(n->dots); fr(n->partials);
    fr(n->eu); fr(n->eau); fr(n->eat); fr(n->egu); fr(n->efaces); fr(n->emask);
    fr(n->dfields); fr(n->dtargets); fr(n->dfaces); fr(n->dmask);
    for (int i = 0; i < kNL; i++) {
        fr(n->wt[i]);
        fr(n->wT[i]);
    }
    (void)hipStreamDestroy(n->stream);
    delete n;
    return WPF_OK;
}

int wpf_net_get_params(wpf_net* n, float* host, int64_t count)
{
    WPF_CHECK(n && host && count == n->np, WPF_ERR_ARG, "parameter count %lld != %lld",
              (long long)count, (long long)(n ? n->np : 0));
    WPF_HIP(hipMemcpyAsync(host, n->params, sizeof(float) * count, hipMemcpyDeviceToHost, n->stream));
    WPF_HIP(hipStreamSynchronize(n->stream));
    return WPF_OK;
}

int wpf_net_set_params(wpf_net* n, const float* host, int64_t count)
{
    WPF_CHECK(n && host && count == n->np, WPF_ERR_ARG, "parameter count %lld != %lld",
              (long long)count, (long long)(n ? n->np : 0));
    WPF_HIP(hipMemcpyAsync(n->params, host, sizeof(float) * count, hipMemcpyHostToDevice, n->stream));
    WPF_HIP(hipStreamSynchronize(n->stream));
    n->packed = false;
    return WPF_OK;
}

int wpf_net_forward(wpf_net* n, int32_t nx, int32_t ny, int32_t nz, const float* fields, float* out)
{
    WPF_CHECK(n && fields && out, WPF_ERR_ARG, "null argument");
    int rc = upload_sample(n, nx, ny, nz, fields, nullptr);
    if (rc) return rc;
    if ((rc = forward(n, n->fields))) return rc;
    WPF_HIP(hipMemcpyAsync(out, n->y, sizeof(float) * n->lv[1].nv, hipMemcpyDeviceToHost, n->stream));
    WPF_HIP(hipStreamSynchronize(n->stream));
    return WPF_OK;
}

int wpf_net_gradient(wpf_net* n, int32_t nx, int32_t ny, int32_t nz, const float* fields,
                     const float* target, double* loss, float* grad)
{
    WPF_CHECK(n && fields && target && grad, WPF_ERR_ARG, "null argument");
    int rc = upload_sample(n, nx, ny, nz, fields, target);
    if (rc) return rc;
    if ((rc = loss_and_grad(n, loss))) return rc;
    WPF_HIP(hipMemcpyAsync(grad, n->grads, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->stream));
    WPF_HIP(hipStreamSynchronize(n->stream));
    return WPF_OK;
}

int wpf_net_train(wpf_net* n, int32_t nx, int32_t ny, int32_t nz, int32_t nsamples,
                  const float* fields, const float* targets, int32_t epochs, float lr, double* losses)
{
    WPF_CHECK(n && fields && targets && nsamples > 0 && epochs > 0, WPF_ERR_ARG, "bad training arguments");
    const size_t nv = (size_t)nx * ny * nz;
    // a solve that started converged leaves a zero correction: no direction
    // to learn from, so such samples are left out
    std::vector<int> use;
    for (int s = 0; s < nsamples; s++) {
        const float* t = targets + nv * s;
        bool any = false;
        for (size_t i = 0; i < nv && !any; i++) any = t[i] != 0.0f;
        if (any) use.push_back(s);
    }
    WPF_CHECK(!use.empty(), WPF_ERR_ARG, "every training target is zero");
    for (int e = 0; e < epochs; e++) {
        double sum = 0.0;
        for (int s : use) {
            int rc = upload_sample(n, nx, ny, nz, fields + 3 * nv * s, targets + nv * s);
            if (rc) return rc;
            double L = 0.0;
            if ((rc = loss_and_grad(n, &L))) return rc;
            sum += L;
            n->adam_t++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
            hipLaunchKernelGGL(k_adam, dim3(nblk(n->np)), dim3(256), 0, n->stream, n->np, n->params,
                               (const float*)n->grads, n->m, n->v, lr, b1, b2, 1e-8f, c1, c2);
            WPF_HIP(hipGetLastError());
            n->packed = false;
        }
        if (losses) losses[e] = sum / (double)use.size();
    }
    WPF_HIP(hipStreamSynchronize(n->stream));
    return WPF_OK;
}

int wpf_net_gradient_energy(wpf_net* n, int32_t nx, int32_t ny, int32_t nz, const float* fields,
                            const float* target, const float* faces, const uint8_t* fixed, double* loss,
                            float* grad)
{
    WPF_CHECK(n && fields && target && faces && fixed && grad, WPF_ERR_ARG, "null argument");
    int rc = upload_sample(n, nx, ny, nz, fields, target);
    if (rc) return rc;
    if ((rc = ensure_energy(n))) return rc;
    const size_t nv = (size_t)nx * ny * nz;
    std::vector<uint8_t> m(nv);
    for (size_t i = 0; i < nv; i++) m[i] = fixed[i] ? WPF_MASK_FIXED : 0;
    WPF_HIP(hipMemcpyAsync(n->efaces, faces, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, n->stream));
    WPF_HIP(hipMemcpyAsync(n->emask, m.data(), nv, hipMemcpyHostToDevice, n->stream));
    if ((rc = energy_loss_and_grad(n, n->fields, n->target, n->efaces, n->emask, loss))) return rc;
    WPF_HIP(hipMemcpyAsync(grad, n->grads, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->stream));
    WPF_HIP(hipStreamSynchronize(n->stream));
    return WPF_OK;
}

int wpf_net_dataset_reserve(wpf_net* n, int32_t nx, int32_t ny, int32_t nz, int32_t max_pairs, int32_t max_ops)
{
    WPF_CHECK(n && max_pairs > 0 && max_ops > 0, WPF_ERR_ARG, "bad dataset size");
    WPF_CHECK(nx % 32 == 0 && ny % 32 == 0 && nz % 32 == 0, WPF_ERR_ARG,
              "network grid %dx%dx%d: every size must be a multiple of 32", nx, ny, nz);
    WPF_HIP(hipSetDevice(n->dev));
    WPF_HIP(hipStreamSynchronize(n->stream));
    fr(n->dfields); fr(n->dtargets); fr(n->dfaces); fr(n->dmask);
    n->dfields = n->dtargets = n->dfaces = nullptr;
    n->dmask = nullptr;
    n->dop.clear();
    n->dn_s = n->dn_o = 0;
    const int64_t nv = (int64_t)nx * ny * nz;
    int rc;
    if ((rc = alloc(&n->dfields, 3 * nv * max_pairs)) || (rc = alloc(&n->dtargets, nv * max_pairs)) ||
        (rc = alloc(&n->dfaces, 3 * nv * max_ops)) || (rc = alloc(&n->dmask, nv * max_ops)))
        return rc;
    n->dnx = nx; n->dny = ny; n->dnz = nz;
    n->dcap_s = max_pairs;
    n->dcap_o = max_ops;
    return WPF_OK;
}

int wpf_net_dataset_add(wpf_net* n, wpf_ctx* c, int32_t stride, int32_t* npairs)
{
    WPF_CHECK(n && c && stride >= 1, WPF_ERR_ARG, "bad argument");
    WPF_CHECK(n->dcap_s > 0, WPF_ERR_STATE, "no dataset reserved");
    WPF_CHECK(c->g.nx == n->dnx && c->g.ny == n->dny && c->g.nz == n->dnz, WPF_ERR_ARG,
              "context grid %dx%dx%d is not the dataset's %dx%dx%d", c->g.nx, c->g.ny, c->g.nz, n->dnx,
              n->dny, n->dnz);
    // the faces still hold the operator of the capturing solve: wpf_step
    // marks them stale for the next step, and only a permeability pass
    // rewrites them
    WPF_CHECK(c->fcap_n > 0, WPF_ERR_STATE, "the context's last solve captured no stage points");
    int add = 0;
    for (int j = 0; j < c->fcap_n; j += stride) add++;
    WPF_CHECK(n->dn_s + add <= n->dcap_s && n->dn_o < n->dcap_o, WPF_ERR_ARG,
              "dataset full (%d + %d pairs of %d, %d operators of %d)", n->dn_s, add, n->dcap_s, n->dn_o,
              n->dcap_o);
    WPF_HIP(hipSetDevice(n->dev));
    const int64_t nv = c->g.n;
    const int op = n->dn_o++;
    hipLaunchKernelGGL(k_faces32, dim3(nblk(nv)), dim3(256), 0, c->stream, nv, c->t_scale, (const double*)c->tx.p,
                       (const double*)c->ty.p, (const double*)c->tz.p, n->dfaces + 3 * nv * op);
    WPF_HIP(hipGetLastError());
    WPF_HIP(hipMemcpyAsync(n->dmask + nv * op, c->mask, nv, hipMemcpyDeviceToDevice, c->stream));
    for (int j = 0; j < c->fcap_n; j += stride) {
        const int s = n->dn_s++;
        WPF_HIP(hipMemcpyAsync(n->dfields + 3 * nv * s, c->fcap_fields[j], sizeof(float) * 3 * nv,
                               hipMemcpyDeviceToDevice, c->stream));
        WPF_HIP(hipMemcpyAsync(n->dtargets + nv * s, c->fcap_target[j], sizeof(float) * nv, hipMemcpyDeviceToDevice,
                               c->stream));
        n->dop.push_back(op);
    }
    WPF_HIP(hipStreamSynchronize(c->stream));
    if (npairs) *npairs = n->dn_s;
    return WPF_OK;
}

int wpf_net_train_energy(wpf_net* n, int32_t epochs, float lr, int32_t seed, double* losses)
{
    WPF_CHECK(n && epochs > 0, WPF_ERR_ARG, "bad training arguments");
    WPF_CHECK(n->dn_s > 0, WPF_ERR_STATE, "the dataset is empty");
    WPF_HIP(hipSetDevice(n->dev));
    int rc = ensure_grid(n, n->dnx, n->dny, n->dnz);
    if (rc) return rc;
    const int64_t nv = (int64_t)n->dnx * n->dny * n->dnz;
    std::vector<int> order(n->dn_s);
    for (int i = 0; i < n->dn_s; i++) order[i] = i;
    uint64_t st = 0x243F6A8885A308D3ull ^ (uint64_t)(uint32_t)seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    for (int e = 0; e < epochs; e++) {
        for (int i = n->dn_s - 1; i > 0; i--) std::swap(order[i], order[next() % (uint64_t)(i + 1)]);
        double sum = 0.0;
        for (int s : order) {
            const int op = n->dop[s];
            double L = 0.0;
            if ((rc = energy_loss_and_grad(n, n->dfields + 3 * nv * s, n->dtargets + nv * s, n->dfaces + 3 * nv * op,
                                           n->dmask + nv * op, &L)))
                return rc;
            sum += L;
            adam_step(n, lr);
            WPF_HIP(hipGetLastError());
        }
        if (losses) losses[e] = sum / (double)n->dn_s;
    }
    WPF_HIP(hipStreamSynchronize(n->stream));
    return WPF_OK;
}

int wpf_attach_net(wpf_ctx* c, wpf_net* n)
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    c->net = n;
    return WPF_OK;
}

}  // extern "C"
