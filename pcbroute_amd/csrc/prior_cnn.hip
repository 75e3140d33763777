// prior_cnn.hip - the learned corridor prior: conv3x3 layers on gfx950's BF16 matrix cores (MFMA 16x16x32, FP32
// accumulation), forward, backward and Adam, behind include/pcbprior.h.
//
// Layout. Activations are pixel-major BF16 with 32 channel slots (NHWC, padded with zeros). A layer's weights are
// kept as FP32 masters in the padded layout w[co][k] with k = (ky * 3 + kx) * 32 + ci, and as two BF16 copies: the
// forward one in that layout and a flipped one wf[ci][k'] = w[co][(8 - d) * 32 + ci] with k' = d * 32 + co, which
// makes the gradient with respect to a layer's input the same implicit GEMM as a forward pass.
//
// Matrix-core fragments (v_mfma_f32_16x16x32_bf16, D = A B + C with A 16 x 32, B 32 x 16): lane l holds A[l % 16][8
// (l / 16) + j] and B[8 (l / 16) + j][l % 16] for j = 0..7 (16 bytes each), and gets D[4 (l / 16) + j][l % 16] for
// j = 0..3. One 3x3 tap of 32 channels is exactly one K step of 32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/pcbprior.h"

namespace {

thread_local std::string g_err;

bool fail(const char* what, hipError_t e = hipSuccess) {
    g_err = what;
    if (e != hipSuccess) { g_err += ": "; g_err += hipGetErrorString(e); }
    return false;
}

#define HIPOK(x)                                   \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return fail(#x, e_); \
    } while (0)

constexpr int C = PCBPRIOR_MAX_CH;  // channel slots of an activation
constexpr int KK = 9 * C;           // the K of a layer's implicit GEMM
constexpr int PX = 64;              // pixels of a row a workgroup takes (4 wavefronts x 16)

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// BF16 from FP32: round to nearest, ties to even (the IEEE rule); NaN stays a quiet NaN
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return uint16_t((u >> 16) | 0x40);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = uint32_t(h) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// 16 bytes at base + off if ok, else zeros: the load is always issued (at base when not ok, which must be valid), so
// a staging loop is straight-line code with all its loads in flight together
__device__ inline uint4 load16_or_0(const uint16_t* base, size_t off, bool ok) {
    const uint4 v = *reinterpret_cast<const uint4*>(base + (ok ? off : 0));
    return ok ? v : make_uint4(0, 0, 0, 0);
}

// D += A B on the matrix cores, the fragments as 16-byte vectors of 8 BF16 values
__device__ inline v4f mfma32(uint4 a, uint4 b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

// ---------------------------------------------------------------- kernels

__global__ void k_to_bf16(const float* x, uint16_t* out, size_t npix, int cin) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npix * C) return;
    const size_t p = i / C;
    const int c = int(i % C);
    out[i] = c < cin ? f2bf(x[p * cin + c]) : uint16_t(0);
}

// y = conv3x3(x, wt) + bias as an implicit GEMM with the weights as A (rows: output channels, K: tap x input
// channel) and the input as B (K: one tap's 32 input channels of a pixel, columns: pixels), so that a lane's result
// is 4 consecutive output channels of one pixel and goes straight to memory. A tile is 64 pixels of CR (8)
// consecutive rows, one row per wavefront, all 32 output channels; tiles are numbered column block first, then row
// block, then image. A block is persistent: it stages the layer's weights once, then takes tiles blockIdx.x,
// blockIdx.x + gridDim.x, ... (the blocks running at once cover neighbouring tiles, which measured faster than each
// block walking its own strip); while a tile's products run, the next tile's (CR + 2) x 66-pixel input patch is
// already on its way into registers, and goes to LDS between two barriers once every wavefront is done with the
// current one. Per tap a wavefront reads 2 A and 4 B fragments (16 bytes each) from LDS for 8 matrix-core products
// into 4 x 2 accumulator tiles (64 pixels x 32 channels). Each lane then finishes its 4 channels of 8 pixel-tiles:
// bias, ReLU, the mask (zero where the mask activation is zero: the derivative of the layer before's ReLU), BF16
// stores of 8 bytes per pixel, or FP32 logits of the first cout channels. With `valid` (rows, columns per image) the
// outputs outside image n's valid extent are 0, so an image padded into a larger batch is computed exactly as alone
// at its valid size (whose conv reads zeros past its edges).
constexpr int CR = 8;         // rows per tile
constexpr int NT = 64 * CR;   // threads per block
constexpr int XSP = 40;       // LDS pitch of a staged input pixel (BF16): 80 bytes, conflict-free 16-byte reads
constexpr int WSP = KK + 8;   // LDS pitch of a staged weight row (BF16): 592 bytes, conflict-free 16-byte reads
constexpr int XS_N = (CR + 2) * (PX + 2) * XSP, WS_N = C * WSP;  // BF16 elements
constexpr int CONV_BLOCKS_PER_CU = 2;  // what the block's LDS allows
__global__ __launch_bounds__(NT) void k_conv(const uint16_t* x, const uint16_t* wt, const float* bias, uint16_t* ybf,
                                              float* yf, const uint16_t* mask, int relu, int cout, int N, int H, int W,
                                              const int32_t* valid) {
    __shared__ __align__(16) uint16_t sm[XS_N + WS_N];
    uint16_t* XS = sm;
    uint16_t* WS = sm + XS_N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = W / PX, tiles_y = (H + CR - 1) / CR, ntiles = tiles_x * tiles_y * N;
    constexpr int NW = C * KK / 8, NX = (CR + 2) * (PX + 2) * (C / 8);
    constexpr int WIT = (NW + NT - 1) / NT, XIT = (NX + NT - 1) / NT;
    uint4 xv[XIT];
    // a tile's input patch into registers (zero outside the image), then into LDS
    auto load_patch = [&](int t) {
        const int x0 = (t % tiles_x) * PX, y0 = (t / tiles_x % tiles_y) * CR;
        const size_t img = size_t(t / (tiles_x * tiles_y)) * H * W;
#pragma unroll
        for (int it = 0; it < XIT; ++it) {
            const int e = tid + it * NT;
            const int r = e / ((PX + 2) * (C / 8)), rem = e % ((PX + 2) * (C / 8)), q = rem / (C / 8), c8 = rem % (C / 8);
            const int py = y0 - 1 + r, px = x0 - 1 + q;
            xv[it] = make_uint4(0, 0, 0, 0);
            if (e < NX && py >= 0 && py < H && px >= 0 && px < W)
                xv[it] = *reinterpret_cast<const uint4*>(x + (img + size_t(py) * W + px) * C + c8 * 8);
        }
    };
    auto store_patch = [&]() {
#pragma unroll
        for (int it = 0; it < XIT; ++it) {
            const int e = tid + it * NT;
            if (e < NX) {
                const int r = e / ((PX + 2) * (C / 8)), rem = e % ((PX + 2) * (C / 8)), q = rem / (C / 8), c8 = rem % (C / 8);
                *reinterpret_cast<uint4*>(XS + (r * (PX + 2) + q) * XSP + c8 * 8) = xv[it];
            }
        }
    };
    {  // the weights and the first tile: every load issued before the first LDS store
        uint4 wv[WIT];
#pragma unroll
        for (int it = 0; it < WIT; ++it) {
            const int e = tid + it * NT;
            wv[it] = make_uint4(0, 0, 0, 0);
            if (e < NW) wv[it] = *reinterpret_cast<const uint4*>(wt + size_t(e / (KK / 8)) * KK + (e % (KK / 8)) * 8);
        }
        if (int(blockIdx.x) < ntiles) load_patch(blockIdx.x);
#pragma unroll
        for (int it = 0; it < WIT; ++it) {
            const int e = tid + it * NT;
            if (e < NW) *reinterpret_cast<uint4*>(WS + (e / (KK / 8)) * WSP + (e % (KK / 8)) * 8) = wv[it];
        }
        store_patch();
    }
    __syncthreads();
    const int pr = lane & 15, kq = lane >> 4;
    // lane (pr, kq) finishes channels 16 t + 4 kq + j of its pixels
    float bv[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) bv[t][j] = bias ? bias[16 * t + 4 * kq + j] : 0.f;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int tn = t + gridDim.x;
        if (tn < ntiles) load_patch(tn);  // in flight during this tile's products
        v4f acc[4][2];
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m][0] = acc[m][1] = v4f{0, 0, 0, 0};
#pragma unroll 3
        for (int d = 0; d < 9; ++d) {  // tap (dy, dx): K = its 32 input channels, lane quarter kq takes 8 of them
            const int dy = d / 3, dx = d % 3;
            const uint4 a0 = *reinterpret_cast<const uint4*>(WS + pr * WSP + d * C + kq * 8);
            const uint4 a1 = *reinterpret_cast<const uint4*>(WS + (16 + pr) * WSP + d * C + kq * 8);
            const uint16_t* xr = XS + ((wave + dy) * (PX + 2) + pr + dx) * XSP + kq * 8;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint4 b = *reinterpret_cast<const uint4*>(xr + m * 16 * XSP);
                acc[m][0] = mfma32(a0, b, acc[m][0]);
                acc[m][1] = mfma32(a1, b, acc[m][1]);
            }
        }
        const int n = t / (tiles_x * tiles_y), x0 = (t % tiles_x) * PX, y = (t / tiles_x % tiles_y) * CR + wave;
        if (y < H) {
            const size_t img = size_t(n) * H * W;
            const bool row_in = !valid || y < valid[2 * n];
            const int cols = valid ? valid[2 * n + 1] : W;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int px = x0 + 16 * m + pr;  // W is a multiple of 64: every pixel is inside
                const size_t p = img + size_t(y) * W + px;
                const bool inside = row_in && px < cols;
#pragma unroll
                for (int tt = 0; tt < 2; ++tt) {
                    const int c0 = 16 * tt + 4 * kq;
                    float v[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        v[j] = acc[m][tt][j] + bv[tt][j];
                        if (relu) v[j] = fmaxf(v[j], 0.f);
                        if (!inside) v[j] = 0.f;
                    }
                    if (yf) {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (c0 + j < cout) yf[p * cout + c0 + j] = v[j];
                    }
                    if (ybf) {
                        if (mask) {
                            const uint2 mk = *reinterpret_cast<const uint2*>(mask + p * C + c0);
#pragma unroll
                            for (int j = 0; j < 4; ++j)  // zero where the ReLU was off
                                if ((((j < 2 ? mk.x : mk.y) >> (16 * (j & 1))) & 0x7FFF) == 0) v[j] = 0.f;
                        }
                        *reinterpret_cast<uint2*>(ybf + p * C + c0) =
                            make_uint2(uint32_t(f2bf(v[0])) | uint32_t(f2bf(v[1])) << 16,
                                       uint32_t(f2bf(v[2])) | uint32_t(f2bf(v[3])) << 16);
                    }
                }
            }
        }
        __syncthreads();  // every wavefront is done with this tile's patch
        if (tn < ntiles) store_patch();
        __syncthreads();
    }
}

// dW[co][k] += sum over the block's pixels of dy[p][co] * x_im2col[p][k]; db[co] += sum of dy[p][co]. The reduction
// runs over pixels, so a lane's 8 operand values are 8 consecutive pixels of one channel: the output gradient and the
// input rows are staged transposed in LDS (a channel's pixels contiguous), and a fragment is one aligned 16-byte read
// (two, and a shift, where the 3x3 tap's column offset misaligns it). A work item is WR rows of 64 columns of one
// image (WR + 2 input rows staged); items are numbered row chunk first, then column block, then image. A block is
// persistent and holds two groups of 4 wavefronts with a staging area each; it takes a contiguous run of items
// (mostly one column strip downwards), each group one half of it in order (so the input rows an item shares with the
// one before are fresh in L2), with its next item's loads in flight during the current item's products.
// Wavefront w of a group accumulates the k tiles w, w + 4, ... against both co tiles in registers over all its
// items; at the end the second group's sums are added to the first's in LDS and the first adds them into dW with
// FP32 atomics, once per block.
constexpr int WR = 4;
constexpr int DTP = WR * PX + 8;         // LDS pitch of a channel of the staged output gradient (BF16): 16 * odd bytes
constexpr int XTW = PX + 16;             // a staged input row: column xb + c at c + 8 (columns xb - 1 .. xb + 64 used)
constexpr int XTP = (WR + 2) * XTW + 8;  // LDS pitch of a channel of the staged input (BF16): 16 * odd bytes
constexpr int WG_NT = 512;               // threads per block: two groups of 256
constexpr int WGRAD_BLOCKS_PER_CU = 1;   // what the block's LDS allows
// row[pos .. pos + 7] as a fragment; sh = pos & 7 is 0, 1 or 7 (the three tap columns) and uniform across the
// wavefront: one aligned 16-byte read, or two and a shift by one BF16 value up or down
__device__ inline uint32_t pair_at(uint32_t lo, uint32_t hi) { return __builtin_amdgcn_alignbyte(hi, lo, 2); }
__device__ inline uint4 ld_px8(const uint16_t* row, int pos, int sh) {
    const uint4 lo = *reinterpret_cast<const uint4*>(row + (pos & ~7));
    if (sh == 0) return lo;
    const uint4 hi = *reinterpret_cast<const uint4*>(row + (pos & ~7) + 8);
    if (sh == 1) return make_uint4(pair_at(lo.x, lo.y), pair_at(lo.y, lo.z), pair_at(lo.z, lo.w), pair_at(lo.w, hi.x));
    return make_uint4(pair_at(lo.w, hi.x), pair_at(hi.x, hi.y), pair_at(hi.y, hi.z), pair_at(hi.z, hi.w));  // sh == 7
}
// 4 pixels x 8 channels (four 16-byte loads, pixel-major) into 8 runs of 4 pixels of one channel (8-byte stores)
__device__ inline void put4(uint16_t* dst, int pitch, const uint4 (&v)[4]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint32_t e[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) e[j] = ((&v[j].x)[k >> 1] >> (16 * (k & 1))) & 0xFFFF;
        *reinterpret_cast<uint2*>(dst + k * pitch) = make_uint2(e[0] | (e[1] << 16), e[2] | (e[3] << 16));
    }
}
// One item's products for wavefront `wave` of a group: its k tiles are kt = wave + 4 u, whose tap is 2 u + wave / 2
// and channel block (wave % 2) * 16, so with WP = wave / 2 known at compile time every tap's row and column offset,
// and so every fragment's shift, is a constant, and the loop is straight-line code. All WR rows are taken: rows past
// the image were staged as zeros and add nothing.
template <int WP>
__device__ inline void wgrad_item(const uint16_t* DT, const uint16_t* XT, int wave, int pr, int kq, v4f (&acc)[2][5]) {
    constexpr int NU = WP == 0 ? 5 : 4;  // k tiles of waves 0, 1: up to 16, 17; of waves 2, 3: up to 14, 15
    const int ci = (wave & 1) * 16 + pr;
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
        for (int ps = 0; ps < PX / 32; ++ps) {
            // A = dy^T (row co, column pixel), B = x_im2col (row pixel, column k); K = 32 pixels of row r
            const int p0 = ps * 32 + kq * 8;
            const uint4 a0 = *reinterpret_cast<const uint4*>(DT + pr * DTP + r * PX + p0);
            const uint4 a1 = *reinterpret_cast<const uint4*>(DT + (16 + pr) * DTP + r * PX + p0);
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                // pixel p's input column is p + dxx - 1, staged at p + dxx + 7
                const int d = 2 * u + WP, dyy = d / 3, dxx = d % 3;
                const uint4 bb = ld_px8(XT + ci * XTP + (r + dyy) * XTW, p0 + dxx + 7, (dxx + 7) & 7);
                acc[0][u] = mfma32(a0, bb, acc[0][u]);
                acc[1][u] = mfma32(a1, bb, acc[1][u]);
            }
        }
}
__global__ __launch_bounds__(WG_NT) void k_wgrad(const uint16_t* x, const uint16_t* dy, float* dw, float* db, int N,
                                                  int H, int W) {
    __shared__ __align__(16) uint16_t DTs[2][C * DTP];
    __shared__ __align__(16) uint16_t XTs[2][C * XTP];
    const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), tid = threadIdx.x & 255, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    uint16_t* DT = DTs[grp];
    uint16_t* XT = XTs[grp];
    const int pr = lane & 15, kq = lane >> 4;
    const int cols_x = W / PX, chunks = (H + WR - 1) / WR, nitems = cols_x * chunks * N;
    const int per = (nitems + gridDim.x - 1) / gridDim.x, it_beg = int(blockIdx.x) * per;
    const int it_end = min(nitems, it_beg + per);
    constexpr int KT = KK / 16;  // 18 k tiles: wavefront w takes w, w + 4, ..., at most 5 of them
    constexpr int NDT = WR * (PX / 4) * (C / 8), DIT = (NDT + 255) / 256;
    constexpr int NXT = (WR + 2) * ((PX + 8) / 4) * (C / 8), XIT = (NXT + 255) / 256;
    uint4 vd[DIT][4], vx[XIT][4];
    // an item's output gradient rows c0 .. c0 + WR - 1 (zero past the image) and input rows c0 - 1 .. c0 + WR in
    // groups of 4 columns xb - 4 + 4 g .. xb - 1 + 4 g, g = 0..17 (zero outside the image)
    auto load_item = [&](int it0) {
        const int xb = (it0 / chunks % cols_x) * PX, c0 = (it0 % chunks) * WR;
        const size_t img = size_t(it0 / (cols_x * chunks)) * H * W;
#pragma unroll
        for (int it = 0; it < DIT; ++it) {
            const int e = tid + it * 256;
            const int r = e / ((PX / 4) * (C / 8)), rem = e % ((PX / 4) * (C / 8)), c8 = rem / (PX / 4), g = rem % (PX / 4);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                vd[it][j] = load16_or_0(dy, (img + size_t(c0 + r) * W + xb + 4 * g + j) * C + c8 * 8, e < NDT && c0 + r < H);
        }
#pragma unroll
        for (int it = 0; it < XIT; ++it) {
            const int e = tid + it * 256;
            const int r = e / (((PX + 8) / 4) * (C / 8)), rem = e % (((PX + 8) / 4) * (C / 8)), c8 = rem / ((PX + 8) / 4),
                      g = rem % ((PX + 8) / 4);
            const int py = c0 - 1 + r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int px = xb - 4 + 4 * g + j;
                vx[it][j] = load16_or_0(x, (img + size_t(py) * W + px) * C + c8 * 8,
                                        e < NXT && py >= 0 && py < H && px >= 0 && px < W);
            }
        }
    };
    auto store_item = [&]() {
#pragma unroll
        for (int it = 0; it < DIT; ++it) {
            const int e = tid + it * 256;
            const int r = e / ((PX / 4) * (C / 8)), rem = e % ((PX / 4) * (C / 8)), c8 = rem / (PX / 4), g = rem % (PX / 4);
            if (e < NDT) put4(DT + (c8 * 8) * DTP + r * PX + 4 * g, DTP, vd[it]);
        }
#pragma unroll
        for (int it = 0; it < XIT; ++it) {
            const int e = tid + it * 256;
            const int r = e / (((PX + 8) / 4) * (C / 8)), rem = e % (((PX + 8) / 4) * (C / 8)), c8 = rem / ((PX + 8) / 4),
                      g = rem % ((PX + 8) / 4);
            if (e < NXT) put4(XT + (c8 * 8) * XTP + r * XTW + 4 + 4 * g, XTP, vx[it]);  // column xb + c at c + 8
        }
    };
    v4f acc[2][5];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 5; ++u) acc[t][u] = v4f{0, 0, 0, 0};
    float bsum = 0.f;
    // group grp takes the grp-th half of the block's run, in order (consecutive row chunks: the input rows an item
    // shares with the one before are fresh in L2); both groups run the same number of rounds (barriers)
    const int rounds = (max(0, it_end - it_beg) + 1) / 2, g_beg = it_beg + grp * rounds;
    const int g_end = min(it_end, g_beg + rounds);
    if (g_beg < g_end) {
        load_item(g_beg);
        store_item();
    }
    __syncthreads();
    for (int k = 0; k < rounds; ++k) {
        const int it0 = g_beg + k, itn = it0 + 1;
        if (itn < g_end) load_item(itn);  // in flight during this item's products
        if (it0 < g_end) {
            const int rows = min(WR, H - (it0 % chunks) * WR);
            {  // the bias gradient: thread t sums 32 pixels of channel t % 32 (the partial sums meet at the end)
                const int ch = tid & (C - 1), q0 = (tid / C) * 32;
#pragma unroll
                for (int q = q0; q < q0 + 32; q += 8)
                    if (q < rows * PX) {
                        const uint4 d8 = *reinterpret_cast<const uint4*>(DT + ch * DTP + q);
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const uint32_t u = (&d8.x)[j];
                            bsum += bf2f(uint16_t(u & 0xFFFF)) + bf2f(uint16_t(u >> 16));
                        }
                    }
            }
            if (wave < 2)
                wgrad_item<0>(DT, XT, wave, pr, kq, acc);
            else
                wgrad_item<1>(DT, XT, wave, pr, kq, acc);
        }
        __syncthreads();  // every wavefront is done with its group's item
        if (itn < g_end) store_item();
        __syncthreads();
    }
    // the second group's sums into the first's (through the free staging space), then one atomic per sum
    float* red = reinterpret_cast<float*>(&XTs[0][0]);  // 2 x 5 x 4 floats per lane of a group: 40 KB
    if (grp == 1) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int u = 0; u < 5; ++u)
#pragma unroll
                for (int j = 0; j < 4; ++j) red[((t * 5 + u) * 4 + j) * 256 + tid] = acc[t][u][j];
    }
    float* bred = reinterpret_cast<float*>(&DTs[0][0]);
    bred[threadIdx.x] = bsum;
    __syncthreads();
    if (grp == 0) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int u = 0; u < 5; ++u) {
                const int kt = wave + 4 * u;
                if (kt < KT)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        atomicAdd(&dw[size_t(16 * t + kq * 4 + j) * KK + kt * 16 + pr],
                                  acc[t][u][j] + red[((t * 5 + u) * 4 + j) * 256 + tid]);
            }
    }
    if (threadIdx.x < C) {
        float t = 0.f;
#pragma unroll
        for (int k = 0; k < WG_NT / C; ++k) t += bred[threadIdx.x + C * k];
        atomicAdd(&db[threadIdx.x], t);
    }
}

// binary cross-entropy of the logits over the masked pixels; dz = (sigmoid(z) - t) m / count into the BF16 slots
// (0 in the padded ones). A thread per pixel; each block adds its partial loss into one of LOSS_SLOTS sums (atomics
// on one address serialize), which the host adds up.
constexpr int LOSS_SLOTS = 64;
__global__ __launch_bounds__(256) void k_loss(const float* z, const float* t, const float* m, uint16_t* dz, float* loss,
                                              size_t npix, int cout, const unsigned* cnt) {
    __shared__ float part[4];
    const float inv = *cnt ? 1.f / (float(*cnt) * float(cout)) : 0.f;
    float l = 0.f;
    if (const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < npix) {
        const bool on = m[p] != 0.f;
        uint32_t w[C / 2];
        for (int c = 0; c < C; c += 2) {
            float g[2] = {0.f, 0.f};
            for (int k = 0; k < 2; ++k)
                if (on && c + k < cout) {
                    const float zz = z[p * cout + c + k], tt = t[p * cout + c + k];
                    g[k] = (1.f / (1.f + expf(-zz)) - tt) * inv;
                    l += (fmaxf(zz, 0.f) - zz * tt + log1pf(expf(-fabsf(zz)))) * inv;
                }
            w[c / 2] = uint32_t(f2bf(g[0])) | (uint32_t(f2bf(g[1])) << 16);
        }
        uint4* o = reinterpret_cast<uint4*>(dz + p * C);
        for (int q = 0; q < C / 8; ++q) o[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&loss[blockIdx.x % LOSS_SLOTS], part[0] + part[1] + part[2] + part[3]);
}

// masked pixels of a batch (the loss's mean runs over them): a grid-stride loop, one atomic per block (a few
// hundred blocks: atomics on one address serialize)
__global__ __launch_bounds__(256) void k_count(const float* m, size_t npix, unsigned* cnt) {
    __shared__ unsigned part[4];
    unsigned c = 0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < npix; i += size_t(gridDim.x) * blockDim.x)
        c += m[i] != 0.f ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(cnt, part[0] + part[1] + part[2] + part[3]);
}

// a batch from the resident training set: sample idx[b] into slot b (inputs in BF16 slots, targets, mask)
__global__ void k_gather(const uint16_t* dsx, const float* dst, const float* dsm, const int32_t* idx, size_t spix,
                         int cout, int batch, uint16_t* x, float* t, float* m) {
    // blockIdx.y: the batch slot; each sample is one contiguous run per array, copied in 16-byte vectors (a pixel's
    // C BF16 slots are 64 bytes) and grid-stride loops
    const int b = blockIdx.y;
    const size_t s = size_t(idx[b]);
    const size_t i0 = size_t(blockIdx.x) * blockDim.x + threadIdx.x, step = size_t(gridDim.x) * blockDim.x;
    const uint4* xs = reinterpret_cast<const uint4*>(dsx + s * spix * C);
    uint4* xd = reinterpret_cast<uint4*>(x + size_t(b) * spix * C);
    for (size_t i = i0; i < spix * C / 8; i += step) xd[i] = xs[i];
    const float* ts = dst + s * spix * cout;
    float* td = t + size_t(b) * spix * cout;
    for (size_t i = i0; i < spix * cout; i += step) td[i] = ts[i];
    for (size_t i = i0; i < spix; i += step) m[size_t(b) * spix + i] = dsm[s * spix + i];
}

__global__ void k_sigmoid(const float* z, float* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 1.f / (1.f + expf(-z[i]));
}

// Adam on the padded FP32 masters; refreshes the forward and flipped BF16 copies
__global__ void k_adam(float* w, const float* g, float* mo, float* ve, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gg = g[i];
    const float m = b1 * mo[i] + (1.f - b1) * gg, v = b2 * ve[i] + (1.f - b2) * gg * gg;
    mo[i] = m;
    ve[i] = v;
    w[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// the BF16 copies of every layer's weights (blockIdx.y: the layer)
__global__ void k_weights_bf16(const float* w, uint16_t* wt, uint16_t* wf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // over C * KK
    if (i >= C * KK) return;
    const size_t L = size_t(blockIdx.y) * C * KK;
    const int co = i / KK, k = i % KK, d = k / C, ci = k % C;
    const uint16_t v = f2bf(w[L + i]);
    wt[L + i] = v;
    wf[L + size_t(ci) * KK + (8 - d) * C + co] = v;  // wf[ci][d' * 32 + co] = w[co][(8 - d') * 32 + ci]
}

// a backward pass's accumulators to zero in one launch: the weight and bias gradients, the loss and the pixel count
__global__ void k_zero_grads(float* g, size_t ng, float* gb, size_t ngb, float* loss, unsigned* cnt) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < ng) g[i] = 0.f;
    if (i < ngb) gb[i] = 0.f;
    if (i < LOSS_SLOTS) loss[i] = 0.f;
    if (i == 0) *cnt = 0u;
}

__global__ void k_mfma_check(const uint16_t* a, const uint16_t* b, float* d) {  // a 16 x 32, b 32 x 16
    const int lane = threadIdx.x, pr = lane & 15, kq = lane >> 4;
    uint32_t fa[4], fb[4];
    for (int j = 0; j < 4; ++j) {  // A[pr][8 kq + 2 j + h] and B[8 kq + 2 j + h][pr], h = 0, 1
        fa[j] = uint32_t(a[pr * 32 + kq * 8 + 2 * j]) | uint32_t(a[pr * 32 + kq * 8 + 2 * j + 1]) << 16;
        fb[j] = uint32_t(b[(kq * 8 + 2 * j) * 16 + pr]) | uint32_t(b[(kq * 8 + 2 * j + 1) * 16 + pr]) << 16;
    }
    const v4f r = mfma32(make_uint4(fa[0], fa[1], fa[2], fa[3]), make_uint4(fb[0], fb[1], fb[2], fb[3]), v4f{0, 0, 0, 0});
    for (int j = 0; j < 4; ++j) d[(kq * 4 + j) * 16 + pr] = r[j];  // D[4 kq + j][pr]
}

unsigned g1(size_t n) { return unsigned((n + 255) / 256); }

template <class T>
struct DB {  // a device array grown on demand
    T* p = nullptr;
    size_t cap = 0;
    bool need(size_t n) {
        if (n == 0) n = 1;
        if (cap >= n) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) { p = nullptr; return fail("hipMalloc"); }
        cap = n;
        return true;
    }
    ~DB() { if (p) (void)hipFree(p); }
};

}  // namespace

struct pcbprior {
    int cin = 0, width = 0, depth = 0, cout = 0;
    int cus = 0;  // compute units of the device: the persistent kernels' grids
    std::vector<int> lin, lout;  // real channels of each layer
    hipStream_t s = nullptr;
    DB<float> w, g, mo, ve, bias, gb, mb, vb;  // per layer: C * KK weights, C biases (padded)
    DB<uint16_t> wt, wf;
    std::vector<DB<uint16_t>> act;  // depth activations: act[0] the input, act[L] the input of layer L
    DB<uint16_t> dya, dyb;
    DB<float> logits, tgt, msk, xin, loss;  // loss: LOSS_SLOTS partial sums
    float loss_h[LOSS_SLOTS] = {};          // their host copy
    DB<unsigned> cnt;
    DB<int32_t> valid;  // per image (rows, columns) of a padded batch
    int step = 0;
    size_t npix = 0;
    // the resident training set: inputs in BF16 slots, targets and masks, all in HBM
    DB<uint16_t> dsx;
    DB<float> dst, dsm;
    DB<int32_t> didx;
    int ds_n = 0, ds_h = 0, ds_w = 0;
    // kernel accounting (events on the network's stream, read at the end of each call)
    bool acct = false;
    std::vector<hipEvent_t> evpool;
    struct Pending { hipEvent_t a, b; int kind; double flop; };
    std::vector<Pending> pend;
    pcbprior_stats st{};
};

namespace {

// every copy goes through the network's own stream (non-blocking: the null stream does not order it)
bool h2d(pcbprior* p, void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, p->s) == hipSuccess || fail("hipMemcpyAsync H2D");
}
bool d2h(pcbprior* p, void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, p->s) == hipSuccess || fail("hipMemcpyAsync D2H");
}
bool sync(pcbprior* p) { return hipStreamSynchronize(p->s) == hipSuccess || fail("hipStreamSynchronize"); }
// the loss's partial sums to the host (ordered on the stream: read them after a sync), and their total
bool loss_to_host(pcbprior* p) { return d2h(p, p->loss_h, p->loss.p, LOSS_SLOTS * 4); }
float loss_sum(const pcbprior* p) {
    float t = 0.f;
    for (int k = 0; k < LOSS_SLOTS; ++k) t += p->loss_h[k];
    return t;
}

bool ensure(pcbprior* p, int n, int h, int w) {
    if (n < 1 || h < 1 || w < 1 || w % PX) return fail("n, h >= 1 and w a positive multiple of 64");
    const size_t np = size_t(n) * h * w;
    p->npix = np;
    if (int(p->act.size()) != p->depth) p->act.resize(p->depth);
    for (auto& a : p->act)
        if (!a.need(np * C)) return false;
    return p->dya.need(np * C) && p->dyb.need(np * C) && p->logits.need(np * p->cout) && p->tgt.need(np * p->cout) &&
           p->msk.need(np) && p->xin.need(np * p->cin) && p->loss.need(LOSS_SLOTS);
}

bool refresh_bf16(pcbprior* p) {
    hipLaunchKernelGGL(k_weights_bf16, dim3(g1(C * KK), unsigned(p->depth)), dim3(256), 0, p->s, p->w.p, p->wt.p,
                       p->wf.p);
    HIPOK(hipGetLastError());
    return true;
}

// kernel accounting: an event pair around a launch of kind 0 (conv: forward or input gradient) or 1 (weight
// gradient) with its algorithmic FLOPs (real channels: 2 x pixels x cout x 9 x cin)
hipEvent_t ev_take(pcbprior* p) {
    if (!p->evpool.empty()) { hipEvent_t e = p->evpool.back(); p->evpool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}
struct Timed {
    pcbprior* p;
    hipEvent_t a = nullptr;
    int kind;
    double flop;
    Timed(pcbprior* p_, int kind_, double flop_) : p(p_), kind(kind_), flop(flop_) {
        if (p->acct) { a = ev_take(p); (void)hipEventRecord(a, p->s); }
    }
    ~Timed() {
        if (!p->acct) return;
        hipEvent_t b = ev_take(p);
        (void)hipEventRecord(b, p->s);
        p->pend.push_back({a, b, kind, flop});
    }
};
void collect(pcbprior* p) {  // after a stream sync
    for (auto& e : p->pend) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e.a, e.b);
        if (e.kind == 0) { ++p->st.conv_launches; p->st.conv_ms += ms; p->st.conv_flop += e.flop; }
        else { ++p->st.wgrad_launches; p->st.wgrad_ms += ms; p->st.wgrad_flop += e.flop; }
        p->evpool.push_back(e.a);
        p->evpool.push_back(e.b);
    }
    p->pend.clear();
}

// persistent grids: at most as many blocks as the device holds at once (k_wgrad: each with an equal run of items)
unsigned run_blocks(long work, long cap) {
    const long per = (work + cap - 1) / cap;
    return unsigned(std::max(1L, (work + per - 1) / per));
}
unsigned conv_blocks(const pcbprior* p, int n, int h, int w) {
    return unsigned(std::max(1L, std::min(long(w / PX) * ((h + CR - 1) / CR) * n, long(CONV_BLOCKS_PER_CU) * p->cus)));
}
unsigned wgrad_blocks(const pcbprior* p, int n, int h, int w) {
    return run_blocks(long(w / PX) * ((h + WR - 1) / WR) * n, long(WGRAD_BLOCKS_PER_CU) * p->cus);
}

// the layers over act[0] (already on the device)
bool forward_dev(pcbprior* p, int n, int h, int w, const int32_t* valid = nullptr) {
    const dim3 grid(conv_blocks(p, n, h, w));
    const double np = double(n) * h * w;
    for (int L = 0; L < p->depth; ++L) {
        const bool last = L == p->depth - 1;
        Timed t(p, 0, 2.0 * np * p->lout[L] * 9 * p->lin[L]);
        hipLaunchKernelGGL(k_conv, grid, dim3(NT), 0, p->s, p->act[L].p, p->wt.p + size_t(L) * C * KK,
                           p->bias.p + L * C, last ? nullptr : p->act[L + 1].p, last ? p->logits.p : nullptr,
                           static_cast<const uint16_t*>(nullptr), last ? 0 : 1, p->cout, n, h, w, valid);
    }
    HIPOK(hipGetLastError());
    return true;
}

bool forward(pcbprior* p, const float* x, int n, int h, int w, const int32_t* valid = nullptr) {
    if (!ensure(p, n, h, w)) return false;
    const size_t np = p->npix;
    HIPOK(hipMemcpyAsync(p->xin.p, x, np * p->cin * 4, hipMemcpyHostToDevice, p->s));
    hipLaunchKernelGGL(k_to_bf16, dim3(g1(np * C)), dim3(256), 0, p->s, p->xin.p, p->act[0].p, np, p->cin);
    if (valid) {
        if (!p->valid.need(size_t(2) * n) || !h2d(p, p->valid.p, valid, size_t(2) * n * 4)) return false;
        return forward_dev(p, n, h, w, p->valid.p);
    }
    return forward_dev(p, n, h, w);
}

// loss and gradients; targets and mask already on the device
bool backward_dev(pcbprior* p, int n, int h, int w) {
    const size_t np = p->npix;
    if (!p->cnt.need(1)) return false;
    const size_t ng = size_t(p->depth) * C * KK, ngb = size_t(p->depth) * C;
    hipLaunchKernelGGL(k_zero_grads, dim3(g1(ng)), dim3(256), 0, p->s, p->g.p, ng, p->gb.p, ngb, p->loss.p, p->cnt.p);
    hipLaunchKernelGGL(k_count, dim3(unsigned(std::min<size_t>(g1(np), size_t(p->cus)))), dim3(256), 0, p->s, p->msk.p,
                       np, p->cnt.p);  // one atomic per block
    hipLaunchKernelGGL(k_loss, dim3(g1(np)), dim3(256), 0, p->s, p->logits.p, p->tgt.p,
                       p->msk.p, p->dya.p, p->loss.p, np, p->cout, static_cast<const unsigned*>(p->cnt.p));
    const dim3 grid(conv_blocks(p, n, h, w)), wgrid(wgrad_blocks(p, n, h, w));
    uint16_t* dy = p->dya.p;
    uint16_t* dx = p->dyb.p;
    for (int L = p->depth - 1; L >= 0; --L) {
        const double fl = 2.0 * double(np) * p->lout[L] * 9 * p->lin[L];
        {
            Timed t(p, 1, fl);
            hipLaunchKernelGGL(k_wgrad, wgrid, dim3(WG_NT), 0, p->s, p->act[L].p, dy, p->g.p + size_t(L) * C * KK,
                               p->gb.p + L * C, n, h, w);
        }
        if (L > 0) {
            // the gradient at layer L's input: the flipped-weight conv of dy, zero where that input's ReLU was off
            Timed t(p, 0, fl);
            hipLaunchKernelGGL(k_conv, grid, dim3(NT), 0, p->s, dy, p->wf.p + size_t(L) * C * KK,
                               static_cast<const float*>(nullptr), dx, static_cast<float*>(nullptr), p->act[L].p, 0, 0,
                               n, h, w, static_cast<const int32_t*>(nullptr));
            std::swap(dy, dx);
        }
    }
    HIPOK(hipGetLastError());
    return true;
}

bool backward(pcbprior* p, const float* target, const float* mask, int n, int h, int w, float* loss_out) {
    const size_t np = p->npix;
    HIPOK(hipMemcpyAsync(p->tgt.p, target, np * p->cout * 4, hipMemcpyHostToDevice, p->s));
    HIPOK(hipMemcpyAsync(p->msk.p, mask, np * 4, hipMemcpyHostToDevice, p->s));
    if (!backward_dev(p, n, h, w)) return false;
    if (loss_out) {
        if (!loss_to_host(p) || !sync(p)) return false;
        *loss_out = loss_sum(p);
        collect(p);
    }
    return true;
}

// one Adam step on the gradients just computed (FP32 masters; BF16 copies refreshed)
bool adam(pcbprior* p, float lr) {
    ++p->step;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, float(p->step)), c2 = 1.f - std::pow(b2, float(p->step));
    const size_t nw = size_t(p->depth) * C * KK, nb = size_t(p->depth) * C;
    hipLaunchKernelGGL(k_adam, dim3(g1(nw)), dim3(256), 0, p->s, p->w.p, p->g.p, p->mo.p, p->ve.p, nw, lr, b1, b2,
                       eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(g1(nb)), dim3(256), 0, p->s, p->bias.p, p->gb.p, p->mb.p, p->vb.p, nb, lr, b1, b2,
                       eps, c1, c2);
    return refresh_bf16(p);
}

// batch idx[0..batch) of the resident set into act[0], tgt and msk
bool gather(pcbprior* p, const int32_t* idx, int batch) {
    if (!p->ds_n) return fail("no resident training set");
    if (batch < 1) return fail("empty batch");
    for (int b = 0; b < batch; ++b)
        if (idx[b] < 0 || idx[b] >= p->ds_n) return fail("sample index out of range");
    if (!ensure(p, batch, p->ds_h, p->ds_w) || !p->didx.need(size_t(batch)) || !h2d(p, p->didx.p, idx, size_t(batch) * 4))
        return false;
    const size_t spix = size_t(p->ds_h) * p->ds_w;
    const unsigned gx = unsigned(std::min<size_t>(512, (spix * C / 8 + 255) / 256));
    hipLaunchKernelGGL(k_gather, dim3(gx, unsigned(batch)), dim3(256), 0, p->s, p->dsx.p, p->dst.p, p->dsm.p,
                       p->didx.p, spix, p->cout, batch, p->act[0].p, p->tgt.p, p->msk.p);
    HIPOK(hipGetLastError());
    return true;
}

// compact parameter order <-> padded device layout
void pack_layout(const pcbprior* p, std::vector<size_t>& wi, std::vector<size_t>& bi) {
    wi.clear();
    bi.clear();
    for (int L = 0; L < p->depth; ++L) {
        for (int co = 0; co < p->lout[L]; ++co)
            for (int d = 0; d < 9; ++d)
                for (int ci = 0; ci < p->lin[L]; ++ci) wi.push_back(size_t(L) * C * KK + size_t(co) * KK + d * C + ci);
        for (int co = 0; co < p->lout[L]; ++co) bi.push_back(size_t(L) * C + co);
    }
}

}  // namespace

extern "C" {

pcbprior* pcbprior_create(int in_ch, int width, int depth, int out_ch, uint64_t seed) {
    g_err.clear();
    if (in_ch < 1 || in_ch > C || width < 1 || width > C || depth < 1 || out_ch < 1 || out_ch > C) {
        g_err = "channels must be in [1, 32]";
        return nullptr;
    }
    auto* p = new pcbprior;
    p->cin = in_ch; p->width = width; p->depth = depth; p->cout = out_ch;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        p->cus < 1) {
        g_err = "no device to run on";
        delete p;
        return nullptr;
    }
    for (int L = 0; L < depth; ++L) {
        p->lin.push_back(L == 0 ? in_ch : width);
        p->lout.push_back(L == depth - 1 ? out_ch : width);
    }
    const size_t nw = size_t(depth) * C * KK, nb = size_t(depth) * C;
    if (hipStreamCreateWithFlags(&p->s, hipStreamNonBlocking) != hipSuccess || !p->w.need(nw) || !p->g.need(nw) ||
        !p->mo.need(nw) || !p->ve.need(nw) || !p->bias.need(nb) || !p->gb.need(nb) || !p->mb.need(nb) ||
        !p->vb.need(nb) || !p->wt.need(nw) || !p->wf.need(nw)) {
        pcbprior_destroy(p);
        return nullptr;
    }
    // He-uniform weights from the seed (host), biases 0; padded slots stay 0
    std::vector<float> w(nw, 0.f), b(nb, 0.f), z(nw, 0.f), zb(nb, 0.f);
    std::mt19937_64 rng(seed);
    for (int L = 0; L < depth; ++L) {
        const float lim = std::sqrt(6.f / float(9 * p->lin[L]));
        std::uniform_real_distribution<float> u(-lim, lim);
        for (int co = 0; co < p->lout[L]; ++co)
            for (int d = 0; d < 9; ++d)
                for (int ci = 0; ci < p->lin[L]; ++ci) w[size_t(L) * C * KK + size_t(co) * KK + d * C + ci] = u(rng);
    }
    if (!h2d(p, p->w.p, w.data(), nw * 4) || !h2d(p, p->bias.p, b.data(), nb * 4) || !h2d(p, p->mo.p, z.data(), nw * 4) ||
        !h2d(p, p->ve.p, z.data(), nw * 4) || !h2d(p, p->mb.p, zb.data(), nb * 4) || !h2d(p, p->vb.p, zb.data(), nb * 4) ||
        !refresh_bf16(p) || !sync(p)) {
        g_err = "initialising the weights";
        pcbprior_destroy(p);
        return nullptr;
    }
    return p;
}

void pcbprior_destroy(pcbprior* p) {
    if (!p) return;
    if (p->s) (void)hipStreamSynchronize(p->s);
    collect(p);
    for (auto e : p->evpool) (void)hipEventDestroy(e);
    if (p->s) (void)hipStreamDestroy(p->s);
    delete p;
}

size_t pcbprior_num_params(const pcbprior* p) {
    if (!p) return 0;
    size_t n = 0;
    for (int L = 0; L < p->depth; ++L) n += size_t(p->lout[L]) * 9 * p->lin[L] + p->lout[L];
    return n;
}

int pcbprior_get_params(const pcbprior* p, float* out) {
    if (!p || !out) return -1;
    const size_t nw = size_t(p->depth) * C * KK, nb = size_t(p->depth) * C;
    std::vector<float> w(nw), b(nb);
    pcbprior* q = const_cast<pcbprior*>(p);
    if (!d2h(q, w.data(), p->w.p, nw * 4) || !d2h(q, b.data(), p->bias.p, nb * 4) || !sync(q)) return -1;
    std::vector<size_t> wi, bi;
    pack_layout(p, wi, bi);
    size_t o = 0, a = 0, c = 0;
    for (int L = 0; L < p->depth; ++L) {
        const size_t nwl = size_t(p->lout[L]) * 9 * p->lin[L];
        for (size_t k = 0; k < nwl; ++k) out[o++] = w[wi[a++]];
        for (int k = 0; k < p->lout[L]; ++k) out[o++] = b[bi[c++]];
    }
    return 0;
}

int pcbprior_set_params(pcbprior* p, const float* in) {
    if (!p || !in) return -1;
    const size_t nw = size_t(p->depth) * C * KK, nb = size_t(p->depth) * C;
    std::vector<float> w(nw, 0.f), b(nb, 0.f);
    std::vector<size_t> wi, bi;
    pack_layout(p, wi, bi);
    size_t o = 0, a = 0, c = 0;
    for (int L = 0; L < p->depth; ++L) {
        const size_t nwl = size_t(p->lout[L]) * 9 * p->lin[L];
        for (size_t k = 0; k < nwl; ++k) w[wi[a++]] = in[o++];
        for (int k = 0; k < p->lout[L]; ++k) b[bi[c++]] = in[o++];
    }
    if (!h2d(p, p->w.p, w.data(), nw * 4) || !h2d(p, p->bias.p, b.data(), nb * 4) || !refresh_bf16(p) ||
        hipStreamSynchronize(p->s) != hipSuccess)
        return -1;
    return 0;
}

int pcbprior_logits(pcbprior* p, const float* x, int n, int h, int w, float* logits) {
    g_err.clear();
    if (!p || !x || !logits) return -1;
    if (!forward(p, x, n, h, w)) return -1;
    if (hipMemcpyAsync(logits, p->logits.p, p->npix * p->cout * 4, hipMemcpyDeviceToHost, p->s) != hipSuccess ||
        hipStreamSynchronize(p->s) != hipSuccess)
        return -1;
    collect(p);
    return 0;
}

int pcbprior_predict(pcbprior* p, const float* x, int n, int h, int w, float* prob) {
    return pcbprior_predict_valid(p, x, n, h, w, nullptr, prob);
}

int pcbprior_predict_valid(pcbprior* p, const float* x, int n, int h, int w, const int32_t* hw, float* prob) {
    g_err.clear();
    if (!p || !x || !prob) return -1;
    if (hw)
        for (int k = 0; k < n; ++k)
            if (hw[2 * k] < 1 || hw[2 * k] > h || hw[2 * k + 1] < 1 || hw[2 * k + 1] > w || hw[2 * k + 1] % PX) {
                g_err = "valid extents: 1 <= rows <= h, columns a multiple of 64 up to w";
                return -1;
            }
    if (!forward(p, x, n, h, w, hw)) return -1;
    const size_t nz = p->npix * p->cout;
    hipLaunchKernelGGL(k_sigmoid, dim3(g1(nz)), dim3(256), 0, p->s, p->logits.p, p->logits.p, nz);
    if (hipMemcpyAsync(prob, p->logits.p, nz * 4, hipMemcpyDeviceToHost, p->s) != hipSuccess ||
        hipStreamSynchronize(p->s) != hipSuccess)
        return -1;
    collect(p);
    return 0;
}

int pcbprior_loss_grad(pcbprior* p, const float* x, const float* target, const float* mask, int n, int h, int w,
                       float* loss, float* grads) {
    g_err.clear();
    if (!p || !x || !target || !mask) return -1;
    if (!forward(p, x, n, h, w) || !backward(p, target, mask, n, h, w, loss)) return -1;
    if (grads) {
        const size_t nw = size_t(p->depth) * C * KK, nb = size_t(p->depth) * C;
        std::vector<float> g(nw), gb(nb);
        if (!d2h(p, g.data(), p->g.p, nw * 4) || !d2h(p, gb.data(), p->gb.p, nb * 4) || !sync(p)) return -1;
        std::vector<size_t> wi, bi;
        pack_layout(p, wi, bi);
        size_t o = 0, a = 0, c = 0;
        for (int L = 0; L < p->depth; ++L) {
            const size_t nwl = size_t(p->lout[L]) * 9 * p->lin[L];
            for (size_t k = 0; k < nwl; ++k) grads[o++] = g[wi[a++]];
            for (int k = 0; k < p->lout[L]; ++k) grads[o++] = gb[bi[c++]];
        }
    }
    return 0;
}

int pcbprior_train_step(pcbprior* p, const float* x, const float* target, const float* mask, int n, int h, int w,
                        float lr, float* loss) {
    g_err.clear();
    if (!p || !x || !target || !mask) return -1;
    if (!forward(p, x, n, h, w) || !backward(p, target, mask, n, h, w, nullptr) || !adam(p, lr)) return -1;
    if (loss && !loss_to_host(p)) return -1;
    if (!sync(p)) return -1;
    if (loss) *loss = loss_sum(p);
    collect(p);
    return 0;
}

int pcbprior_dataset(pcbprior* p, const float* x, const float* target, const float* mask, int n, int h, int w) {
    g_err.clear();
    if (!p || !x || !target || !mask || n < 1 || h < 1 || w < 1 || w % PX) { g_err = "bad training set"; return -1; }
    const size_t spix = size_t(h) * w, tot = size_t(n) * spix;
    if (!p->dsx.need(tot * C) || !p->dst.need(tot * p->cout) || !p->dsm.need(tot) || !p->xin.need(spix * p->cin)) return -1;
    // inputs go over a sample at a time and become BF16 slots on the device
    for (int k = 0; k < n; ++k) {
        if (!h2d(p, p->xin.p, x + size_t(k) * spix * p->cin, spix * p->cin * 4)) return -1;
        hipLaunchKernelGGL(k_to_bf16, dim3(g1(spix * C)), dim3(256), 0, p->s, p->xin.p, p->dsx.p + size_t(k) * spix * C,
                           spix, p->cin);
        if (!sync(p)) return -1;  // xin is reused by the next sample
    }
    if (!h2d(p, p->dst.p, target, tot * p->cout * 4) || !h2d(p, p->dsm.p, mask, tot * 4) || !sync(p)) return -1;
    p->ds_n = n; p->ds_h = h; p->ds_w = w;
    return 0;
}

int pcbprior_train_resident(pcbprior* p, const int32_t* idx, int batch, float lr, float* loss) {
    g_err.clear();
    if (!p || !idx) return -1;
    if (!gather(p, idx, batch) || !forward_dev(p, batch, p->ds_h, p->ds_w) || !backward_dev(p, batch, p->ds_h, p->ds_w) ||
        !adam(p, lr))
        return -1;
    if (loss && !loss_to_host(p)) return -1;
    if (!sync(p)) return -1;
    if (loss) *loss = loss_sum(p);
    collect(p);
    return 0;
}

int pcbprior_forward_resident(pcbprior* p, const int32_t* idx, int batch, float* logits) {
    g_err.clear();
    if (!p || !idx) return -1;
    if (!gather(p, idx, batch) || !forward_dev(p, batch, p->ds_h, p->ds_w)) return -1;
    if (logits && !d2h(p, logits, p->logits.p, p->npix * p->cout * 4)) return -1;
    if (!sync(p)) return -1;
    collect(p);
    return 0;
}

void pcbprior_stats_enable(pcbprior* p, int on) {
    if (p) p->acct = on != 0;
}

void pcbprior_stats_read(pcbprior* p, pcbprior_stats* st, int reset) {
    if (!p) return;
    if (st) *st = p->st;
    if (reset) p->st = pcbprior_stats{};
}

int pcbprior_mfma_check(const float* a, const float* b, float* d) {
    uint16_t ha[512], hb[512];
    for (int i = 0; i < 512; ++i) { ha[i] = f2bf(a[i]); hb[i] = f2bf(b[i]); }
    uint16_t *da = nullptr, *dbb = nullptr;
    float* dd = nullptr;
    bool ok = hipMalloc(&da, 1024) == hipSuccess && hipMalloc(&dbb, 1024) == hipSuccess && hipMalloc(&dd, 1024) == hipSuccess &&
              hipMemcpy(da, ha, 1024, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(dbb, hb, 1024, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_mfma_check, dim3(1), dim3(64), 0, 0, da, dbb, dd);
        ok = hipGetLastError() == hipSuccess && hipMemcpy(d, dd, 1024, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (da) (void)hipFree(da);
    if (dbb) (void)hipFree(dbb);
    if (dd) (void)hipFree(dd);
    return ok ? 0 : -1;
}

const char* pcbprior_last_error(void) { return g_err.c_str(); }

}  // extern "C"
