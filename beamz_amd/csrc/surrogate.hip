// BF16-MFMA / FP32-accumulate kernels of the CNN surrogate (new; no reference
// counterpart): 3x3 convolutions forward, backward-data and backward-weights
// as implicit GEMMs on v_mfma_f32_32x32x16_bf16, plus the small layers
// (first convolution with one input channel, ReLU mask, 2x2 average pooling,
// global average pooling + fully connected head, Adam).
//
// Layout: activations NHWC in bfloat16 (stored as uint16 bit patterns),
// convolution weights [Co][3][3][Ci] (K = tap*Ci + ci contiguous), FP32 master
// weights, gradients and biases.
//
// GEMM tiling: a workgroup of 4 waves computes a 64 x 64 tile of C = A * B^T
// with A [M][K] and B [N][K] both staged K-contiguous in LDS, each wave one
// 32 x 32 block: 64-wide K slices (4 MFMAs per wave and slice) for forward and
// backward-data, 128-pixel slices (8 MFMAs) for the weight gradient; the next
// slice's global loads are in flight during the current slice's MFMAs. The
// operand / result lane layout of the 32x32x16 bf16 MFMA is: lane l supplies
// A[l%32][8*(l/32)+t] and B[8*(l/32)+t][l%32], t = 0..7, and holds
// C[8*(r/4) + 4*(l/32) + r%4][l%32] in accumulator r = 0..15.
//   forward: M = pixels, N = Co, K = 9*Ci, A = im2col(x) built on the fly;
//   dgrad:   M = pixels, N = Ci, K = 9*Co, A = im2col(gz) with mirrored taps,
//            B = weights transposed to [Ci][3][3][Co];
//   wgrad:   M = Co, N = 9*Ci, K = pixels (split over workgroups, FP32
//            atomics), both operands transposed in registers (8 x 8 blocks)
//            before they are stored in LDS.
#include "common.h"

#include <algorithm>
#include <cstdint>
#include <string>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int TM = 64, TN = 64;   // output tile of a workgroup (4 waves, one 32 x 32 block each)

__device__ __forceinline__ float b2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2b(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);     // round to nearest even (finite values)
    return (uint16_t)(u >> 16);
}

struct Geo {
    int B, H, W;
};

__device__ __forceinline__ void pixel(int p, const Geo& g, int& b, int& y, int& x) {
    x = p % g.W;
    const int q = p / g.W;
    y = q % g.H;
    b = q / g.H;
}

// Forward (sign = +1) or backward-data (sign = -1) 3x3 convolution as an implicit GEMM.
// src: [P][Cs] bf16 (x, or the masked upstream gradient), wk: [N][9][Cs] bf16, out: [P][N] bf16.
// K slices of FK = 32 or 64 elements, CPR = FK / 8 chunks of 16 bytes per tile row: thread t stages chunk t % CPR
// of rows t / CPR + h 256 / CPR (h < NH) of both tiles (the tap of a chunk is k / Cs; Cs % 8 == 0, so a chunk
// never straddles two taps, and chunks past K are zero). The next slice's loads are issued before the current
// slice's MFMAs.
// Tile rows are padded to FK + 8 elements, so the consecutive lanes of an MFMA operand read (rows r .. r + 15,
// one chunk each) fall on distinct banks; with FK = 64 the eight lanes storing a row write its 128 contiguous
// bytes (four lanes per row, two chunks each, put two rows on the same banks).
// TAP (Cs % FK == 0): a slice lies within one tap, so the tap and its pixel shift are computed once per slice
// (uniform) and each lane adds offsets to its pixels' precomputed bases, instead of two integer divisions per
// chunk and lane (the per-lane index arithmetic outweighed the MFMAs several times over).
template <int FK, bool TAP>
__global__ __launch_bounds__(256, TAP ? 8 : 1) void k_conv_gemm(const uint16_t* __restrict__ src, const uint16_t* __restrict__ wk,
                                                   const float* __restrict__ bias, uint16_t* __restrict__ out,
                                                   const uint16_t* __restrict__ mask, Geo g, int Cs, int N, int sign,
                                                   int relu) {
    constexpr int CPR = FK / 8, NH = FK / 32, RSTEP = 256 / CPR;
    constexpr int FLD = FK + 8;
    __shared__ __attribute__((aligned(16))) uint16_t As[TM][FLD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][FLD];
    const int P = g.B * g.H * g.W;
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
    const int t = threadIdx.x;
    const int pr = t / CPR, cq = t % CPR, kq = cq * 8;
    const int K = 9 * Cs;
    // per staged row h: the pixel (TAP: its element offset, P * Cs < 2^31, checked by the entry points, and in
    // bits 16 h + tap of tapok whether the tap's source pixel is inside the image; else its b, y, x)
    int py[NH], px[NH], pofs[NH], wofs[NH];
    unsigned tapok = 0;
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        wofs[h] = min(n0 + pr + h * RSTEP, N - 1) * K + kq;   // N * K < 2^31 (checked by the entry points)
        int pb;
        pixel(min(m0 + pr + h * RSTEP, P - 1), g, pb, py[h], px[h]);
        pofs[h] = TAP ? ((pb * g.H + py[h]) * g.W + px[h]) * Cs + kq : pb;
        if (TAP && m0 + pr + h * RSTEP < P) {
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) {
                const int yy = py[h] + sign * (tap / 3 - 1), xx = px[h] + sign * (tap % 3 - 1);
                if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W) tapok |= 1u << (16 * h + tap);
            }
        }
    }
    uint4 a[NH], bw[NH];
    auto load = [&](int k0) {
        if (TAP) {
            const int tap = k0 / Cs;
            const int ky = tap / 3, kx = tap - 3 * ky;
            const int dy = sign * (ky - 1), dx = sign * (kx - 1);
            const int cs = k0 - tap * Cs;
            const int shift = (dy * g.W + dx) * Cs;
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const bool ok = (tapok >> (16 * h + tap)) & 1u;
                // every lane loads from a valid address (its own pixel where the tap falls outside) and selects 0
                const uint4 v = *reinterpret_cast<const uint4*>(src + pofs[h] + cs + (ok ? shift : 0));
                const uint4 w = *reinterpret_cast<const uint4*>(wk + wofs[h] + k0);
                a[h] = ok ? v : make_uint4(0, 0, 0, 0);
                bw[h] = n0 + pr + h * RSTEP < N ? w : make_uint4(0, 0, 0, 0);
            }
            return;
        }
        const int k = k0 + kq;
        const int tap = k / Cs, c = k - tap * Cs;
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            a[h] = make_uint4(0, 0, 0, 0);
            bw[h] = make_uint4(0, 0, 0, 0);
            if (k < K) {
                const int yy = py[h] + sign * (ky - 1), xx = px[h] + sign * (kx - 1);
                if (m0 + pr + h * RSTEP < P && yy >= 0 && yy < g.H && xx >= 0 && xx < g.W)
                    a[h] = *reinterpret_cast<const uint4*>(src + (((long)pofs[h] * g.H + yy) * g.W + xx) * Cs + c);
                const int nrow = n0 + pr + h * RSTEP;
                if (nrow < N) bw[h] = *reinterpret_cast<const uint4*>(wk + (long)nrow * K + k);
            }
        }
    };
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const int l = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int ra = wm * 32 + (l & 31), rb = wn * 32 + (l & 31);
    load(0);
    for (int k0 = 0; k0 < K; k0 += FK) {
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const int r = pr + h * RSTEP;
            *reinterpret_cast<uint4*>(&As[r][kq]) = a[h];
            *reinterpret_cast<uint4*>(&Bs[r][kq]) = bw[h];
        }
        __syncthreads();
        if (k0 + FK < K) load(k0 + FK);
#pragma unroll
        for (int kk = 0; kk < FK; kk += 16) {
            const uint4 av = *reinterpret_cast<const uint4*>(&As[ra][kk + 8 * (l >> 5)]);
            const uint4 bv = *reinterpret_cast<const uint4*>(&Bs[rb][kk + 8 * (l >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                           __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int n = n0 + wn * 32 + (l & 31);
    if (n >= N) return;
    const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
        if (m < P) {
            float v = acc[r] + bn;
            if (relu) v = fmaxf(v, 0.0f);
            const long o = (long)m * N + n;
            out[o] = (mask && !(b2f(mask[o]) > 0.0f)) ? (uint16_t)0 : f2b(v);
        }
    }
}

// Forward (sign = +1) or backward-data (sign = -1) 3x3 convolution by a march down image rows (W = 64 or 128):
//   out[p][n] = act(bias[n] + sum_{ky, kx, c} src[p + sign (ky - 1, kx - 1)][c] * wk[n][ky][kx][c]).
// A workgroup owns CB output channels (blockIdx.y) and a band of output rows (blockIdx.z, rows flattened as
// b * H + y). It keeps the weights of its channels (all taps, all CIN input channels) in LDS for the whole band,
// and a ring of four input rows, pixel-major with a zero pixel on each side: the rows r - 1 .. r + 1 that
// output row r reads and the row being stored. The rows a step needs two steps later are loaded into registers
// at its start and stored after the next step (one barrier per row), so each input row is read from memory
// once per band and channel block. Wave w computes the 32-pixel block w % (W / 32) and the 32-channel block
// w / (W / 32) of every row: MFMAs with M = pixels, N = output channels, K = input channels, nine taps of
// CIN / 16 steps. With a mask (the forward input of a backward-data step, an activation after ReLU), outputs
// where the mask is not positive are stored as 0: the ReLU derivative of the layer below, fused. A tap shifts whole pixels, so every fragment is one aligned 16-byte LDS read; pixel strides
// of CIN + 8 and weight rows of 9 CIN + 8 elements put the 16 rows of a read on distinct banks.
template <int W, int CIN, int CB, bool MASK>
__global__ __launch_bounds__(64 * (W / 32) * (CB / 32)) void k_conv_march(const uint16_t* __restrict__ src,
                                                                          const uint16_t* __restrict__ wk,
                                                                          const float* __restrict__ bias,
                                                                          uint16_t* __restrict__ out,
                                                                          const uint16_t* __restrict__ mask, int H,
                                                                          int rows_total, int Cout, int sign,
                                                                          int relu, int band) {
    constexpr int NT = 64 * (W / 32) * (CB / 32);
    constexpr int WS = 9 * CIN + 8, XC = CIN + 8, XSLOT = (W + 2) * XC;
    constexpr int CHUNKS = W * CIN / 8, NCH = (CHUNKS + NT - 1) / NT;   // 16-byte chunks of one input row
    __shared__ __attribute__((aligned(16))) uint16_t lds[CB * WS + 4 * XSLOT];
    uint16_t* wl = lds;
    uint16_t* ring = lds + CB * WS;
    const int t = threadIdx.x, l = t & 63, wave = t >> 6;
    const int n0 = blockIdx.y * CB;
    const int r0 = blockIdx.z * band, r1 = min(rows_total, r0 + band);
    for (int q = t; q < 4 * XSLOT / 8; q += NT) reinterpret_cast<uint4*>(ring)[q] = make_uint4(0, 0, 0, 0);
    for (int q = t; q < CB * (9 * CIN / 8); q += NT) {
        const int n = q / (9 * CIN / 8), k8 = q - n * (9 * CIN / 8);
        *reinterpret_cast<uint4*>(wl + n * WS + 8 * k8) =
            *reinterpret_cast<const uint4*>(wk + (long)(n0 + n) * 9 * CIN + 8 * k8);
    }
    uint4 va[NCH], vb[NCH];
    auto load = [&](uint4 (&v)[NCH], int q) {      // input row q (contiguous W * CIN elements), zeros outside
        const bool ok = q >= 0 && q < rows_total;
#pragma unroll
        for (int i = 0; i < NCH; ++i) {
            const int ch = t + i * NT;
            v[i] = make_uint4(0, 0, 0, 0);
            if (ok && ch < CHUNKS) v[i] = *reinterpret_cast<const uint4*>(src + (long)q * W * CIN + 8 * ch);
        }
    };
    auto store = [&](const uint4 (&v)[NCH], int q) {
        uint16_t* slot = ring + ((q + 4) & 3) * XSLOT;
#pragma unroll
        for (int i = 0; i < NCH; ++i) {
            const int ch = t + i * NT;
            if (ch < CHUNKS) {
                const int p = ch / (CIN / 8), c8 = ch - p * (CIN / 8);
                *reinterpret_cast<uint4*>(slot + (p + 1) * XC + 8 * c8) = v[i];
            }
        }
    };
    __syncthreads();
    for (int q = r0 - 1; q <= r0 + 1; ++q) {
        load(va, q);
        store(va, q);
    }
    if (r0 + 1 < r1) load(va, r0 + 2);
    __syncthreads();
    const int mb = wave % (W / 32), nb = wave / (W / 32);
    const int m = l & 31, g = l >> 5;
    const int n = n0 + nb * 32 + m;
    const float bn = bias ? bias[n] : 0.0f;
    const uint16_t* wrow = wl + (nb * 32 + m) * WS + 8 * g;
    auto compute = [&](int r) {
        const int y = r % H;
        const long o0 = ((long)r * W + mb * 32) * Cout + n;
        // the row's mask values load before the MFMAs, which hide their latency
        uint16_t mk[16];
        if (MASK)
#pragma unroll
            for (int e = 0; e < 16; ++e) mk[e] = mask[o0 + (long)(8 * (e >> 2) + 4 * g + (e & 3)) * Cout];
        f32x16 acc;
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int dy = sign * (ky - 1);
            if (y + dy < 0 || y + dy >= H) continue;                   // zero padding above / below
            const uint16_t* xrow = ring + ((r + dy + 4) & 3) * XSLOT + 8 * g;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const uint16_t* a = xrow + (mb * 32 + m + sign * (kx - 1) + 1) * XC;
                const uint16_t* b = wrow + (3 * ky + kx) * CIN;
#pragma unroll
                for (int k = 0; k < CIN; k += 16)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(a + k)),
                        __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(b + k)), acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            float v = acc[e] + bn;
            if (relu) v = fmaxf(v, 0.0f);
            const long o = o0 + (long)(8 * (e >> 2) + 4 * g + (e & 3)) * Cout;
            if (MASK) out[o] = b2f(mk[e]) > 0.0f ? f2b(v) : (uint16_t)0;
            else out[o] = f2b(v);
        }
    };
    for (int r = r0; r < r1; r += 2) {
        if (r + 2 < r1) load(vb, r + 3);
        compute(r);
        if (r + 1 < r1) store(va, r + 2);
        __syncthreads();
        if (r + 1 >= r1) break;
        if (r + 3 < r1) load(va, r + 4);
        compute(r + 1);
        if (r + 2 < r1) store(vb, r + 3);
        __syncthreads();
    }
}

// Forward / backward-data 3x3 convolution with the input halo staged once per channel chunk.
// A workgroup computes 64 consecutive pixels (one image; a 64-pixel row segment if W >= 64, else 64 / W
// whole rows) x 64 output channels. For each chunk of CC input channels it loads the (rows + 2) x (cols + 2)
// pixel halo of the tile into LDS (zero outside the image), then forms all nine taps from it: the A
// fragment of output pixel m at tap (ky, kx) is the halo slot of pixel m shifted by sign * (ky - 1, kx - 1).
// Each activation is read from memory once per chunk instead of once per tap. The B fragments (weights,
// [N][9][Cs], K-contiguous) come straight from global memory (L2), the next tap's while the current tap's
// MFMAs run. Needs H*W % 64 == 0 and W % 64 == 0 or 64 % W == 0, Cs % CC == 0.
// Used for 32-channel inputs (level-0 layers: 172 -> 139 us for batch 128 at 128 x 128). With 64 or more
// channels the direct weight reads (1 KB per MFMA and wave from L2) cost more than the tap re-reads they
// replace (92 -> 125 us), so those layers keep k_conv_gemm; staging the weights in LDS is the next step.
template <int CC>
__global__ __launch_bounds__(256) void k_conv_halo(const uint16_t* __restrict__ src, const uint16_t* __restrict__ wk,
                                                   const float* __restrict__ bias, uint16_t* __restrict__ out, const uint16_t* __restrict__ mask, Geo g,
                                                   int Cs, int N, int sign, int relu) {
    constexpr int SLD = CC + 8;              // halfwords per halo slot (16-byte pad: conflict-free reads)
    constexpr int KS = CC / 16;              // MFMA k-steps per tap and chunk
    __shared__ __attribute__((aligned(16))) uint16_t halo[198][SLD];
    const int HW = g.H * g.W;
    const int p0 = blockIdx.x * 64, n0 = blockIdx.y * TN;
    const int b = p0 / HW, rem = p0 - b * HW;
    const int y0 = rem / g.W, x0 = rem - y0 * g.W;
    const int ncols = g.W < 64 ? g.W : 64, nrows = 64 / ncols;
    const int hc = ncols + 2, slots = (nrows + 2) * hc;
    const int t = threadIdx.x, l = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int K = 9 * Cs;
    // this lane's output pixel of the A fragment and its B row
    const int m = wm * 32 + (l & 31);
    const int my = m / ncols, mx = m - my * ncols;
    const int n = n0 + wn * 32 + (l & 31);
    const int kq = 8 * (l >> 5);
    const uint16_t* wrow = wk + (long)min(n, N - 1) * K;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    uint4 bq[2][KS];
    auto load_b = [&](uint4* dst, int tap, int c0) {
#pragma unroll
        for (int s = 0; s < KS; ++s)
            dst[s] = n < N ? *reinterpret_cast<const uint4*>(wrow + tap * Cs + c0 + 16 * s + kq) : make_uint4(0, 0, 0, 0);
    };
    for (int c0 = 0; c0 < Cs; c0 += CC) {
        // halo of the tile for channels c0 .. c0 + CC - 1
        for (int idx = t; idx < slots * (CC / 8); idx += 256) {
            const int slot = idx / (CC / 8), cg = idx - slot * (CC / 8);
            const int hy = slot / hc, hx = slot - hy * hc;
            const int py = y0 + hy - 1, px = x0 + hx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (py >= 0 && py < g.H && px >= 0 && px < g.W)
                v = *reinterpret_cast<const uint4*>(src + (((long)b * g.H + py) * g.W + px) * Cs + c0 + 8 * cg);
            *reinterpret_cast<uint4*>(&halo[slot][8 * cg]) = v;
        }
        load_b(bq[0], 0, c0);
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            if (tap < 8) load_b(bq[(tap + 1) & 1], tap + 1, c0);
            const int ky = tap / 3, kx = tap - 3 * (tap / 3);
            const int slot = (my + 1 + sign * (ky - 1)) * hc + (mx + 1 + sign * (kx - 1));
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const uint4 av = *reinterpret_cast<const uint4*>(&halo[slot][16 * s + kq]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                               __builtin_bit_cast(bf16x8, bq[tap & 1][s]), acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (n >= N) return;
    const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int mm = p0 + wm * 32 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
        float v = acc[r] + bn;
        if (relu) v = fmaxf(v, 0.0f);
        const long o = (long)mm * N + n;
        out[o] = (mask && !(b2f(mask[o]) > 0.0f)) ? (uint16_t)0 : f2b(v);
    }
}

inline bool halo_ok(const Geo& g) {
    return (g.H * g.W) % 64 == 0 && (g.W % 64 == 0 || 64 % g.W == 0) && (64 / (g.W < 64 ? g.W : 64) + 2) *
           ((g.W < 64 ? g.W : 64) + 2) <= 198;
}

// Weight gradient gw[co][tap][ci] += sum_p gz[p][co] * x[p + tap][ci] over this workgroup's pixel range
// (a multiple of WK pixels, the last one clipped at P). GEMM with M = Co, N = 9*Ci, K = pixels.
// Staging of one WK-pixel slice: threads 0..127 take the gz operand, 128..255 the im2col(x) operand; each
// loads 8 pixels x 8 channels (eight 16-byte rows), transposes that block in registers and stores 8
// channel rows of 8 pixels (16 bytes each) into the K-contiguous LDS tiles. The next slice's loads are
// issued before the current slice's MFMAs.
constexpr int WK = 128, WLD = WK + 8;

__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }

__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ gz,
                                                    float* __restrict__ gw, float* __restrict__ part, Geo g, int Ci,
                                                    int Co, int chunk) {
    __shared__ __attribute__((aligned(16))) uint16_t As[TM][WLD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][WLD];
    const int P = g.B * g.H * g.W;
    const int N = 9 * Ci;
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
    const int pstart = blockIdx.z * chunk, pend = min(P, pstart + chunk);
    const int t = threadIdx.x;
    const bool opb = t >= 128;
    const int u = t & 127;
    const int pg = u >> 3, q8 = (u & 7) * 8;      // 8-pixel group of the slice, 8-wide channel group
    const int nb = n0 + q8;
    const int tap = nb / Ci, ci = nb - tap * Ci;
    const int ky = tap / 3, kx = tap - 3 * (tap / 3);
    const bool colok = opb ? nb < N : m0 + q8 < Co;
    uint4 v[8];
    auto load = [&](int p0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int p = p0 + pg * 8 + j;
            uint4 a = make_uint4(0, 0, 0, 0);
            if (p < pend && colok) {
                if (!opb) {
                    a = *reinterpret_cast<const uint4*>(gz + (long)p * Co + m0 + q8);
                } else {
                    int pb, py, px;
                    pixel(p, g, pb, py, px);
                    const int yy = py + ky - 1, xx = px + kx - 1;
                    if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W)
                        a = *reinterpret_cast<const uint4*>(x + (((long)pb * g.H + yy) * g.W + xx) * Ci + ci);
                }
            }
            v[j] = a;
        }
    };
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    uint16_t (*dst)[WLD] = opb ? Bs : As;
    load(pstart);
    for (int p0 = pstart; p0 < pend; p0 += WK) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            uint16_t h[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) h[j] = reinterpret_cast<const uint16_t*>(&v[j])[e];
            *reinterpret_cast<uint4*>(&dst[q8 + e][pg * 8]) =
                make_uint4(pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7]));
        }
        __syncthreads();
        if (p0 + WK < pend) load(p0 + WK);
        const int l = t & 63, wave = t >> 6;
        const int wm = wave & 1, wn = wave >> 1;
#pragma unroll
        for (int kk = 0; kk < WK; kk += 16) {
            const uint4 av = *reinterpret_cast<const uint4*>(&As[wm * 32 + (l & 31)][kk + 8 * (l >> 5)]);
            const uint4 bv = *reinterpret_cast<const uint4*>(&Bs[wn * 32 + (l & 31)][kk + 8 * (l >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                           __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int l = t & 63, wave = t >> 6;
    const int n = n0 + (wave >> 1) * 32 + (l & 31);
    if (n >= N) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + (wave & 1) * 32 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
        if (m < Co) {
            if (part)
                part[((long)blockIdx.z * Co + m) * N + n] = acc[r];   // summed in split order by k_sum_parts
            else
                atomicAdd(gw + (long)m * N + n, acc[r]);
        }
    }
}

// Weight and bias gradients by a march down image rows (W = 16, 32, 64 or 128 pixels per row):
//   dW[co][ky][kx][ci] = sum_p gz[p][co] * x[p + (ky - 1, kx - 1)][ci],   db[co] = sum_p gz[p][co]
// over the pixels of one band of rows (rows flattened as b * H + y), as MFMA GEMMs with K = pixels.
// A workgroup owns one block of CB = 32 NCO output and IB = 32 NCI input channels and computes all nine taps
// and the bias for it, so each activation and each gradient entry of its band is read from memory once. A
// step covers R rows (R W >= 64 pixels); LDS holds a ring of 2R + 2 input rows (the R + 2 rows a step reads and
// the R rows stored for the next step) and 2R gradient rows, all transposed to channel-major rows of pixels.
// The rows a step needs two steps later are loaded into registers at its start and stored into LDS after the
// next step (one barrier per step).
// The workgroup has 3 NCO NCI waves: wave w takes the 32 x 32 channel sub-block w / 3 and the kernel row
// ky = w % 3 (three taps, 3 x 16 accumulators per lane) over every pixel; the ky = 1 waves of the ci-block 0
// also accumulate the bias (a product with a ones operand). The +-1 pixel shifts of kx = 0 and 2 are formed
// from the aligned 16-byte read of kx = 1 and one neighbouring dword by byte alignment (the LDS rows carry
// zero pads of 8 and 16 elements). Image edges: a wave skips the row y - 1 < 0 or y + 1 >= H (zero padding);
// pixel edges read the pads. Partial sums go to part[band][Co][9 Ci] and bpart[band][Co] (blockIdx.y == 0).
// LDS rows are padded to strides of 12 mod 64 (inputs) and 4 mod 64 (gradients) dwords, so the 16-byte
// fragment reads of 16 consecutive rows touch distinct banks.
template <int NCO, int NCI, int W, int R>
__global__ __launch_bounds__(192 * NCO * NCI) void k_wgrad_march(const uint16_t* __restrict__ x,
                                                                 const uint16_t* __restrict__ gz,
                                                                 float* __restrict__ part, float* __restrict__ bpart,
                                                                 int H, int rows_total, int Ci, int Co, int band) {
    constexpr int CB = 32 * NCO, IB = 32 * NCI, NT = 192 * NCO * NCI;
    // neighbour pixels by lane exchange, except where its four more live registers would spill (the 2 x 2
    // channel blocks, 3 waves per SIMD, with more than one 16-pixel chunk per row)
    constexpr bool XCH = !(NCO == 2 && NCI == 2 && W > 16);
    constexpr int XS = W + 24, GS = W + 8;                  // row strides (elements) of the x and gz tiles
    constexpr int XSLOT = IB * XS, GSLOT = CB * GS, NXS = 2 * R + 2, NGS = 2 * R;
    constexpr int NBX = (W / 8) * (IB / 8), NBG = (W / 8) * (CB / 8);
    static_assert(R * (NBX + NBG) <= NT, "one staging block per thread and set");
    __shared__ __attribute__((aligned(16))) uint16_t lds[NXS * XSLOT + NGS * GSLOT];
    uint16_t* xs = lds;
    uint16_t* gs = lds + NXS * XSLOT;
    const int t = threadIdx.x, l = t & 63, wave = t >> 6;
    const int co_base = blockIdx.x * CB, ci_base = blockIdx.y * IB;
    const int r0 = blockIdx.z * band, r1 = min(rows_total, r0 + band);
    // zero both tiles once: the pads stay zero, the staged rows overwrite the rest
    for (int q = t; q < (NXS * XSLOT + NGS * GSLOT) / 8; q += NT)
        reinterpret_cast<uint4*>(lds)[q] = make_uint4(0, 0, 0, 0);
    // staging role: one 8-pixel x 8-channel block of one of R input rows (t < R NBX) or gradient rows
    const bool isx = t < R * NBX, isg = !isx && t < R * (NBX + NBG);
    const int u = isx ? t : t - R * NBX;
    const int nb = isx ? NBX : NBG, nch = isx ? IB / 8 : CB / 8;
    // row of the set, channel chunk, pixel group: consecutive lanes take consecutive pixel groups of one channel
    // chunk, so the transposed stores of eight lanes write contiguous bytes of one LDS row
    const int rr = u / nb, pg = (u % nb) % (W / 8), cc = (u % nb) / (W / 8);
    // two register sets in flight: rows loaded at the start of a step are stored after the next step
    uint4 va[8], vb[8];
    auto load = [&](uint4 (&v)[8], int q) {   // row q of x (isx) or gz (isg), zeros outside 0 .. rows_total-1
        const bool ok = (isx || isg) && q >= 0 && q < rows_total;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            v[j] = make_uint4(0, 0, 0, 0);
            if (ok) {
                const long pix = (long)q * W + pg * 8 + j;
                v[j] = isx ? *reinterpret_cast<const uint4*>(x + pix * Ci + ci_base + cc * 8)
                           : *reinterpret_cast<const uint4*>(gz + pix * Co + co_base + cc * 8);
            }
        }
    };
    auto store = [&](const uint4 (&v)[8], int q) {   // transpose the 8 x 8 block to channel rows, store in q's slot
        if (!(isx || isg)) return;
        uint16_t* dst = isx ? xs + ((q + NXS) % NXS) * XSLOT + (cc * 8) * XS + 8 + pg * 8
                            : gs + (q % NGS) * GSLOT + (cc * 8) * GS + pg * 8;
        const int stride = isx ? XS : GS;
        const uint32_t* w = reinterpret_cast<const uint32_t*>(v);   // w[4j + d]: pixel j, channels 2d, 2d+1
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint32_t sel = (c & 1) ? 0x07060302u : 0x05040100u;   // high or low halves of two dwords
            const int d = c >> 1;
            uint4 o;
            o.x = __builtin_amdgcn_perm(w[4 * 1 + d], w[4 * 0 + d], sel);
            o.y = __builtin_amdgcn_perm(w[4 * 3 + d], w[4 * 2 + d], sel);
            o.z = __builtin_amdgcn_perm(w[4 * 5 + d], w[4 * 4 + d], sel);
            o.w = __builtin_amdgcn_perm(w[4 * 7 + d], w[4 * 6 + d], sel);
            *reinterpret_cast<uint4*>(dst + c * stride) = o;
        }
    };
    // this thread's row of the set a step stores: input rows rs + R + 1 + rr, gradient rows rs + R + rr
    auto next_row = [&](int rs) { return isx ? rs + R + 1 + rr : rs + R + rr; };
    __syncthreads();
    // prologue: input rows r0 - 1 .. r0 + R, gradient rows r0 .. r0 + R - 1, then the set for the second step
    if (isx)
        for (int q = r0 - 1 + rr; q <= r0 + R; q += R) {
            load(va, q);
            store(va, q);
        }
    if (isg) {
        load(va, r0 + rr);
        store(va, r0 + rr);
    }
    if (r0 + R < r1) load(va, next_row(r0));
    __syncthreads();
    const int s = wave / 3, ky = wave % 3;
    const int co32 = s % NCO, ci32 = s / NCO;
    const int row = l & 31, g = l >> 5;
    const bool with_bias = blockIdx.y == 0 && ci32 == 0 && ky == 1;
    f32x16 acc[3], accb;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[k][e] = 0.0f;
#pragma unroll
    for (int e = 0; e < 16; ++e) accb[e] = 0.0f;
    const uint32_t one2 = 0x3F803F80u;                        // two bf16 ones
    const bf16x8 ones = __builtin_bit_cast(bf16x8, make_uint4(one2, one2, one2, one2));
    auto compute = [&](int rs) {
        for (int r = rs; r < min(rs + R, r1); ++r) {
            const int y = r % H;
            const uint16_t* grow = gs + (r % NGS) * GSLOT + (co32 * 32 + row) * GS + 8 * g;
            // this wave's input row r + ky - 1 exists (else zero padding: nothing to add)
            if (!((ky == 0 && y == 0) || (ky == 2 && y == H - 1))) {
                const uint16_t* xrow = xs + ((r + ky - 1 + NXS) % NXS) * XSLOT + (ci32 * 32 + row) * XS + 8 + 8 * g;
                // the pixel pairs beside a lane's eight (p - 2, p - 1 and p + 8, p + 9) are the first or last dword
                // of lane l +- 32's chunk, of this k or the next: exchanged with permlane32_swap (in result 0 the
                // upper lanes hold the lower lanes' second operand, in result 1 the lower lanes hold the upper
                // lanes' first operand) instead of two 4-byte LDS reads that put 32 rows on 8 bank groups; the
                // row pads give the zeros at both ends
                uint32_t carry = 0;                          // lower lanes: pixels p - 2, p - 1 of the next k
                uint4 curn = XCH ? *reinterpret_cast<const uint4*>(xrow) : make_uint4(0, 0, 0, 0);
#pragma unroll
                for (int k = 0; k < W / 16; ++k) {
                    const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(grow + 16 * k));
                    if (with_bias) accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, ones, accb, 0, 0, 0);
                    uint4 cur;
                    uint32_t prev, next;
                    if (XCH) {
                        cur = curn;
                        curn = k + 1 < W / 16 ? *reinterpret_cast<const uint4*>(xrow + 16 * (k + 1))
                                              : make_uint4(0, 0, 0, 0);
                        const auto s1 = __builtin_amdgcn_permlane32_swap(cur.x, cur.w, false, false);
                        const auto s2 = __builtin_amdgcn_permlane32_swap(cur.w, curn.x, false, false);
                        prev = g ? (uint32_t)s1[0] : carry;
                        next = g ? (uint32_t)s2[0] : (uint32_t)s1[1];
                        carry = s2[1];
                    } else {
                        const uint16_t* xr = xrow + 16 * k;
                        cur = *reinterpret_cast<const uint4*>(xr);
                        prev = *reinterpret_cast<const uint32_t*>(xr - 2);
                        next = *reinterpret_cast<const uint32_t*>(xr + 8);
                    }
                    const uint32_t m1 = __builtin_amdgcn_alignbyte(cur.y, cur.x, 2);   // pixels 1, 2
                    const uint32_t m2 = __builtin_amdgcn_alignbyte(cur.z, cur.y, 2);   // 3, 4
                    const uint32_t m3 = __builtin_amdgcn_alignbyte(cur.w, cur.z, 2);   // 5, 6
                    const uint4 left = make_uint4(__builtin_amdgcn_alignbyte(cur.x, prev, 2), m1, m2, m3);   // p - 1
                    const uint4 right = make_uint4(m1, m2, m3, __builtin_amdgcn_alignbyte(next, cur.w, 2));  // p + 1
                    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, left), acc[0], 0,
                                                                     0, 0);
                    acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, cur), acc[1], 0,
                                                                     0, 0);
                    acc[2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, right), acc[2], 0,
                                                                     0, 0);
                }
            } else if (with_bias) {
#pragma unroll
                for (int k = 0; k < W / 16; ++k)
                    accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(grow + 16 * k)), ones, accb, 0, 0,
                        0);
            }
        }
    };
    // step rs: the set for step rs + 2R is loaded while the other set (loaded during the previous step, for
    // step rs + R) is stored after the compute, into ring slots no read of step rs touches
    for (int rs = r0; rs < r1; rs += 2 * R) {
        if (rs + 2 * R < r1) load(vb, next_row(rs + R));
        compute(rs);
        if (rs + R < r1) store(va, next_row(rs));
        __syncthreads();
        if (rs + R >= r1) break;
        if (rs + 3 * R < r1) load(va, next_row(rs + 2 * R));
        compute(rs + R);
        if (rs + 2 * R < r1) store(vb, next_row(rs + R));
        __syncthreads();
    }
    const int N = 9 * Ci;
    const int co = co_base + co32 * 32, ci = ci_base + ci32 * 32 + row;
#pragma unroll
    for (int kx = 0; kx < 3; ++kx)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int m = co + 8 * (e >> 2) + 4 * g + (e & 3);
            part[((long)blockIdx.z * Co + m) * N + (3 * ky + kx) * Ci + ci] = acc[kx][e];
        }
    if (with_bias && row == 0)
#pragma unroll
        for (int e = 0; e < 16; ++e) bpart[(long)blockIdx.z * Co + co + 8 * (e >> 2) + 4 * g + (e & 3)] = accb[e];
}

// Deterministic second pass of the split reductions: Z partial rows of `count` floats are added in a fixed
// tree order. k_sum_group adds the rows 16g .. 16g+15 of `in` into row g of `outp`; k_sum_parts adds the
// last (at most 16) rows into out, in row order.
constexpr int SUM_FAN = 16;

__global__ void k_sum_group(const float* __restrict__ in, int Z, long sin, long count, float* __restrict__ outp,
                            long sout) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;   // element (x), group of rows (y)
    const long g = blockIdx.y;
    const long z0 = g * SUM_FAN;
    if (i >= count || z0 >= Z) return;
    const long z1 = min((long)Z, z0 + SUM_FAN);
    float s = in[z0 * sin + i];
    for (long z = z0 + 1; z < z1; ++z) s += in[z * sin + i];
    outp[g * sout + i] = s;
}

__global__ void k_sum_parts(const float* __restrict__ part, int Z, long stride, long count, float* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = out[i];
    for (int z = 0; z < Z; ++z) s += part[(long)z * stride + i];
    out[i] = s;
}

// The whole tree of k_sum_group levels and the closing k_sum_parts in one launch, for up to two segments (a
// layer's weight and bias partials): the same additions in the same order, so the same bits. A workgroup
// takes SE consecutive elements of one segment. Its 256 / SE lanes per element add the first level's groups
// of SUM_FAN rows (group g on lane g mod (256 / SE)) into LDS; one lane per element then adds the further
// levels in place and the output, in order.
struct SumSeg {
    float* in;        // the partial rows (workspace; the fallback through tree_sum overwrites them)
    long stride, count;
    float* out;
    int Z;
};
constexpr int SUM_LDS_FLOATS = 16384;     // first-level sums one workgroup holds (64 KB)

template <int SE>
__global__ __launch_bounds__(256) void k_tree_sum(SumSeg a, SumSeg b, int nblk_a) {
    extern __shared__ float lv[];     // [first-level group][SE]
    constexpr int ST = 256 / SE;
    const bool second = (int)blockIdx.x >= nblk_a;
    const SumSeg S = second ? b : a;
    const long i = (long)(second ? (int)blockIdx.x - nblk_a : (int)blockIdx.x) * SE + threadIdx.x % SE;
    const int e = threadIdx.x % SE, t = threadIdx.x / SE;
    const bool ok = i < S.count;
    // rows z0 .. z1-1 (at most SUM_FAN) of element i added in order: all loads issued first, then the adds
    // (a row past z1 is skipped, not added as 0, which would turn -0 into +0)
    auto group = [&](int z0, int z1, float s, bool first) {
        float v[SUM_FAN];
#pragma unroll
        for (int q = 0; q < SUM_FAN; ++q) v[q] = z0 + q < z1 ? S.in[(long)(z0 + q) * S.stride + i] : 0.0f;
        if (first) s = v[0];
#pragma unroll
        for (int q = 0; q < SUM_FAN; ++q)
            if ((q > 0 || !first) && z0 + q < z1) s += v[q];
        return s;
    };
    if (S.Z <= SUM_FAN) {             // uniform over the workgroup
        if (t == 0 && ok) S.out[i] = group(0, S.Z, S.out[i], false);
        return;
    }
    const int G1 = (S.Z + SUM_FAN - 1) / SUM_FAN;
    if (ok)
        for (int g = t; g < G1; g += ST) lv[g * SE + e] = group(g * SUM_FAN, min(S.Z, (g + 1) * SUM_FAN), 0.0f, true);
    __syncthreads();
    if (t != 0 || !ok) return;
    int Zc = G1;
    while (Zc > SUM_FAN) {
        const int Zn = (Zc + SUM_FAN - 1) / SUM_FAN;
        for (int h = 0; h < Zn; ++h) {            // group h reads entries 16h.. and writes entry h <= 16h
            const int z0 = h * SUM_FAN, z1 = min(Zc, z0 + SUM_FAN);
            float s = lv[z0 * SE + e];
            for (int z = z0 + 1; z < z1; ++z) s += lv[z * SE + e];
            lv[h * SE + e] = s;
        }
        Zc = Zn;
    }
    float s = S.out[i];
    for (int z = 0; z < Zc; ++z) s += lv[z * SE + e];
    S.out[i] = s;
}

// First convolution (one input channel): y[p][co] = relu(b[co] + sum_tap x[p+tap] w[co][tap]).
__global__ void k_conv_first(const uint16_t* __restrict__ x, const float* __restrict__ w, const float* __restrict__ b,
                             uint16_t* __restrict__ y, Geo g, int Co) {
    const int P = g.B * g.H * g.W;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)P * Co) return;
    const int p = (int)(idx / Co), co = (int)(idx - (long)p * Co);
    int pb, py, px;
    pixel(p, g, pb, py, px);
    float s = 0.0f;
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
        if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W)
            s += b2f(x[((long)pb * g.H + yy) * g.W + xx]) * b2f(f2b(w[co * 9 + tap]));
    }
    y[idx] = f2b(fmaxf(s + b[co], 0.0f));
}

// First convolution, backward: gw[co][tap] and gb[co] (block partial sums + atomics), gx (input gradient).
__global__ void k_conv_first_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ gz,
                                   float* __restrict__ gw, float* __restrict__ gb, Geo g, int Co, int chunk) {
    // one workgroup per (co, pixel chunk): threads reduce over pixels
    __shared__ float red[10][256];
    const int co = blockIdx.y;
    const int P = g.B * g.H * g.W;
    const int pstart = blockIdx.x * chunk, pend = min(P, pstart + chunk);
    float s[10];
    for (int q = 0; q < 10; ++q) s[q] = 0.0f;
    for (int p = pstart + threadIdx.x; p < pend; p += blockDim.x) {
        const float gv = b2f(gz[(long)p * Co + co]);
        int pb, py, px;
        pixel(p, g, pb, py, px);
        for (int tap = 0; tap < 9; ++tap) {
            const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
            if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W) s[tap] += gv * b2f(x[((long)pb * g.H + yy) * g.W + xx]);
        }
        s[9] += gv;
    }
    for (int q = 0; q < 10; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < 10; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 9) atomicAdd(gw + co * 9 + threadIdx.x, red[threadIdx.x][0]);
    if (threadIdx.x == 9) atomicAdd(gb + co, red[9][0]);
}

// First convolution, one pixel per thread, all Co channels (Co % 8 == 0, Co <= 256): the 9 input taps are
// read once, the bf16-rounded weights come from LDS (same address across the wave: broadcast), and the
// outputs go out as 16-byte rows. Same sum order as k_conv_first (taps 0..8, out-of-range taps add nothing).
__global__ __launch_bounds__(256) void k_conv_first8(const uint16_t* __restrict__ x, const float* __restrict__ w,
                                                     const float* __restrict__ b, uint16_t* __restrict__ y, Geo g,
                                                     int Co) {
    __shared__ float ws[9 * 256];
    __shared__ float bs[256];
    for (int i = threadIdx.x; i < 9 * Co; i += blockDim.x) ws[i] = b2f(f2b(w[i]));
    for (int i = threadIdx.x; i < Co; i += blockDim.x) bs[i] = b[i];
    __syncthreads();
    const int P = g.B * g.H * g.W;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int pb, py, px;
    pixel(p, g, pb, py, px);
    float xv[9];
    bool in[9];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
        in[tap] = yy >= 0 && yy < g.H && xx >= 0 && xx < g.W;
        xv[tap] = in[tap] ? b2f(x[((long)pb * g.H + yy) * g.W + xx]) : 0.0f;
    }
    for (int c = 0; c < Co; c += 8) {
        float s[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] = 0.0f;
#pragma unroll
        for (int tap = 0; tap < 9; ++tap)
            if (in[tap])
#pragma unroll
                for (int e = 0; e < 8; ++e) s[e] += xv[tap] * ws[(c + e) * 9 + tap];
        uint4 o;
        uint16_t* oh = reinterpret_cast<uint16_t*>(&o);
#pragma unroll
        for (int e = 0; e < 8; ++e) oh[e] = f2b(fmaxf(s[e] + bs[c + e], 0.0f));
        *reinterpret_cast<uint4*>(y + (long)p * Co + c) = o;
    }
}

// First convolution, weight and bias gradients, 8 channels per thread: G = Co / 8 channel groups (a power of
// two, at most 16), R = 256 / G pixel rows per pass. Each thread sums gz * x(tap) for its channels over every
// R-th pixel of the workgroup's chunk; lanes of one wave with the same channel group are added by butterfly
// shuffles, the 4 waves in LDS, and one atomic per weight and workgroup goes to gw / gb.
__global__ __launch_bounds__(256) void k_conv_first_wgrad8(const uint16_t* __restrict__ x, const uint16_t* __restrict__ gz,
                                                           float* __restrict__ gw, float* __restrict__ gb,
                                                           float* __restrict__ part, Geo g, int Co, long chunk) {
    __shared__ float red[4][16][80];
    const int G = Co / 8, R = 256 / G;
    const int t = threadIdx.x, gi = t % G, r = t / G;
    const int c = 8 * gi;
    const long P = (long)g.B * g.H * g.W;
    const long p0 = (long)blockIdx.x * chunk, p1 = min(P, p0 + chunk);
    float acc[10][8];
#pragma unroll
    for (int q = 0; q < 10; ++q)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[q][e] = 0.0f;
    for (long p = p0 + r; p < p1; p += R) {
        const uint4 v = *reinterpret_cast<const uint4*>(gz + p * Co + c);
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
        float gv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) gv[e] = b2f(h[e]);
        int pb, py, px;
        pixel((int)p, g, pb, py, px);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
            const float xv = (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W) ? b2f(x[((long)pb * g.H + yy) * g.W + xx]) : 0.0f;
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[tap][e] += gv[e] * xv;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[9][e] += gv[e];
    }
    // lanes l and l ^ off (off a multiple of G) hold the same channel group
    for (int off = G; off < 64; off <<= 1)
#pragma unroll
        for (int q = 0; q < 10; ++q)
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[q][e] += __shfl_xor(acc[q][e], off);
    const int wave = t >> 6, lane = t & 63;
    if (lane < G)
#pragma unroll
        for (int q = 0; q < 10; ++q)
#pragma unroll
            for (int e = 0; e < 8; ++e) red[wave][lane][q * 8 + e] = acc[q][e];
    __syncthreads();
    for (int k = t; k < G * 80; k += 256) {
        const int gg = k / 80, qe = k % 80, q = qe / 8, e = qe % 8;
        const float sum = red[0][gg][qe] + red[1][gg][qe] + red[2][gg][qe] + red[3][gg][qe];
        // partials laid out as [workgroup][gw (Co x 9) | gb (Co)]
        const long o = q < 9 ? (long)(8 * gg + e) * 9 + q : 9L * Co + 8 * gg + e;
        if (part)
            part[(long)blockIdx.x * 10 * Co + o] = sum;
        else if (q < 9)
            atomicAdd(gw + o, sum);
        else
            atomicAdd(gb + 8 * gg + e, sum);
    }
}

// k_conv_first8 on chunks of RB whole rows of one image (blockIdx.x): the rows y0-1 .. y0+RB of x with a zero
// column on each side are staged in LDS ([RB + 2][W + 2] floats, dynamic); thread t takes the 8 channels
// 8 (t % G) .. of every (256 / G)-th pixel, with their 72 BF16-rounded weights in registers. Each channel adds
// its nine taps in tap order, as k_conv_first8 does; a zero-padded tap adds a zero product to a sum that
// starts at +0, which leaves it as k_conv_first8's skipped tap does, so the outputs are the same bits.
__global__ __launch_bounds__(256) void k_conv_first_rows(const uint16_t* __restrict__ x, const float* __restrict__ w,
                                                         const float* __restrict__ b, uint16_t* __restrict__ y,
                                                         Geo g, int Co, int RB) {
    extern __shared__ float xs[];
    const int G = Co / 8, R = 256 / G;
    const int t = threadIdx.x, gi = t % G, r = t / G;
    const int c = 8 * gi;
    const int W = g.W, WP = W + 2;
    const int row0 = blockIdx.x * RB;
    const int bi = row0 / g.H, y0 = row0 - bi * g.H;
    for (int q = t; q < (RB + 2) * WP; q += 256) {
        const int ry = q / WP, rx = q - ry * WP;
        const int yy = y0 + ry - 1, xx = rx - 1;
        xs[q] = (yy >= 0 && yy < g.H && xx >= 0 && xx < W) ? b2f(x[((long)bi * g.H + yy) * W + xx]) : 0.0f;
    }
    float wv[8][9], bv[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) wv[e][tap] = b2f(f2b(w[(c + e) * 9 + tap]));
        bv[e] = b[c + e];
    }
    __syncthreads();
    uint16_t* __restrict__ yc = y + (long)row0 * W * Co + c;
    int ly = r / W, lx = r - (r / W) * W;
    for (int q = r; q < RB * W; q += R) {
        const float* xw = xs + ly * WP + lx;
        float xv[9];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) xv[tap] = xw[(tap / 3) * WP + tap % 3];
        uint4 o;
        uint16_t* oh = reinterpret_cast<uint16_t*>(&o);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float s = 0.0f;
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) s += xv[tap] * wv[e][tap];
            oh[e] = f2b(fmaxf(s + bv[e], 0.0f));
        }
        *reinterpret_cast<uint4*>(yc + (long)q * Co) = o;
        lx += R;
        while (lx >= W) {
            lx -= W;
            ++ly;
        }
    }
}

// k_conv_first_wgrad8 for a chunk of RB whole rows of one image: the rows y0-1 .. y0+RB of x, with a zero
// column on each side, are staged in LDS once ([RB + 2][W + 2] floats, dynamic), so each pixel's nine taps
// are LDS reads without bounds tests, and each thread steps through its pixels by row and column instead of
// dividing. Partials as k_conv_first_wgrad8 (the same workgroups, layout and reduction).
__global__ __launch_bounds__(256) void k_conv_first_wgrad_rows(const uint16_t* __restrict__ x,
                                                               const uint16_t* __restrict__ gz,
                                                               float* __restrict__ gw, float* __restrict__ gb,
                                                               float* __restrict__ part, Geo g, int Co, int RB) {
    extern __shared__ float xs[];
    __shared__ float red[4][16][80];
    const int G = Co / 8, R = 256 / G;
    const int t = threadIdx.x, gi = t % G, r = t / G;
    const int c = 8 * gi;
    const int W = g.W, WP = W + 2;
    const int row0 = blockIdx.x * RB;           // flattened image row b * H + y of the chunk's first row
    const int b = row0 / g.H, y0 = row0 - b * g.H;
    for (int q = t; q < (RB + 2) * WP; q += 256) {
        const int ry = q / WP, rx = q - ry * WP;
        const int yy = y0 + ry - 1, xx = rx - 1;
        xs[q] = (yy >= 0 && yy < g.H && xx >= 0 && xx < W) ? b2f(x[((long)b * g.H + yy) * W + xx]) : 0.0f;
    }
    __syncthreads();
    float acc[10][8];
#pragma unroll
    for (int q = 0; q < 10; ++q)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[q][e] = 0.0f;
    const uint16_t* __restrict__ gzc = gz + (long)row0 * W * Co + c;
    int ly = r / W, lx = r - (r / W) * W;
    for (int q = r; q < RB * W; q += R) {
        const uint4 v = *reinterpret_cast<const uint4*>(gzc + (long)q * Co);
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
        float gv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) gv[e] = b2f(h[e]);
        const float* xw = xs + ly * WP + lx;    // tap (ky, kx) of this pixel is xw[ky * WP + kx]
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const float xv = xw[(tap / 3) * WP + tap % 3];
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[tap][e] += gv[e] * xv;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[9][e] += gv[e];
        lx += R;
        while (lx >= W) {
            lx -= W;
            ++ly;
        }
    }
    for (int off = G; off < 64; off <<= 1)
#pragma unroll
        for (int q = 0; q < 10; ++q)
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[q][e] += __shfl_xor(acc[q][e], off);
    const int wave = t >> 6, lane = t & 63;
    if (lane < G)
#pragma unroll
        for (int q = 0; q < 10; ++q)
#pragma unroll
            for (int e = 0; e < 8; ++e) red[wave][lane][q * 8 + e] = acc[q][e];
    __syncthreads();
    for (int k = t; k < G * 80; k += 256) {
        const int gg = k / 80, qe = k % 80, q = qe / 8, e = qe % 8;
        const float sum = red[0][gg][qe] + red[1][gg][qe] + red[2][gg][qe] + red[3][gg][qe];
        const long o = q < 9 ? (long)(8 * gg + e) * 9 + q : 9L * Co + 8 * gg + e;
        if (part)
            part[(long)blockIdx.x * 10 * Co + o] = sum;
        else if (q < 9)
            atomicAdd(gw + o, sum);
        else
            atomicAdd(gb + 8 * gg + e, sum);
    }
}

__global__ void k_conv_first_dgrad(const uint16_t* __restrict__ gz, const float* __restrict__ w, float* __restrict__ gx,
                                   Geo g, int Co) {
    const int P = g.B * g.H * g.W;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int pb, py, px;
    pixel(p, g, pb, py, px);
    float s = 0.0f;
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = py - (tap / 3 - 1), xx = px - (tap % 3 - 1);
        if (yy < 0 || yy >= g.H || xx < 0 || xx >= g.W) continue;
        const uint16_t* gp = gz + (((long)pb * g.H + yy) * g.W + xx) * Co;
        for (int co = 0; co < Co; ++co) s += b2f(gp[co]) * b2f(f2b(w[co * 9 + tap]));
    }
    gx[p] = s;
}

// k_conv_first_dgrad with 16-byte loads of each tap's channels (Co % 8 == 0, Co <= 64) and the weights rounded
// to bf16 once, in LDS: the same products added in the same order (tap, then channel), so the same bits.
__global__ __launch_bounds__(256) void k_conv_first_dgrad8(const uint16_t* __restrict__ gz, const float* __restrict__ w,
                                                           float* __restrict__ gx, Geo g, int Co) {
    __shared__ float ws[9][64];
    for (int q = threadIdx.x; q < 9 * Co; q += blockDim.x) {
        const int co = q / 9, tap = q - 9 * co;
        ws[tap][co] = b2f(f2b(w[q]));
    }
    __syncthreads();
    const int P = g.B * g.H * g.W;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int pb, py, px;
    pixel(p, g, pb, py, px);
    float s = 0.0f;
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = py - (tap / 3 - 1), xx = px - (tap % 3 - 1);
        if (yy < 0 || yy >= g.H || xx < 0 || xx >= g.W) continue;
        const uint4* gp = reinterpret_cast<const uint4*>(gz + (((long)pb * g.H + yy) * g.W + xx) * Co);
        for (int c8 = 0; c8 < Co / 8; ++c8) {
            const uint4 v = gp[c8];
            const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int e = 0; e < 8; ++e) s += b2f(h[e]) * ws[tap][8 * c8 + e];
        }
    }
    gx[p] = s;
}

// Column sums (bias gradient): gb[c] += sum_p g[p][c] over the workgroup's `chunk` pixel rows. Thread t
// reads channels 8*(t % G) .. +7 (G = ceil(C/8) groups, 16-byte loads when C % 8 == 0) of every R-th row
// (R = 256 / G), the R partial sums of each channel are added in LDS, and one atomic per channel and
// workgroup goes to gb. C <= 2048.
__global__ __launch_bounds__(256) void k_colsum(const uint16_t* __restrict__ gz, float* __restrict__ gb,
                                                float* __restrict__ part, long P, int C, long chunk) {
    __shared__ float red[2048];
    const int G = (C + 7) / 8, R = 256 / G;
    const int t = threadIdx.x, gi = t % G, r = t / G;
    const int c0 = 8 * gi;
    float s[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) s[e] = 0.0f;
    if (r < R) {
        const long p0 = (long)blockIdx.x * chunk, p1 = min(P, p0 + chunk);
        if ((C & 7) == 0) {
            for (long p = p0 + r; p < p1; p += R) {
                const uint4 v = *reinterpret_cast<const uint4*>(gz + p * C + c0);
                const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                for (int e = 0; e < 8; ++e) s[e] += b2f(h[e]);
            }
        } else {
            for (long p = p0 + r; p < p1; p += R)
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (c0 + e < C) s[e] += b2f(gz[p * C + c0 + e]);
        }
    }
    // every slot red[q * 8G + c] with q < R is written (unused channels c >= C hold 0); the sums read q < R only
    if (r < R)
#pragma unroll
        for (int e = 0; e < 8; ++e) red[r * 8 * G + c0 + e] = s[e];
    __syncthreads();
    for (int c = t; c < C; c += 256) {
        float a = 0.0f;
        for (int q = 0; q < R; ++q) a += red[q * 8 * G + c];
        if (part)
            part[(long)blockIdx.x * C + c] = a;
        else
            atomicAdd(gb + c, a);
    }
}

// 8 consecutive elements per thread (n % 8 == 0): zero the gradient where the ReLU output is not positive.
__global__ void k_relu_mask8(uint16_t* __restrict__ gz, const uint16_t* __restrict__ y, long n8) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    uint4 g = reinterpret_cast<const uint4*>(gz)[i];
    const uint4 v = reinterpret_cast<const uint4*>(y)[i];
    uint16_t* gh = reinterpret_cast<uint16_t*>(&g);
    const uint16_t* vh = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
    for (int e = 0; e < 8; ++e)
        if (!(b2f(vh[e]) > 0.0f)) gh[e] = 0;
    reinterpret_cast<uint4*>(gz)[i] = g;
}

__global__ void k_relu_mask(uint16_t* __restrict__ gz, const uint16_t* __restrict__ y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(b2f(y[i]) > 0.0f)) gz[i] = 0;
}

// 2x2 average pooling and its backward, 8 channels per thread (C % 8 == 0).
// Thread x of a row handles pixel x / C8 and channels 8 (x % C8) .. +7; rows (b, y) come from blockIdx.y in
// strides of gridDim.y, so the row's image and line are uniform and no thread divides 64-bit indices.
__global__ void k_avgpool2_8(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, Geo g, int C) {
    const int Ho = g.H / 2, Wo = g.W / 2, C8 = C / 8;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Wo * C8) return;
    const int xo = t / C8, c = (t - xo * C8) * 8;
    for (int row = blockIdx.y; row < g.B * Ho; row += gridDim.y) {
        const int b = row / Ho, yo = row - b * Ho;
        const long base = (((long)b * g.H + 2 * yo) * g.W + 2 * xo) * C + c;
        const uint4 v0 = *reinterpret_cast<const uint4*>(x + base);
        const uint4 v1 = *reinterpret_cast<const uint4*>(x + base + C);
        const uint4 v2 = *reinterpret_cast<const uint4*>(x + base + (long)g.W * C);
        const uint4 v3 = *reinterpret_cast<const uint4*>(x + base + (long)g.W * C + C);
        const uint16_t* h0 = reinterpret_cast<const uint16_t*>(&v0);
        const uint16_t* h1 = reinterpret_cast<const uint16_t*>(&v1);
        const uint16_t* h2 = reinterpret_cast<const uint16_t*>(&v2);
        const uint16_t* h3 = reinterpret_cast<const uint16_t*>(&v3);
        uint4 o;
        uint16_t* oh = reinterpret_cast<uint16_t*>(&o);
#pragma unroll
        for (int e = 0; e < 8; ++e) oh[e] = f2b(0.25f * (b2f(h0[e]) + b2f(h1[e]) + b2f(h2[e]) + b2f(h3[e])));
        *reinterpret_cast<uint4*>(y + ((long)row * Wo + xo) * C + c) = o;
    }
}

// Backward of the pooling, on pooled rows as in k_avgpool2_8. RELU: 0 where y (the pooled layer's own
// output, after its ReLU) is not positive, the ReLU derivative of that layer fused; the same bits as the two
// passes.
template <bool RELU>
__global__ void k_avgpool2_bwd8(const uint16_t* __restrict__ gy, const uint16_t* __restrict__ y,
                                uint16_t* __restrict__ gx, Geo g, int C) {
    // one pooled row per grid row: its gradient chunk is loaded once for the two full-resolution rows below it
    const int Ho = g.H / 2, Wo = g.W / 2, C8 = C / 8;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.W * C8) return;
    const int xx = t / C8, c = (t - xx * C8) * 8;
    for (int prow = blockIdx.y; prow < g.B * Ho; prow += gridDim.y) {
        const int b = prow / Ho, yo = prow - b * Ho;
        const uint4 v = *reinterpret_cast<const uint4*>(gy + ((long)prow * Wo + xx / 2) * C + c);
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
        const long q0 = ((long)b * g.H + 2 * yo) * g.W + xx, q1 = q0 + g.W;
        uint4 o0, o1;
        uint16_t* oh0 = reinterpret_cast<uint16_t*>(&o0);
        uint16_t* oh1 = reinterpret_cast<uint16_t*>(&o1);
        if (RELU) {
            const uint4 m0 = *reinterpret_cast<const uint4*>(y + q0 * C + c);
            const uint4 m1 = *reinterpret_cast<const uint4*>(y + q1 * C + c);
            const uint16_t* mh0 = reinterpret_cast<const uint16_t*>(&m0);
            const uint16_t* mh1 = reinterpret_cast<const uint16_t*>(&m1);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint16_t q = f2b(0.25f * b2f(h[e]));
                oh0[e] = b2f(mh0[e]) > 0.0f ? q : (uint16_t)0;
                oh1[e] = b2f(mh1[e]) > 0.0f ? q : (uint16_t)0;
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) oh0[e] = oh1[e] = f2b(0.25f * b2f(h[e]));
        }
        *reinterpret_cast<uint4*>(gx + q0 * C + c) = o0;
        *reinterpret_cast<uint4*>(gx + q1 * C + c) = o1;
    }
}

__global__ void k_avgpool2(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, Geo g, int C) {
    const int Ho = g.H / 2, Wo = g.W / 2;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.B * Ho * Wo * C) return;
    const int c = (int)(idx % C);
    const long q = idx / C;
    const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), b = (int)(q / ((long)Wo * Ho));
    const long base = (((long)b * g.H + 2 * yo) * g.W + 2 * xo) * C + c;
    const float s = b2f(x[base]) + b2f(x[base + C]) + b2f(x[base + (long)g.W * C]) + b2f(x[base + (long)g.W * C + C]);
    y[idx] = f2b(0.25f * s);
}

// gx (H x W) from gy (H/2 x W/2): every input cell receives a quarter of its pool's gradient.
__global__ void k_avgpool2_bwd(const uint16_t* __restrict__ gy, uint16_t* __restrict__ gx, Geo g, int C) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.B * g.H * g.W * C) return;
    const int c = (int)(idx % C);
    const long q = idx / C;
    const int xx = (int)(q % g.W), yy = (int)((q / g.W) % g.H), b = (int)(q / ((long)g.W * g.H));
    const int Ho = g.H / 2, Wo = g.W / 2;
    gx[idx] = f2b(0.25f * b2f(gy[(((long)b * Ho + yy / 2) * Wo + xx / 2) * C + c]));
}

// Head: gap[b][c] = mean_p a[b][p][c]; out[b][j] = fb[j] + sum_c fw[j][c] gap[b][c]. One workgroup per sample.
__global__ void k_head(const uint16_t* __restrict__ a, const float* __restrict__ fw, const float* __restrict__ fb,
                       float* __restrict__ gap, float* __restrict__ out, int HW, int C, int nout) {
    extern __shared__ float sg[];
    const int b = blockIdx.x;
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        float s = 0.0f;
        for (int p = 0; p < HW; ++p) s += b2f(a[((long)b * HW + p) * C + c]);
        sg[c] = s / (float)HW;
        gap[(long)b * C + c] = sg[c];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nout; j += blockDim.x) {
        float s = fb[j];
        for (int c = 0; c < C; ++c) s += fw[j * C + c] * sg[c];
        out[(long)b * nout + j] = s;
    }
}

// The head with 8 channels per thread: a workgroup of 256 threads per sample splits its pixels into S = 256 / (C / 8)
// interleaved slices (thread t: channels 8 (t % (C / 8)) .. +7 by 16-byte loads, pixels t / (C / 8) + S k); the
// slices' sums are added in slice order in LDS (deterministic). C % 8 == 0 and 256 % (C / 8) == 0.
__global__ __launch_bounds__(256) void k_head8(const uint16_t* __restrict__ a, const float* __restrict__ fw,
                                               const float* __restrict__ fb, float* __restrict__ gap,
                                               float* __restrict__ out, int HW, int C, int nout) {
    extern __shared__ float sh[];
    const int G = C / 8, S = 256 / G;
    float* red = sh;             // [S][C]
    float* sg = sh + S * C;      // [C]
    const int b = blockIdx.x, t = threadIdx.x, cg = t % G, sl = t / G;
    float acc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = 0.0f;
    const uint16_t* ab = a + (long)b * HW * C + 8 * cg;
#pragma unroll 8
    for (int p = sl; p < HW; p += S) {     // unrolled: the loads of 8 pixels in flight, added in pixel order
        const uint4 v = *reinterpret_cast<const uint4*>(ab + (long)p * C);
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[e] += b2f(h[e]);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) red[sl * C + 8 * cg + e] = acc[e];
    __syncthreads();
    for (int c = t; c < C; c += 256) {
        float v = 0.0f;
        for (int q = 0; q < S; ++q) v += red[q * C + c];
        sg[c] = v / (float)HW;
        gap[(long)b * C + c] = sg[c];
    }
    __syncthreads();
    for (int j = t; j < nout; j += 256) {
        float v = fb[j];
        for (int c = 0; c < C; ++c) v += fw[j * C + c] * sg[c];
        out[(long)b * nout + j] = v;
    }
}

// Head backward: gfw += gout^T gap, gfb += sum gout, ga[b][p][c] = (gout[b] . fw[:, c]) / HW.
__global__ void k_head_bwd(const float* __restrict__ gout, const float* __restrict__ gap, const float* __restrict__ fw,
                           float* __restrict__ gfw, float* __restrict__ gfb, uint16_t* __restrict__ ga, int HW, int C,
                           int nout) {
    const int b = blockIdx.x;
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        float s = 0.0f;
        for (int j = 0; j < nout; ++j) s += gout[(long)b * nout + j] * fw[j * C + c];
        const uint16_t v = f2b(s / (float)HW);
        for (int p = 0; p < HW; ++p) ga[((long)b * HW + p) * C + c] = v;
    }
}

// Head parameter gradients, summed over the samples in order: gfw[j][c] += sum_b gout[b][j] gap[b][c],
// gfb[j] += sum_b gout[b][j]. One thread per output (nout x C weights, then nout biases).
__global__ void k_head_wgrad(const float* __restrict__ gout, const float* __restrict__ gap, float* __restrict__ gfw,
                             float* __restrict__ gfb, int B, int C, int nout) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (long)nout * C) {
        const int j = (int)(i / C), c = (int)(i % C);
        float s = gfw[i];
#pragma unroll 16
        for (int b = 0; b < B; ++b) s += gout[(long)b * nout + j] * gap[(long)b * C + c];   // loads in flight together
        gfw[i] = s;
    } else if (i < (long)nout * C + nout) {
        const int j = (int)(i - (long)nout * C);
        float s = gfb[j];
#pragma unroll 16
        for (int b = 0; b < B; ++b) s += gout[(long)b * nout + j];
        gfb[j] = s;
    }
}

// bf16 copies of FP32 weights [Co][9][Ci]: w16 in the same order, wt16 with the channel axes swapped,
// [Ci][9][Co] (same tap order).
__device__ __forceinline__ void weight_copy(const float* __restrict__ w, uint16_t* __restrict__ w16,
                                            uint16_t* __restrict__ wt16, int Co, int Ci, int idx) {
    // 32-bit index arithmetic: a layer's Co * 9 * Ci weights are far below 2^31
    if (idx >= Co * 9 * Ci) return;
    const int q = idx / Ci, ci = idx - q * Ci;
    const int co = q / 9, tap = q - co * 9;
    const uint16_t v = f2b(w[idx]);
    w16[idx] = v;
    if (wt16) wt16[((long)ci * 9 + tap) * Co + co] = v;   // the dgrad GEMM mirrors the tap offset itself
}

__global__ void k_weights(const float* __restrict__ w, uint16_t* __restrict__ w16, uint16_t* __restrict__ wt16, int Co,
                          int Ci) {
    weight_copy(w, w16, wt16, Co, Ci, blockIdx.x * blockDim.x + threadIdx.x);
}

// every layer's copies in one launch: layer j takes blocks start[j] .. start[j + 1] - 1
constexpr int WMAX = 16;
struct WCopies {
    const float* w[WMAX];
    uint16_t* w16[WMAX];
    uint16_t* wt16[WMAX];
    int co[WMAX], ci[WMAX], start[WMAX + 1];
    int n;
};

__global__ void k_weights_n(const WCopies c) {
    int j = 0;
    while (j + 1 < c.n && (int)blockIdx.x >= c.start[j + 1]) ++j;
    weight_copy(c.w[j], c.w16[j], c.wt16[j], c.co[j], c.ci[j], (blockIdx.x - c.start[j]) * blockDim.x + threadIdx.x);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_f2b(const float* __restrict__ x, uint16_t* __restrict__ y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2b(x[i]);
}

inline dim3 grid1(long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }
// (threads of one row, rows): x covers a row, y strides over the rows (at most 65535 workgroups in y)
inline dim3 rows_grid(long per_row, long rows) {
    return dim3((unsigned)((per_row + 255) / 256), (unsigned)(rows < 65535 ? rows : 65535));
}

}  // namespace

// Row-march convolution for rows of 64 or 128 pixels: input channels 32 (or 64 for 64-pixel rows), output
// channels in blocks of 32 (64 where 32 input channels meet 64-pixel rows); bands of at least 16 rows, about
// 1024 workgroups. Returns false where no shape applies (the caller takes the tiled GEMM).
static bool conv_march(const uint16_t* src, const uint16_t* wk, const float* bias, uint16_t* out, const uint16_t* mask,
                       int B, int H, int W, int Cin, int Cout, int sign, int relu, hipStream_t st) {
    int cb;
    if (W == 128 && Cin == 32 && Cout % 32 == 0) cb = 32;
    else if (W == 64 && Cin == 32 && Cout % 64 == 0) cb = 64;
    else if (W == 64 && (Cin == 32 || Cin == 64) && Cout % 32 == 0) cb = 32;
    else return false;
    const long rows = (long)B * H;
    const long nblk = Cout / cb;
    long nz = std::max(1L, std::min(rows / 16, 1024 / nblk));
    const int band = (int)((rows + nz - 1) / nz);
    nz = (rows + band - 1) / band;
    const dim3 grid(1, (unsigned)nblk, (unsigned)nz);
    const int rt = (int)rows;
    if (mask) {
        if (W == 128)
            hipLaunchKernelGGL((k_conv_march<128, 32, 32, true>), grid, dim3(256), 0, st, src, wk, bias, out, mask, H, rt,
                               Cout, sign, relu, band);
        else if (Cin == 32 && cb == 64)
            hipLaunchKernelGGL((k_conv_march<64, 32, 64, true>), grid, dim3(256), 0, st, src, wk, bias, out, mask, H, rt,
                               Cout, sign, relu, band);
        else if (Cin == 32)
            hipLaunchKernelGGL((k_conv_march<64, 32, 32, true>), grid, dim3(128), 0, st, src, wk, bias, out, mask, H, rt,
                               Cout, sign, relu, band);
        else
            hipLaunchKernelGGL((k_conv_march<64, 64, 32, true>), grid, dim3(128), 0, st, src, wk, bias, out, mask, H, rt,
                               Cout, sign, relu, band);
        return true;
    }
    if (W == 128)
        hipLaunchKernelGGL((k_conv_march<128, 32, 32, false>), grid, dim3(256), 0, st, src, wk, bias, out, mask, H, rt, Cout, sign,
                           relu, band);
    else if (Cin == 32 && cb == 64)
        hipLaunchKernelGGL((k_conv_march<64, 32, 64, false>), grid, dim3(256), 0, st, src, wk, bias, out, mask, H, rt, Cout, sign,
                           relu, band);
    else if (Cin == 32)
        hipLaunchKernelGGL((k_conv_march<64, 32, 32, false>), grid, dim3(128), 0, st, src, wk, bias, out, mask, H, rt, Cout, sign,
                           relu, band);
    else
        hipLaunchKernelGGL((k_conv_march<64, 64, 32, false>), grid, dim3(128), 0, st, src, wk, bias, out, mask, H, rt, Cout, sign,
                           relu, band);
    return true;
}

extern "C" int bz_nn_conv3x3(const uint16_t* x, const uint16_t* w16, const float* bias, uint16_t* y, int B, int H,
                             int W, int Ci, int Co, int relu, void* stream) {
    BZ_REQUIRE(x && w16 && y && B > 0 && H > 0 && W > 0, "bz_nn_conv3x3: bad arguments");
    BZ_REQUIRE(Ci % 32 == 0 && Co % 32 == 0, "bz_nn_conv3x3: Ci must be a multiple of 32 and Co of 32");
    const long P = (long)B * H * W;
    BZ_REQUIRE(P * (long)(Ci > Co ? Ci : Co) < (1L << 31), "bz_nn_conv3x3: tensor too large");
    Geo g{B, H, W};
    if (conv_march(x, w16, bias, y, nullptr, B, H, W, Ci, Co, 1, relu, (hipStream_t)stream))
        ;
    else if (halo_ok(g) && Ci == 32)
        hipLaunchKernelGGL(k_conv_halo<32>, dim3((unsigned)(P / 64), (Co + TN - 1) / TN), dim3(256), 0,
                           (hipStream_t)stream, x, w16, bias, y, (const uint16_t*)nullptr, g, Ci, Co, 1, relu);
    else
        hipLaunchKernelGGL((Ci % 64 == 0 ? k_conv_gemm<64, true>
                                         : ((9 * Ci) % 64 == 0 ? k_conv_gemm<64, false> : k_conv_gemm<32, false>)),
                           dim3((unsigned)((P + TM - 1) / TM), (Co + TN - 1) / TN), dim3(256), 0, (hipStream_t)stream,
                           x, w16, bias, y, (const uint16_t*)nullptr, g, Ci, Co, 1, relu);
    BZ_CHECK_LAUNCH("bz_nn_conv3x3");
    return BZ_OK;
}

static int conv_dgrad(const uint16_t* gz, const uint16_t* wt16, const uint16_t* mask, uint16_t* gx, int B, int H,
                      int W, int Ci, int Co, void* stream, const char* name) {
    if (!(gz && wt16 && gx && B > 0 && H > 0 && W > 0)) {
        bz::set_error(std::string(name) + ": bad arguments");
        return BZ_EINVAL;
    }
    if (Co % 32 || Ci % 32) {
        bz::set_error(std::string(name) + ": Co must be a multiple of 32 and Ci of 32");
        return BZ_EINVAL;
    }
    const long P = (long)B * H * W;
    if (P * (long)(Ci > Co ? Ci : Co) >= (1L << 31)) {
        bz::set_error(std::string(name) + ": tensor too large");
        return BZ_EINVAL;
    }
    Geo g{B, H, W};
    if (conv_march(gz, wt16, nullptr, gx, mask, B, H, W, Co, Ci, -1, 0, (hipStream_t)stream))
        ;
    else if (halo_ok(g) && Co == 32)
        hipLaunchKernelGGL(k_conv_halo<32>, dim3((unsigned)(P / 64), (Ci + TN - 1) / TN), dim3(256), 0,
                           (hipStream_t)stream, gz, wt16, (const float*)nullptr, gx, mask, g, Co, Ci, -1, 0);
    else
        hipLaunchKernelGGL((Co % 64 == 0 ? k_conv_gemm<64, true>
                                         : ((9 * Co) % 64 == 0 ? k_conv_gemm<64, false> : k_conv_gemm<32, false>)),
                           dim3((unsigned)((P + TM - 1) / TM), (Ci + TN - 1) / TN), dim3(256), 0, (hipStream_t)stream,
                           gz, wt16, (const float*)nullptr, gx, mask, g, Co, Ci, -1, 0);
    BZ_CHECK_LAUNCH(name);
    return BZ_OK;
}

extern "C" int bz_nn_conv3x3_dgrad(const uint16_t* gz, const uint16_t* wt16, uint16_t* gx, int B, int H, int W, int Ci,
                                   int Co, void* stream) {
    return conv_dgrad(gz, wt16, nullptr, gx, B, H, W, Ci, Co, stream, "bz_nn_conv3x3_dgrad");
}

extern "C" int bz_nn_conv3x3_dgrad_relu(const uint16_t* gz, const uint16_t* wt16, const uint16_t* x, uint16_t* gx, int B,
                                        int H, int W, int Ci, int Co, void* stream) {
    if (!x) {
        bz::set_error("bz_nn_conv3x3_dgrad_relu: null mask");
        return BZ_EINVAL;
    }
    return conv_dgrad(gz, wt16, x, gx, B, H, W, Ci, Co, stream, "bz_nn_conv3x3_dgrad_relu");
}

// Launch geometry of bz_nn_conv3x3_wgrad: pixel chunk and number of splits of the weight GEMM, and pixel
// chunk and workgroup count of the bias column sums (about 2048 workgroups of at least 16 rows per thread).
struct WgradPlan {
    int chunk, nz;
    long cchunk, nblk;
};
inline WgradPlan wgrad_plan(long P, int Co, int splits) {
    WgradPlan q;
    q.chunk = (int)(((P + splits - 1) / splits + WK - 1) / WK * WK);
    q.nz = (int)((P + q.chunk - 1) / q.chunk);
    const long rows = 256 / ((Co + 7) / 8);
    q.cchunk = std::max(rows * 16, (P + 2047) / 2048);
    q.nblk = (P + q.cchunk - 1) / q.cchunk;
    return q;
}

// Rows of the extra region the tree needs for Z partial rows (the first level's output, the largest).
inline long tree_rows(long Z) { return Z > SUM_FAN ? (Z + SUM_FAN - 1) / SUM_FAN : 0; }

// out[0 .. count) += the Z rows of part (rows `stride` apart) in a fixed tree order; scratch holds
// tree_rows(Z) rows of `count` floats. Levels alternate between scratch and the head of part, which the
// previous level has finished reading (stream order).
static void tree_sum(float* part, long Z, long stride, long count, float* scratch, float* out, hipStream_t st) {
    float* cur = part;
    long sc = stride;
    bool in_part = true;
    while (Z > SUM_FAN) {
        const long Zn = (Z + SUM_FAN - 1) / SUM_FAN;
        float* dst = in_part ? scratch : part;
        hipLaunchKernelGGL(k_sum_group, dim3(grid1(count).x, (unsigned)Zn), dim3(256), 0, st, cur, (int)Z, sc, count,
                           dst, count);
        cur = dst;
        sc = count;
        in_part = !in_part;
        Z = Zn;
    }
    hipLaunchKernelGGL(k_sum_parts, grid1(count), dim3(256), 0, st, cur, (int)Z, sc, count, out);
}

// out += the Z rows of each segment's partials, in tree_sum's order, in one launch (k_tree_sum); b.count may be
// 0. First levels too large for one workgroup's LDS go through tree_sum (scratch as there).
static void tree_sum2(const SumSeg& a, const SumSeg& b, float* scratch, hipStream_t st) {
    auto g1 = [](const SumSeg& q) { return q.count > 0 && q.Z > SUM_FAN ? (q.Z + SUM_FAN - 1) / SUM_FAN : 0; };
    const int G1 = std::max(g1(a), g1(b));
    // lanes per element (256 / se) about the number of first-level groups
    const int se = G1 >= 32 ? 8 : (G1 >= 4 ? 32 : (G1 >= 2 ? 128 : 256));
    if ((long)G1 * se > SUM_LDS_FLOATS) {
        tree_sum(a.in, a.Z, a.stride, a.count, scratch, a.out, st);
        if (b.count > 0) tree_sum(b.in, b.Z, b.stride, b.count, scratch, b.out, st);
        return;
    }
    const long na = (a.count + se - 1) / se, nb = b.count > 0 ? (b.count + se - 1) / se : 0;
    const size_t lds = (size_t)G1 * se * sizeof(float);
    const dim3 grid((unsigned)(na + nb));
    switch (se) {
        case 8: hipLaunchKernelGGL(k_tree_sum<8>, grid, dim3(256), lds, st, a, b, (int)na); break;
        case 32: hipLaunchKernelGGL(k_tree_sum<32>, grid, dim3(256), lds, st, a, b, (int)na); break;
        case 128: hipLaunchKernelGGL(k_tree_sum<128>, grid, dim3(256), lds, st, a, b, (int)na); break;
        default: hipLaunchKernelGGL(k_tree_sum<256>, grid, dim3(256), lds, st, a, b, (int)na); break;
    }
}

// The row march (k_wgrad_march) serves rows of 16, 32, 64 or 128 pixels with channel counts in multiples of 32,
// R = max(1, 64 / W) rows per step: 32-wide channel blocks on 128-pixel rows, 64-wide where the counts allow
// otherwise (16- and 32-pixel rows need 64-wide output blocks, 16-pixel rows 64-wide input blocks too; on
// the deep levels they measured faster than 32-wide blocks with fewer bands). The rows are cut into nz bands
// (multiples of R rows) for about 2048 workgroups, at least 16 rows per band, at most 8 M (64- and 128-pixel
// rows) or 16 M partial floats (the ordered sums read them all).
struct MarchPlan {
    bool ok;
    int nco, nci, rows_per_step, nz, band;
};
inline MarchPlan march_plan(int B, int H, int W, int Ci, int Co) {
    MarchPlan m{false, 1, 1, 1, 0, 0};
    if (Ci % 32 || Co % 32 || (W != 16 && W != 32 && W != 64 && W != 128)) return m;
    if (W <= 32 && Co % 64) return m;
    if (W == 16 && Ci % 64) return m;
    const int nco = (W <= 64 && Co % 64 == 0) ? 2 : 1;
    const int nci = (W <= 64 && Ci % 64 == 0) ? 2 : 1;
    const int R = W >= 64 ? 1 : 64 / W;
    const long rows = (long)B * H;
    const long nblk = (long)(Co / (32 * nco)) * (Ci / (32 * nci));
    const long nw = (long)Co * 9 * Ci;
    long nz = std::max(1L, 2048 / nblk);
    nz = std::min(nz, std::max(1L, rows / 16));
    nz = std::min(nz, std::max(1L, (W >= 64 ? 8L << 20 : 16L << 20) / nw));
    long band = (rows + nz - 1) / nz;
    band = (band + R - 1) / R * R;
    m.ok = true;
    m.nco = nco;
    m.nci = nci;
    m.rows_per_step = R;
    m.band = (int)band;
    m.nz = (int)((rows + band - 1) / band);
    return m;
}

extern "C" long bz_nn_wgrad_work(int B, int H, int W, int Ci, int Co, int splits) {
    if (B <= 0 || H <= 0 || W <= 0 || Ci <= 0 || Co <= 0 || splits <= 0) return -1;
    const long nw = (long)Co * 9 * Ci;
    const MarchPlan mp = march_plan(B, H, W, Ci, Co);
    if (mp.ok) return mp.nz * nw + mp.nz * Co + tree_rows(mp.nz) * nw;
    const WgradPlan q = wgrad_plan((long)B * H * W, Co, splits);
    return q.nz * nw + q.nblk * Co + std::max(tree_rows(q.nz) * nw, tree_rows(q.nblk) * Co);
}

template <int NCO, int NCI, int W>
static void launch_march(const MarchPlan& mp, const uint16_t* x, const uint16_t* gz, float* part, float* bpart,
                         int B, int H, int Ci, int Co, hipStream_t st) {
    constexpr int R = W >= 64 ? 1 : 64 / W;
    hipLaunchKernelGGL((k_wgrad_march<NCO, NCI, W, R>), dim3(Co / (32 * NCO), Ci / (32 * NCI), mp.nz),
                       dim3(192 * NCO * NCI), 0, st, x, gz, part, bpart, H, B * H, Ci, Co, mp.band);
}

extern "C" int bz_nn_conv3x3_wgrad(const uint16_t* x, const uint16_t* gz, float* gw, float* gb, int B, int H, int W,
                                   int Ci, int Co, int splits, float* work, long nwork, void* stream) {
    BZ_REQUIRE(x && gz && gw && gb && B > 0 && H > 0 && W > 0 && splits > 0, "bz_nn_conv3x3_wgrad: bad arguments");
    BZ_REQUIRE(Ci % 8 == 0 && Co % 8 == 0, "bz_nn_conv3x3_wgrad: channel counts must be multiples of 8");
    BZ_REQUIRE(Co <= 2048, "bz_nn_conv3x3_wgrad: at most 2048 output channels");
    const long P = (long)B * H * W;
    const WgradPlan q = wgrad_plan(P, Co, splits);
    const long nw = (long)Co * 9 * Ci;
    const long scratch_rows = std::max(tree_rows(q.nz) * nw, tree_rows(q.nblk) * Co);
    Geo g{B, H, W};
    const hipStream_t st = (hipStream_t)stream;
    const MarchPlan mp = march_plan(B, H, W, Ci, Co);
    if (work && mp.ok) {
        BZ_REQUIRE(nwork >= mp.nz * nw + mp.nz * Co + tree_rows(mp.nz) * nw,
                   "bz_nn_conv3x3_wgrad: workspace smaller than bz_nn_wgrad_work");
        float* wpart = work;
        float* bpart = work + (long)mp.nz * nw;
        const int key = W * 4 + (mp.nco - 1) * 2 + (mp.nci - 1);
        switch (key) {
            case 128 * 4: launch_march<1, 1, 128>(mp, x, gz, wpart, bpart, B, H, Ci, Co, st); break;
            case 64 * 4: launch_march<1, 1, 64>(mp, x, gz, wpart, bpart, B, H, Ci, Co, st); break;
            case 64 * 4 + 2: launch_march<2, 1, 64>(mp, x, gz, wpart, bpart, B, H, Ci, Co, st); break;
            case 64 * 4 + 1: launch_march<1, 2, 64>(mp, x, gz, wpart, bpart, B, H, Ci, Co, st); break;
            case 64 * 4 + 3: launch_march<2, 2, 64>(mp, x, gz, wpart, bpart, B, H, Ci, Co, st); break;
            case 32 * 4 + 2: launch_march<2, 1, 32>(mp, x, gz, wpart, bpart, B, H, Ci, Co, st); break;
            case 32 * 4 + 3: launch_march<2, 2, 32>(mp, x, gz, wpart, bpart, B, H, Ci, Co, st); break;
            case 16 * 4 + 3: launch_march<2, 2, 16>(mp, x, gz, wpart, bpart, B, H, Ci, Co, st); break;
            default: bz::set_error("bz_nn_conv3x3_wgrad: no row-march shape"); return BZ_EINVAL;
        }
        BZ_CHECK_LAUNCH("bz_nn_conv3x3_wgrad(march)");
        float* scratch = bpart + (long)mp.nz * Co;
        tree_sum2(SumSeg{wpart, nw, nw, gw, mp.nz}, SumSeg{bpart, Co, Co, gb, mp.nz}, scratch, st);
        BZ_CHECK_LAUNCH("bz_nn_conv3x3_wgrad(sums)");
        return BZ_OK;
    }
    BZ_REQUIRE(!work || nwork >= q.nz * nw + q.nblk * Co + scratch_rows,
               "bz_nn_conv3x3_wgrad: workspace smaller than bz_nn_wgrad_work");
    float* wpart = work;
    float* bpart = work ? work + q.nz * nw : nullptr;
    hipLaunchKernelGGL(k_conv_wgrad, dim3((Co + TM - 1) / TM, (9 * Ci + TN - 1) / TN, q.nz), dim3(256), 0, st, x, gz,
                       gw, wpart, g, Ci, Co, q.chunk);
    BZ_CHECK_LAUNCH("bz_nn_conv3x3_wgrad");
    hipLaunchKernelGGL(k_colsum, dim3((unsigned)q.nblk), dim3(256), 0, st, gz, gb, bpart, P, Co, q.cchunk);
    BZ_CHECK_LAUNCH("bz_nn_conv3x3_wgrad(bias)");
    if (work) {
        float* scratch = bpart + q.nblk * Co;
        tree_sum2(SumSeg{wpart, nw, nw, gw, q.nz}, SumSeg{bpart, Co, Co, gb, (int)q.nblk}, scratch, st);
        BZ_CHECK_LAUNCH("bz_nn_conv3x3_wgrad(sums)");
    }
    return BZ_OK;
}

extern "C" int bz_nn_conv_first(const uint16_t* x, const float* w, const float* b, uint16_t* y, int B, int H, int W,
                                int Co, void* stream) {
    BZ_REQUIRE(x && w && b && y && B > 0 && H > 0 && W > 0 && Co > 0, "bz_nn_conv_first: bad arguments");
    Geo g{B, H, W};
    // chunks of whole rows of one image (about 2048 pixels) take the row-staged kernel
    int rb = std::max(1, std::min(H, 2048 / W));
    while (H % rb) --rb;
    const size_t lds = (size_t)(rb + 2) * (W + 2) * sizeof(float);
    const bool rows_ok = Co % 8 == 0 && 256 % (Co / 8) == 0 && lds <= 32768 && ((uintptr_t)y & 15) == 0;
    if (rows_ok)
        hipLaunchKernelGGL(k_conv_first_rows, dim3((unsigned)((long)B * H / rb)), dim3(256), lds, (hipStream_t)stream,
                           x, w, b, y, g, Co, rb);
    else if (Co % 8 == 0 && Co <= 256 && ((uintptr_t)y & 15) == 0)
        hipLaunchKernelGGL(k_conv_first8, grid1((long)B * H * W), dim3(256), 0, (hipStream_t)stream, x, w, b, y, g, Co);
    else
        hipLaunchKernelGGL(k_conv_first, grid1((long)B * H * W * Co), dim3(256), 0, (hipStream_t)stream, x, w, b, y, g,
                           Co);
    BZ_CHECK_LAUNCH("bz_nn_conv_first");
    return BZ_OK;
}

// The 8-channel first-layer weight gradient applies (Co / 8 a power of two up to 16): about 1024 workgroups.
inline bool first8(int Co) {
    const int G = Co / 8;
    return Co % 8 == 0 && G <= 16 && (G & (G - 1)) == 0;
}
inline long first8_chunk(long P, int Co) { return std::max((long)(256 / (Co / 8)) * 8, (P + 1023) / 1024); }

extern "C" long bz_nn_first_work(int B, int H, int W, int Co) {
    if (B <= 0 || H <= 0 || W <= 0 || Co <= 0) return -1;
    if (!first8(Co)) return 0;
    const long P = (long)B * H * W, chunk = first8_chunk(P, Co), nblk = (P + chunk - 1) / chunk;
    return (nblk + tree_rows(nblk)) * 10L * Co;
}

extern "C" int bz_nn_conv_first_bwd(const uint16_t* x, const uint16_t* gz, const float* w, float* gw, float* gb,
                                    float* gx, int B, int H, int W, int Co, float* work, long nwork, void* stream) {
    BZ_REQUIRE(x && gz && w && gw && gb && B > 0 && H > 0 && W > 0 && Co > 0, "bz_nn_conv_first_bwd: bad arguments");
    Geo g{B, H, W};
    const long P = (long)B * H * W;
    const hipStream_t st = (hipStream_t)stream;
    if (first8(Co) && ((uintptr_t)gz & 15) == 0) {
        const long chunk = first8_chunk(P, Co);
        const long nblk = (P + chunk - 1) / chunk;
        BZ_REQUIRE(!work || nwork >= (nblk + tree_rows(nblk)) * 10 * Co,
                   "bz_nn_conv_first_bwd: workspace smaller than bz_nn_first_work");
        // chunks of whole rows of one image take the row-staged kernel
        const long rb = chunk / W;
        const size_t lds = (size_t)(rb + 2) * (W + 2) * sizeof(float);
        if (rb > 0 && chunk % W == 0 && H % rb == 0 && lds <= 32768)
            hipLaunchKernelGGL(k_conv_first_wgrad_rows, dim3((unsigned)nblk), dim3(256), lds, st, x, gz, gw, gb, work,
                               g, Co, (int)rb);
        else
            hipLaunchKernelGGL(k_conv_first_wgrad8, dim3((unsigned)nblk), dim3(256), 0, st, x, gz, gw, gb, work, g,
                               Co, chunk);
        if (work)   // rows of 10 Co: the 9 Co weight columns into gw, the Co bias columns into gb
            tree_sum2(SumSeg{work, 10L * Co, 9L * Co, gw, (int)nblk},
                      SumSeg{work + 9L * Co, 10L * Co, (long)Co, gb, (int)nblk}, work + nblk * 10L * Co, st);
    } else {
        const int chunk = 8192;
        hipLaunchKernelGGL(k_conv_first_wgrad, dim3((unsigned)((P + chunk - 1) / chunk), Co), dim3(256), 0, st, x, gz,
                           gw, gb, g, Co, chunk);
    }
    BZ_CHECK_LAUNCH("bz_nn_conv_first_bwd(w)");
    if (gx) {
        if (Co % 8 == 0 && Co <= 64 && ((uintptr_t)gz & 15) == 0)
            hipLaunchKernelGGL(k_conv_first_dgrad8, grid1(P), dim3(256), 0, st, gz, w, gx, g, Co);
        else
            hipLaunchKernelGGL(k_conv_first_dgrad, grid1(P), dim3(256), 0, st, gz, w, gx, g, Co);
        BZ_CHECK_LAUNCH("bz_nn_conv_first_bwd(x)");
    }
    return BZ_OK;
}

extern "C" int bz_nn_relu_mask(uint16_t* gz, const uint16_t* y, long n, void* stream) {
    BZ_REQUIRE(gz && y && n >= 0, "bz_nn_relu_mask: bad arguments");
    if (n == 0) return BZ_OK;
    if (n % 8 == 0 && ((uintptr_t)gz & 15) == 0 && ((uintptr_t)y & 15) == 0)
        hipLaunchKernelGGL(k_relu_mask8, grid1(n / 8), dim3(256), 0, (hipStream_t)stream, gz, y, n / 8);
    else
        hipLaunchKernelGGL(k_relu_mask, grid1(n), dim3(256), 0, (hipStream_t)stream, gz, y, n);
    BZ_CHECK_LAUNCH("bz_nn_relu_mask");
    return BZ_OK;
}

extern "C" int bz_nn_avgpool2(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, void* stream) {
    BZ_REQUIRE(x && y && B > 0 && H % 2 == 0 && W % 2 == 0 && H > 0 && W > 0 && C > 0, "bz_nn_avgpool2: bad arguments");
    Geo g{B, H, W};
    if (C % 8 == 0 && ((uintptr_t)x & 15) == 0 && ((uintptr_t)y & 15) == 0)
        hipLaunchKernelGGL(k_avgpool2_8, rows_grid((long)(W / 2) * (C / 8), (long)B * (H / 2)), dim3(256), 0,
                           (hipStream_t)stream, x, y, g, C);
    else
        hipLaunchKernelGGL(k_avgpool2, grid1((long)B * (H / 2) * (W / 2) * C), dim3(256), 0, (hipStream_t)stream, x, y,
                           g, C);
    BZ_CHECK_LAUNCH("bz_nn_avgpool2");
    return BZ_OK;
}

extern "C" int bz_nn_avgpool2_bwd(const uint16_t* gy, uint16_t* gx, int B, int H, int W, int C, void* stream) {
    BZ_REQUIRE(gy && gx && B > 0 && H % 2 == 0 && W % 2 == 0 && H > 0 && W > 0 && C > 0,
               "bz_nn_avgpool2_bwd: bad arguments");
    Geo g{B, H, W};
    if (C % 8 == 0 && ((uintptr_t)gy & 15) == 0 && ((uintptr_t)gx & 15) == 0)
        hipLaunchKernelGGL(k_avgpool2_bwd8<false>, rows_grid((long)W * (C / 8), (long)B * (H / 2)), dim3(256), 0,
                           (hipStream_t)stream, gy, (const uint16_t*)nullptr, gx, g, C);
    else
        hipLaunchKernelGGL(k_avgpool2_bwd, grid1((long)B * H * W * C), dim3(256), 0, (hipStream_t)stream, gy, gx, g, C);
    BZ_CHECK_LAUNCH("bz_nn_avgpool2_bwd");
    return BZ_OK;
}

extern "C" int bz_nn_avgpool2_bwd_relu(const uint16_t* gy, const uint16_t* y, uint16_t* gx, int B, int H, int W, int C,
                                       void* stream) {
    BZ_REQUIRE(gy && y && gx && B > 0 && H % 2 == 0 && W % 2 == 0 && H > 0 && W > 0 && C % 8 == 0 &&
                   ((uintptr_t)gy & 15) == 0 && ((uintptr_t)y & 15) == 0 && ((uintptr_t)gx & 15) == 0,
               "bz_nn_avgpool2_bwd_relu: bad arguments (C a multiple of 8, 16-byte aligned tensors)");
    Geo g{B, H, W};
    hipLaunchKernelGGL(k_avgpool2_bwd8<true>, rows_grid((long)W * (C / 8), (long)B * (H / 2)), dim3(256), 0,
                       (hipStream_t)stream, gy, y, gx, g, C);
    BZ_CHECK_LAUNCH("bz_nn_avgpool2_bwd_relu");
    return BZ_OK;
}

extern "C" int bz_nn_head(const uint16_t* a, const float* fw, const float* fb, float* gap, float* out, int B, int HW,
                          int C, int nout, void* stream) {
    BZ_REQUIRE(a && fw && fb && gap && out && B > 0 && HW > 0 && C > 0 && C <= 4096 && nout > 0,
               "bz_nn_head: bad arguments");
    if (C % 8 == 0 && 256 % (C / 8) == 0)
        hipLaunchKernelGGL(k_head8, dim3(B), dim3(256), (256 / (C / 8) + 1) * C * sizeof(float), (hipStream_t)stream,
                           a, fw, fb, gap, out, HW, C, nout);
    else
        hipLaunchKernelGGL(k_head, dim3(B), dim3(256), C * sizeof(float), (hipStream_t)stream, a, fw, fb, gap, out, HW,
                           C, nout);
    BZ_CHECK_LAUNCH("bz_nn_head");
    return BZ_OK;
}

extern "C" int bz_nn_head_bwd(const float* gout, const float* gap, const float* fw, float* gfw, float* gfb,
                              uint16_t* ga, int B, int HW, int C, int nout, void* stream) {
    BZ_REQUIRE(gout && gap && fw && gfw && gfb && ga && B > 0 && HW > 0 && C > 0 && nout > 0 && nout <= 256,
               "bz_nn_head_bwd: bad arguments");
    hipLaunchKernelGGL(k_head_bwd, dim3(B), dim3(256), 0, (hipStream_t)stream, gout, gap, fw, gfw, gfb, ga, HW, C, nout);
    BZ_CHECK_LAUNCH("bz_nn_head_bwd");
    hipLaunchKernelGGL(k_head_wgrad, grid1((long)nout * C + nout), dim3(256), 0, (hipStream_t)stream, gout, gap, gfw,
                       gfb, B, C, nout);
    BZ_CHECK_LAUNCH("bz_nn_head_bwd(weights)");
    return BZ_OK;
}

extern "C" int bz_nn_weights(const float* w, uint16_t* w16, uint16_t* wt16, int Co, int Ci, void* stream) {
    BZ_REQUIRE(w && w16 && Co > 0 && Ci > 0 && (long)Co * 9 * Ci < (1L << 31), "bz_nn_weights: bad arguments");
    hipLaunchKernelGGL(k_weights, grid1((long)Co * 9 * Ci), dim3(256), 0, (hipStream_t)stream, w, w16, wt16, Co, Ci);
    BZ_CHECK_LAUNCH("bz_nn_weights");
    return BZ_OK;
}

extern "C" int bz_nn_weights_n(const float* const* w, uint16_t* const* w16, uint16_t* const* wt16, const int* Co,
                               const int* Ci, int n, void* stream) {
    BZ_REQUIRE(w && w16 && wt16 && Co && Ci && n >= 1 && n <= WMAX, "bz_nn_weights_n: bad arguments (1..16 layers)");
    WCopies c{};
    c.n = n;
    long blocks = 0;
    for (int j = 0; j < n; ++j) {
        BZ_REQUIRE(w[j] && w16[j] && Co[j] > 0 && Ci[j] > 0 && (long)Co[j] * 9 * Ci[j] < (1L << 31),
                   "bz_nn_weights_n: bad layer");
        c.w[j] = w[j];
        c.w16[j] = w16[j];
        c.wt16[j] = wt16[j];
        c.co[j] = Co[j];
        c.ci[j] = Ci[j];
        c.start[j] = (int)blocks;
        blocks += ((long)Co[j] * 9 * Ci[j] + 255) / 256;
        BZ_REQUIRE(blocks < (1L << 31), "bz_nn_weights_n: too many weights");
    }
    c.start[n] = (int)blocks;
    hipLaunchKernelGGL(k_weights_n, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, c);
    BZ_CHECK_LAUNCH("bz_nn_weights_n");
    return BZ_OK;
}

extern "C" int bz_nn_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                          int step, void* stream) {
    BZ_REQUIRE(p && g && m && v && n >= 0 && step >= 1, "bz_nn_adam: bad arguments");
    if (n == 0) return BZ_OK;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, grid1(n), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    BZ_CHECK_LAUNCH("bz_nn_adam");
    return BZ_OK;
}

extern "C" int bz_nn_to_bf16(const float* x, uint16_t* y, long n, void* stream) {
    BZ_REQUIRE(x && y && n >= 0, "bz_nn_to_bf16: bad arguments");
    if (n == 0) return BZ_OK;
    hipLaunchKernelGGL(k_f2b, grid1(n), dim3(256), 0, (hipStream_t)stream, x, y, n);
    BZ_CHECK_LAUNCH("bz_nn_to_bf16");
    return BZ_OK;
}
