// 2D density filter of the topology parametrisation, on B maps of H x W FP32 values:
//
//   out[b][y][x] = oscale[y][x] * sum_{|dy|,|dx| <= r} K[dy + r][dx + r] * iscale[y+dy][x+dx] * in[b][y+dy][x+dx]
//
// with zero padding outside the map (iscale, oscale optional, per pixel, shared by the B maps). The normalised
// cone filter y = (K * x) / (K * 1) is oscale = 1 / (K * 1); its transpose, the gradient, is the correlation
// with the flipped kernel and iscale = 1 / (K * 1). The host code that replaces examples/4_topo.py:120-129
// (blur of the latent density before the projection) calls it in both directions.
//
// A workgroup takes a 32 x 8 pixel tile of one map: the tile and its r-wide halo are staged in LDS once (scaled
// by iscale), then each thread adds its (2r+1)^2 products in row-major kernel order.
#include "common.h"

namespace {

constexpr int TX = 32, TY = 8;
constexpr int RMAX = 8;

__global__ __launch_bounds__(TX* TY) void k_filter2d(const float* __restrict__ in, float* __restrict__ out, int H,
                                                    int W, const float* __restrict__ K, int r,
                                                    const float* __restrict__ iscale,
                                                    const float* __restrict__ oscale) {
    __shared__ float tile[TY + 2 * RMAX][TX + 2 * RMAX];
    __shared__ float ks[(2 * RMAX + 1) * (2 * RMAX + 1)];
    const int n = 2 * r + 1;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const long plane = (long)H * W;
    const float* __restrict__ src = in + blockIdx.z * plane;
    for (int q = threadIdx.x; q < n * n; q += TX * TY) ks[q] = K[q];
    const int tw = TX + 2 * r, th = TY + 2 * r;
    for (int q = threadIdx.x; q < tw * th; q += TX * TY) {
        const int ty = q / tw, tx = q - ty * tw;
        const int y = y0 + ty - r, x = x0 + tx - r;
        float v = 0.0f;
        if (y >= 0 && y < H && x >= 0 && x < W) {
            const long o = (long)y * W + x;
            v = src[o];
            if (iscale) v *= iscale[o];
        }
        tile[ty][tx] = v;
    }
    __syncthreads();
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= W || y >= H) return;
    float s = 0.0f;
    for (int dy = 0; dy < n; ++dy)
        for (int dx = 0; dx < n; ++dx) s += ks[dy * n + dx] * tile[ly + dy][lx + dx];
    const long o = (long)y * W + x;
    if (oscale) s *= oscale[o];
    out[blockIdx.z * plane + o] = s;
}

}  // namespace

extern "C" int bz_filter2d(const float* in, float* out, int B, int H, int W, const float* K, int r,
                           const float* iscale, const float* oscale, void* stream) {
    BZ_REQUIRE(in && out && K && in != out, "bz_filter2d: null or aliased pointer");
    BZ_REQUIRE(B > 0 && H > 0 && W > 0 && B <= 65535, "bz_filter2d: bad shape");
    BZ_REQUIRE(r >= 0 && r <= RMAX, "bz_filter2d: radius must be 0..8");
    const dim3 grid((W + TX - 1) / TX, (H + TY - 1) / TY, B);
    hipLaunchKernelGGL(k_filter2d, grid, dim3(TX * TY), 0, (hipStream_t)stream, in, out, H, W, K, r, iscale, oscale);
    BZ_CHECK_LAUNCH("bz_filter2d");
    return BZ_OK;
}
