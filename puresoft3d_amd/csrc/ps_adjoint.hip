// ps_adjoint.hip -- differentiable texture sampling and shading on a geometry
// buffer, forward and hand-written adjoint (psgpu_diff_forward/_backward).
//
// Model, per covered pixel p with geometry (N, wp, u, v):
//   n = N/|N|, L = norm(light - wp), E = norm(camera - wp), H = norm(L + E)
//   lam = max(n.L, 0), s = max(n.H, 0), spec = s > 0 ? ks * s^e : 0
//   T = sum_k w_k * tex[texel_k]        (bilinear, 4 texels, wrapping)
//   c = a * T * lam + spec              (per RGB channel)
// Adjoint, with g = dLoss/dc:
//   dLoss/dtex[texel_k] += a * lam * w_k * g       (atomics: texels are shared)
//   dLoss/da            += sum_c g_c * T_c * lam
//   dLoss/de            += sum_c g_c * spec * ln(s)
// Visibility is the geometry buffer's: it does not move with the parameters.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/psgpu.h"

namespace {

struct DiffArgs {
    int W, H, tw, th;
    float light[3], camera[3];
    float ks;
};

struct Geo {
    bool covered;
    float lam, s, spec, u, v;
};

__device__ __forceinline__ void norm3(float* a)
{
    float r = 1.0f / sqrtf((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
    a[0] *= r; a[1] *= r; a[2] *= r;
}

__device__ __forceinline__ float dot3(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

__device__ __forceinline__ Geo geometry(const DiffArgs& A, const float* __restrict__ gbuf, int64_t p, float e)
{
    Geo g;
    const float4* gp = (const float4*)(gbuf + p * 12);
    float4 g0 = gp[0], g1 = gp[1], g2 = gp[2];
    g.covered = g0.w != 0.0f;
    if (!g.covered) return g;
    float n[3] = {g0.x, g0.y, g0.z};
    norm3(n);
    float L[3], E[3], Hh[3];
    L[0] = A.light[0] - g1.x; L[1] = A.light[1] - g1.y; L[2] = A.light[2] - g1.z;
    E[0] = A.camera[0] - g1.x; E[1] = A.camera[1] - g1.y; E[2] = A.camera[2] - g1.z;
    norm3(L);
    norm3(E);
    Hh[0] = L[0] + E[0]; Hh[1] = L[1] + E[1]; Hh[2] = L[2] + E[2];
    norm3(Hh);
    float lam = dot3(n, L), s = dot3(n, Hh);
    g.lam = lam > 0.0f ? lam : 0.0f;
    g.s = s > 0.0f ? s : 0.0f;
    g.spec = g.s > 0.0f ? A.ks * powf(g.s, e) : 0.0f;
    g.u = g2.x;
    g.v = g2.y;
    return g;
}

// bilinear footprint: texel j covers u = j/tw; wrapping addresses
__device__ __forceinline__ void footprint(const DiffArgs& A, float u, float v, int64_t* idx, float* w)
{
    float fx = u * (float)A.tw, fy = v * (float)A.th;
    if (!(fabsf(fx) < 1073741824.0f)) fx = 0.0f;
    if (!(fabsf(fy) < 1073741824.0f)) fy = 0.0f;
    float x0 = floorf(fx), y0 = floorf(fy);
    float ax = fx - x0, ay = fy - y0;
    int ix = (int)x0 % A.tw, iy = (int)y0 % A.th;
    if (ix < 0) ix += A.tw;
    if (iy < 0) iy += A.th;
    int ix1 = ix + 1 == A.tw ? 0 : ix + 1, iy1 = iy + 1 == A.th ? 0 : iy + 1;
    idx[0] = (int64_t)iy * A.tw + ix;
    idx[1] = (int64_t)iy * A.tw + ix1;
    idx[2] = (int64_t)iy1 * A.tw + ix;
    idx[3] = (int64_t)iy1 * A.tw + ix1;
    w[0] = (1.0f - ax) * (1.0f - ay);
    w[1] = ax * (1.0f - ay);
    w[2] = (1.0f - ax) * ay;
    w[3] = ax * ay;
}

__global__ void __launch_bounds__(256) k_diff_forward(DiffArgs A, const float* __restrict__ gbuf,
                                                      const float* __restrict__ tex, const float* __restrict__ params,
                                                      float* __restrict__ out)
{
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= (int64_t)A.W * A.H) return;
    float a = params[0], e = params[1];
    Geo g = geometry(A, gbuf, p, e);
    float c[3] = {0.0f, 0.0f, 0.0f};
    if (g.covered) {
        int64_t idx[4];
        float w[4];
        footprint(A, g.u, g.v, idx, w);
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float t = ((w[0] * tex[idx[0] * 3 + ch] + w[1] * tex[idx[1] * 3 + ch]) + w[2] * tex[idx[2] * 3 + ch]) +
                      w[3] * tex[idx[3] * 3 + ch];
            c[ch] = a * t * g.lam + g.spec;
        }
    }
    out[p * 3 + 0] = c[0];
    out[p * 3 + 1] = c[1];
    out[p * 3 + 2] = c[2];
}

__global__ void __launch_bounds__(256) k_diff_backward(DiffArgs A, const float* __restrict__ gbuf,
                                                       const float* __restrict__ tex, const float* __restrict__ params,
                                                       const float* __restrict__ gout, float* __restrict__ gtex,
                                                       float* __restrict__ gparams)
{
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    float ga = 0.0f, ge = 0.0f;
    if (p < (int64_t)A.W * A.H) {
        float a = params[0], e = params[1];
        Geo g = geometry(A, gbuf, p, e);
        if (g.covered) {
            int64_t idx[4];
            float w[4];
            footprint(A, g.u, g.v, idx, w);
            float go[3] = {gout[p * 3], gout[p * 3 + 1], gout[p * 3 + 2]};
            float al = a * g.lam;
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                float t = ((w[0] * tex[idx[0] * 3 + ch] + w[1] * tex[idx[1] * 3 + ch]) + w[2] * tex[idx[2] * 3 + ch]) +
                          w[3] * tex[idx[3] * 3 + ch];
                ga += go[ch] * t * g.lam;
                float gc = go[ch] * al;
                if (gc != 0.0f) {
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (w[k] != 0.0f) unsafeAtomicAdd(gtex + idx[k] * 3 + ch, gc * w[k]);
                }
            }
            if (g.s > 0.0f) ge = ((go[0] + go[1]) + go[2]) * g.spec * logf(g.s);
        }
    }
    // wave sums, one atomic per wave and parameter
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ga += __shfl_xor(ga, o, 64);
        ge += __shfl_xor(ge, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        unsafeAtomicAdd(gparams, ga);
        unsafeAtomicAdd(gparams + 1, ge);
    }
}

int make_args(const psgpu_diff_desc* d, DiffArgs& A)
{
    if (!d || d->width < 1 || d->height < 1 || d->tex_w < 1 || d->tex_h < 1) return PSGPU_ERR_INVALID;
    A.W = d->width; A.H = d->height; A.tw = d->tex_w; A.th = d->tex_h;
    for (int k = 0; k < 3; k++) { A.light[k] = d->light[k]; A.camera[k] = d->camera[k]; }
    A.ks = d->ks;
    return PSGPU_OK;
}

}  // namespace

// These kernels keep no state of their own and touch only the buffers they
// are given; they run on the caller's stream (NULL is HIP's null stream).
extern "C" int psgpu_diff_forward(psgpu_ctx* ctx, const psgpu_diff_desc* desc, const float* gbuf, const float* tex,
                                  const float* params, float* out_rgb, void* stream)
{
    DiffArgs A;
    if (!ctx || !gbuf || !tex || !params || !out_rgb || make_args(desc, A)) return PSGPU_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int64_t n = (int64_t)A.W * A.H;
    hipLaunchKernelGGL(k_diff_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A, gbuf, tex, params, out_rgb);
    return hipGetLastError() == hipSuccess ? PSGPU_OK : PSGPU_ERR_DEVICE;
}

extern "C" int psgpu_diff_backward(psgpu_ctx* ctx, const psgpu_diff_desc* desc, const float* gbuf, const float* tex,
                                   const float* params, const float* grad_out, float* grad_tex, float* grad_params,
                                   void* stream)
{
    DiffArgs A;
    if (!ctx || !gbuf || !tex || !params || !grad_out || !grad_tex || !grad_params || make_args(desc, A))
        return PSGPU_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int64_t n = (int64_t)A.W * A.H;
    hipLaunchKernelGGL(k_diff_backward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A, gbuf, tex, params,
                       grad_out, grad_tex, grad_params);
    return hipGetLastError() == hipSuccess ? PSGPU_OK : PSGPU_ERR_DEVICE;
}
