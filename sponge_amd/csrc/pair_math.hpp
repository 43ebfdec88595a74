// Device math shared by the pair kernel and its adjoint.
#pragma once
#include "common.hpp"

namespace snb {

constexpr float kTwoOverSqrtPi = 1.1283791670955126f;

// LJ type (< 64) packed into the low two bits of the three uint coordinates
// of a slot record.
__device__ __forceinline__ int unpack_type(uint4 v) { return (int)((v.x & 3u) | ((v.y & 3u) << 2) | ((v.z & 3u) << 4)); }

// erfc(x) = exp(-x^2) * erfcx(x) with erfcx(x) ~ t P7(t), t = 1/(1 + p x):
// the form of Abramowitz & Stegun 7.1.26 with eight coefficients refitted
// (iteratively reweighted least squares on [0, 5]) for a relative error of
// 6.3e-8 in erfcx; evaluated in FP32 the relative error of erfc and of the
// force factor 2x/sqrt(pi) e^{-x^2} + erfc(x) stays below 5.2e-7 on [0, 4.5].
__device__ __forceinline__ float erfcx_poly(float x) {
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3195f, x, 1.0f));
    float p = -0.181163783781557f;
    p = fmaf(p, t, 0.7263733451877996f);
    p = fmaf(p, t, -0.8054935211563752f);
    p = fmaf(p, t, 0.8311706379959746f);
    p = fmaf(p, t, -0.18734797244699758f);
    p = fmaf(p, t, 0.27142876592060167f);
    p = fmaf(p, t, 0.16357488112143515f);
    p = fmaf(p, t, 0.18145758631968023f);
    return t * p;
}

}  // namespace snb
