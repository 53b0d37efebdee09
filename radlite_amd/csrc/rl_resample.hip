// Resampling of a circular image cube onto a rectangular grid: the
// interpolation step of imcir_conv (PRO/imcir_conv.pro:161-181, IDL
// triangulate + trigrid: linear interpolation on a triangulation of the
// image's (ring, angle) points).  The triangulation and each pixel's
// triangle and barycentric weights depend only on the geometry and are
// formed once on the host; this kernel applies them to every channel.
//
// img: [ray][channel] as the trace returns it (ray 0 the central beam,
// then ring-major, angle fastest); out: [channel][pixel], pixel = ix +
// nx*iy, the order of IDL's imrect[ix, iy, ifr].
#include "rl_engine.h"

#include <hip/hip_runtime.h>

namespace rl {

// one thread per (pixel, channel), pixels fastest: the writes of a wave
// are contiguous; the few vertex rows a block reads stay in L2
__global__ void __launch_bounds__(256)
k_rect_gather(int npix, int nfr, int nv, const int *__restrict__ vtx,
              const double *__restrict__ w, const double *__restrict__ img,
              double *__restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int ch = blockIdx.y;
    if (p >= npix) return;
    double acc = 0.0;
    for (int k = 0; k < nv; k++) {
        const int v = vtx[(size_t)p * nv + k];
        if (v >= 0) acc = acc + w[(size_t)p * nv + k] * img[(size_t)v * nfr + ch];
    }
    out[(size_t)ch * npix + p] = acc;
}

void launch_rect_gather(hipStream_t st, int npix, int nfr, int nv, const int *vtx,
                        const double *w, const double *img, double *out) {
    if (npix <= 0 || nfr <= 0) return;
    hipLaunchKernelGGL(k_rect_gather, dim3((npix + 255) / 256, nfr), dim3(256), 0, st, npix, nfr,
                       nv, vtx, w, img, out);
}

}  // namespace rl
