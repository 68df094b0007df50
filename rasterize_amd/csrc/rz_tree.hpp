// Kernels of the scene tree's layer composition: Layer::compose
// (src/scene.rs:528-560) over the overlap of two layers, for the Opacity and
// Clip nodes of Pipeline::render_rec (src/scene.rs:436-457).
#pragma once
#include "rz_paint.hpp"

namespace rzk {

// A rectangle of layer pixels: dst pixel (x + dx, y + dy) takes src pixel
// (x + sx, y + sy) for x < w, y < h (both relative to their layer origins).
struct ComposeRect {
  int w, h;
  int dx, dy, sx, sy;
};

// dst = dst.blend_over(src * opacity): an Opacity node's child layer onto its
// parent (opacity 1 for a Clip node's masked child: a product by 1 is exact).
__global__ __launch_bounds__(256) void k_compose_over(rzp::C4* dst, size_t dst_stride,
                                                      const rzp::C4* src, size_t src_stride,
                                                      ComposeRect r, float opacity) {
  const size_t n = (size_t)r.w * r.h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t y = i / (size_t)r.w, x = i - y * (size_t)r.w;
    rzp::C4* d = dst + (y + r.dy) * dst_stride + (x + r.dx);
    const rzp::C4 s = rzp::mul(src[(y + r.sy) * src_stride + (x + r.sx)], opacity);
    *d = rzp::blend_over(*d, s);
  }
}

// dst = dst * mask: a Clip node's child layer times the clip path's coverage
// (all four lanes).
__global__ __launch_bounds__(256) void k_compose_mask(rzp::C4* dst, size_t dst_stride,
                                                      const float* mask, size_t mask_stride,
                                                      ComposeRect r) {
  const size_t n = (size_t)r.w * r.h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t y = i / (size_t)r.w, x = i - y * (size_t)r.w;
    rzp::C4* d = dst + (y + r.dy) * dst_stride + (x + r.dx);
    *d = rzp::mul(*d, mask[(y + r.sy) * mask_stride + (x + r.sx)]);
  }
}

}  // namespace rzk
