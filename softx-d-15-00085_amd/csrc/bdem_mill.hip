/*
 * bdem_mill.hip - the mill's turning lifters, on the device.
 *
 * The reference turns its lifters on the host before every step
 * (Rotate_VolObject_Inc, Input/Processing_Mills.h:7-60, called from
 * UpdateSim, main.cpp:114-121) and re-uploads the whole lifter table to
 * constant memory (Device_Update_DObjectPostions,
 * Device/DeviceInterface.cu:2826-2831).  Here one 64-thread launch turns
 * the lifters inside the parameter block the step kernel reads, on the
 * step's own stream, so a mill run never leaves the device.
 *
 * Arithmetic as on the host: the cosine and sine of the double angle
 * (computed once on the host), the two products and their sum in double,
 * rounded once to float; the face centroids, normals and areas re-derived in
 * float in the reference's order.  The build turns off FMA contraction and
 * uses correctly rounded division and square root, so the results are the
 * host's bit for bit.
 */
#include "bdem_kernels.hpp"

namespace bdem {

__device__ __forceinline__ void turn_point(bdem_float3& v, double c, double s, float ax, float ay, float az) {
  const float px = v.x - ax, py = v.y - ay;
  /* vertex.x = cos*point.x + sin*point.y; vertex.y = -sin*point.x + cos*point.y;
   * vertex.z kept; then vertex += about */
  const float x = (float)(c * (double)px + s * (double)py);
  const float y = (float)(-s * (double)px + c * (double)py);
  v.x = x + ax;
  v.y = y + ay;
  v.z = v.z + az;
}

__global__ __launch_bounds__(64) void k_turn_lifters(DevParams* __restrict__ P, double c, double s, float ax,
                                                     float ay, float az) {
  const int k = threadIdx.x;
  if (k >= P->num_dobj || k >= MAX_DOBJ) return;
  bdem_DOBJ& d = P->dobj[k];
  if (!d.is_attached_DWorld) return;
  for (int v = 0; v < d.num_vertex; ++v) turn_point(d.vertex[v], c, s, ax, ay, az);
  for (int j = 0; j < d.num_faces; ++j) {
    bdem_Surface& f = d.faces[j];
    f3 sum = mk3(0.0f, 0.0f, 0.0f);
    for (int q = 0; q < f.num_vertex; ++q) sum = sum + cvt(d.vertex[f.vertex_Order[q]]);
    const f3 cen = (1.0f / (float)f.num_vertex) * sum;
    const f3 v0 = cvt(d.vertex[f.vertex_Order[0]]), v1 = cvt(d.vertex[f.vertex_Order[1]]);
    const f3 v2 = cvt(d.vertex[f.vertex_Order[2]]), v3 = cvt(d.vertex[f.vertex_Order[3]]);
    const f3 V1 = v1 - v0, V2 = v2 - v1, V3 = v3 - v2, V4 = v0 - v3;
    const f3 n = cross(V2, V1);
    const f3 nn = n / length(n);
    f.centroid = bdem_float3{cen.x, cen.y, cen.z};
    f.normal = bdem_float3{nn.x, nn.y, nn.z};
    f.area = 0.500000f * length(n) + 0.500000f * length(cross(V3, V4));
  }
  turn_point(d.COM, c, s, ax, ay, az);
}

void launch_turn_lifters(DevParams* dP, double c, double s, float ax, float ay, float az, hipStream_t st) {
  hipLaunchKernelGGL(k_turn_lifters, dim3(1), dim3(64), 0, st, dP, c, s, ax, ay, az);
}

}  // namespace bdem
