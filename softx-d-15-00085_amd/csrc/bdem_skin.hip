/*
 * bdem_skin.hip - which stencil rows the sphere step tests in full.
 *
 * The reference tests every candidate of the 27-cell stencil for every
 * particle every step (Kernel_BroadCollisionDetection_NonSymmetry,
 * Collision_Detection_BroadPhase.cuh:402-513).  Most candidates are far:
 * the step keeps, per particle and stencil row, a mask of the candidates
 * within ra + rb + skin (MaskedRows, bdem_device.hpp) and tests only those.
 * A row's mask stays exact while
 *   (1) no particle entered or left the row's three cells -- the run of
 *       sorted positions then holds the same particles in the same order
 *       (the re-sort is stable), and
 *   (2) the particle and any candidate moved less than the skin together
 *       since the mask was built.
 * k_skin_mark enforces (1): for every particle that changed cell it flags,
 * in every particle of the 27-cell blocks around its old and its new cell,
 * the one row that contains that cell, and all nine rows of the particle
 * itself.  (2) is kept by epochs: the step kernel bounds each particle's
 * displacement, the maxima are summed from the epoch start, and when twice
 * the sum reaches the skin every row of every particle is tested in full.
 * Both are conservative; contacts are found in the reference order either
 * way.  Cell coordinates are recovered from keys, so the masks are used
 * only when the hash is injective (Nx == Nz).
 */
#include "bdem_kernels.hpp"

namespace bdem {

__global__ __launch_bounds__(256) void k_skin_mark(const DevParams* __restrict__ Pp, int n,
                                                   const uint2* __restrict__ cells,
                                                   const uint32_t* __restrict__ meta,
                                                   const unsigned long long* __restrict__ sbuf,
                                                   const uint32_t* __restrict__ old_keys,
                                                   const uint32_t* __restrict__ inv_map,
                                                   uint32_t* __restrict__ rowflag, uint32_t* __restrict__ vmeta,
                                                   int par, int full_host, float skin) {
  const DevParams& P = *Pp;
  /* epoch: the per-step displacement maxima summed since every row was last
   * tested in full, rounded up */
  const float U = __uint_as_float(vmeta[VM_U + par]);
  const float d = __uint_as_float(vmeta[VM_DMAX + (par ^ 1)]);
  const float Un = (U + d) * 1.00000095367431640625f;
  const uint32_t ns = min(meta[META_COUNT], (uint32_t)n);
  /* many cell changes at once: testing everything is cheaper than flagging */
  const bool full = full_host != 0 || !(2.0f * Un < skin) || ns > (uint32_t)n / 16u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    vmeta[VM_U + (par ^ 1)] = __float_as_uint(full ? 0.0f : Un);
    vmeta[VM_FULL] = full ? 1u : 0u;
    vmeta[VM_DMAX + par] = 0u;   /* this step's kernel takes the maximum into it */
  }
  if (full) return;
  const uint32_t nxz = (uint32_t)P.nx * (uint32_t)P.nz;
  /* work items: (changed particle, old/new cell, row of that cell's block),
   * plus one per changed particle for its own rows */
  const uint32_t work = ns * 19u;
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < work; w += gridDim.x * blockDim.x) {
    const uint32_t c = w / 19u, rem = w - c * 19u;
    const unsigned long long e = sbuf[c];
    const uint32_t old_slot = (uint32_t)e;
    if (old_slot >= (uint32_t)n) continue;
    if (rem == 18u) {   /* the particle itself: all its rows are new */
      const uint32_t np = inv_map[old_slot];
      if (np < (uint32_t)n) atomicOr(&rowflag[np], 0x1FFu);
      continue;
    }
    const uint32_t which = rem / 9u, r = rem - which * 9u;
    const uint32_t key = which ? old_keys[old_slot] : (uint32_t)(e >> 32);
    if (key >= P.C) continue;   /* outside the grid or dead: no block */
    const int y = (int)(key / nxz);
    const uint32_t rxz = key - (uint32_t)y * nxz;
    const int z = (int)(rxz / (uint32_t)P.nz);
    const int x = (int)(rxz - (uint32_t)z * (uint32_t)P.nz);
    const int dy = (int)(r / 3u) - 1, dz = (int)(r % 3u) - 1;
    const int yy = y + dy, zz = z + dz;
    if (yy < 0 || yy >= P.ny || zz < 0 || zz >= P.nz) continue;
    const int xs = max(x - 1, 0), xe = min(x + 1, P.nx - 1);
    const uint32_t h0 = hash_of(P, xs, yy, zz), h1 = hash_of(P, xe, yy, zz);
    if (h1 >= P.C) continue;
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (uint32_t h = h0; h <= h1; ++h) {
      const uint2 ce = cells[h];
      if (ce.x != 0xFFFFFFFFu) { lo = min(lo, ce.x); hi = ce.y; }
    }
    if (lo == 0xFFFFFFFFu) continue;
    hi = min(hi, (uint32_t)n);
    /* seen from a particle of this row, the changed cell lies in its row
     * (-dy, -dz) */
    const uint32_t bit = 1u << (uint32_t)((1 - dy) * 3 + (1 - dz));
    for (uint32_t s = lo; s < hi; ++s) atomicOr(&rowflag[s], bit);
  }
}

void launch_skin(const DevParams* dP, const StepBuffers& b, const uint32_t* old_keys, int par, bool full,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_skin_mark, dim3(64), dim3(256), 0, s, dP, b.n, b.cells, b.meta, b.sbuf, old_keys,
                     b.inv_map, b.rowflag, b.vmeta, par, full ? 1 : 0, b.skin);
}

}  // namespace bdem
