// Device hash of oct keys and the resolution of a (level, key) to an oct:
// shared by the mesh set-up kernels and the refinement-flag kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rh {

// ---------------------------------------------------------------------------
// Device hash: open addressing, linear probing, 64-bit keys.
// key = key_off(lev) + ix + iy*nx + iz*nx^2 (gpu/gpu_hash.cuf:240-245), nx = 2^(lev-1).
// ---------------------------------------------------------------------------
constexpr uint64_t kEmpty = ~0ull;

__device__ __forceinline__ uint64_t hmix(uint64_t k)
{
    k ^= k >> 31;
    k *= 0x7fb5d329728ea185ull;
    k ^= k >> 27;
    k *= 0x81dadef4bc2dd44dull;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ uint64_t oct_key(int64_t koff, int nx, int ndim, int ix, int iy, int iz)
{
    uint64_t k = (uint64_t)koff + (uint64_t)ix;
    if (ndim > 1) k += (uint64_t)iy * (uint64_t)nx;
    if (ndim > 2) k += (uint64_t)iz * (uint64_t)nx * (uint64_t)nx;
    return k;
}

// Key of an oct outside the computational box (one key beyond a
// non-periodic face): bit 63, log2(nx) in bits 57-61, and the three key
// components + 1 in 18-bit fields.  Disjoint from oct_key (< 2^62) and from
// kEmpty; such octs are ghosts interpolated from the coarser level's boundary
// octs (a stencil oct outside the box whose boundary reference oct is absent).
__device__ __forceinline__ uint64_t ext_key(int nx, int ix, int iy, int iz)
{
    const uint64_t l = (uint64_t)(31 - __builtin_clz((unsigned)nx));
    return (1ull << 63) | (l << 57) | ((uint64_t)(ix + 1) & 0x3ffffull) | (((uint64_t)(iy + 1) & 0x3ffffull) << 18) |
           (((uint64_t)(iz + 1) & 0x3ffffull) << 36);
}

__device__ __forceinline__ int hash_get(const uint64_t *__restrict__ keys, const int32_t *__restrict__ vals,
                                        uint64_t cap, uint64_t key)
{
    uint64_t h = hmix(key) & (cap - 1);
    for (uint64_t n = 0; n < cap; n++) {
        const uint64_t k = keys[h];
        if (k == key) return vals[h];
        if (k == kEmpty) return 0;
        h = (h + 1) & (cap - 1);
    }
    return 0;
}

// Resolve the oct at (possibly out-of-domain) key k of a level.  Returns the
// oct index, -(reference oct) for a boundary ghost, 0 if absent.  Periodic
// directions wrap modulo the box (hydro/godunov_fine.f90:463-468).
struct ResolveCtx {
    const uint64_t *keys;
    const int32_t *vals;
    uint64_t cap;
    int64_t koff;
    int nx, ndim;
    int ngridmax;  // octs above are ghosts (not reference octs of boundary ghosts)
    int boxmin[3], boxmax[3], periodic[3];
};

__device__ __forceinline__ int resolve_oct(const ResolveCtx &c, int k0, int k1, int k2)
{
    int k[3] = {k0, k1, k2};
    for (int d = 0; d < c.ndim; d++) {
        if (c.periodic[d]) {
            const int n = c.boxmax[d] - c.boxmin[d];
            int r = (k[d] - c.boxmin[d]) % n;
            if (r < 0) r += n;
            k[d] = c.boxmin[d] + r;
        }
    }
    int out_dim = -1, side = 0;
    for (int d = c.ndim - 1; d >= 0; d--) {
        if (k[d] < c.boxmin[d]) { out_dim = d; side = 0; }
        else if (k[d] >= c.boxmax[d]) { out_dim = d; side = 1; }
    }
    if (out_dim < 0) return hash_get(c.keys, c.vals, c.cap, oct_key(c.koff, c.nx, c.ndim, k[0], k[1], k[2]));
    // boundary ghost: the oct one step back into the domain along out_dim
    const int kout = k[out_dim];
    k[out_dim] += side == 0 ? 1 : -1;
    for (int d = 0; d < c.ndim; d++)
        if (k[d] < c.boxmin[d] || k[d] >= c.boxmax[d]) return 0;
    const int r = hash_get(c.keys, c.vals, c.cap, oct_key(c.koff, c.nx, c.ndim, k[0], k[1], k[2]));
    if (r > 0 && r <= c.ngridmax) return -r;
    // no reference oct among the level's own octs (an interpolated ghost is
    // not one: the reference keeps no missing octs in its cache): a ghost
    // interpolated from the coarse boundary octs, if one was made
    k[out_dim] = kout;
    if (c.ndim < 3) return 0;
    return hash_get(c.keys, c.vals, c.cap, ext_key(c.nx, k[0], k[1], k[2]));
}

}  // namespace rh
