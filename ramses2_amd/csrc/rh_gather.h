// Cell gather of the 3D kernels: the address of a stencil cell from a
// window or table value (an oct, a boundary ghost synthesised from its
// reference oct, or none) and its load.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rh {

__device__ __forceinline__ int reflexive_mask(const int btype[3][2])
{
    int refl = 0;
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int sd = 0; sd < 2; sd++) refl |= (btype[d][sd] == 1) << (2 * d + sd);
    return refl;
}

// gather_cell_m for 3D split into the address computation and the loads, so
// that the loads of the next plane can be issued a plane ahead: the cell's
// five values start at p (stride 8; null: no oct, zero state), its refined
// flag is *rf (null: 0), and momentum component flip (0..2, or -1) is negated
// when the values are used (reflexive boundary ghost).
struct CellRef {
    const double *p;
    const uint8_t *rf;
    int flip;
};

__device__ __forceinline__ CellRef cell_ref3(const double *uold, const uint8_t *refined, int v, const int okey[3],
                                             const int boxmin[3], const int boxmax[3], int refl, const int c[3])
{
    CellRef r;
    r.flip = -1;
    if (v > 0) {
        const int ind = c[0] + 2 * c[1] + 4 * c[2];
        r.p = uold + (size_t)(v - 1) * 40 + ind;
        r.rf = refined + (size_t)(v - 1) * 8 + ind;
        return r;
    }
    if (v == 0) {
        r.p = nullptr;
        r.rf = nullptr;
        return r;
    }
    int dir = 0, side = 0;
    for (int d = 2; d >= 0; d--) {
        if (okey[d] < boxmin[d]) { dir = d; side = 0; }
        else if (okey[d] >= boxmax[d]) { dir = d; side = 1; }
    }
    const int type = (refl >> (2 * dir + side)) & 1 ? 1 : 2;
    int s3[3] = {c[0], c[1], c[2]}, rs[3] = {c[0], c[1], c[2]};
    const bool rvalid = side == 0 ? c[dir] == 1 : c[dir] == 0;
    rs[dir] = side == 0 ? 0 : 1;
    s3[dir] = type == 1 ? 1 - c[dir] : (side == 0 ? 0 : 1);
    const int ref = -v;
    r.p = uold + (size_t)(ref - 1) * 40 + (s3[0] + 2 * s3[1] + 4 * s3[2]);
    r.rf = rvalid ? refined + (size_t)(ref - 1) * 8 + rs[0] + 2 * rs[1] + 4 * rs[2] : nullptr;
    r.flip = type == 1 ? dir : -1;
    return r;
}

__device__ __forceinline__ void cell_load3(const CellRef &r, double u[5], uint8_t &rf)
{
    if (r.p) {
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = r.p[8 * n];
    } else {
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = 0.0;
    }
    rf = r.rf ? *r.rf : (uint8_t)0;
}

__device__ __forceinline__ void cell_flip(double u[5], int flip)
{
    u[1] = flip == 0 ? -u[1] : u[1];
    u[2] = flip == 1 ? -u[2] : u[2];
    u[3] = flip == 2 ? -u[3] : u[3];
}

}  // namespace rh
