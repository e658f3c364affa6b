// pp_noise.h -- the reference's Gaussian white noise on the device.
//
// fill_noise_threefry! / _threefry_gaussian (src/InitialConditions/
// RandomField.jl:46-81): cell idx (1-based, x fastest) of an N^3 grid takes
// the first normal of counter pair (idx-1) >> 1 when idx-1 is even, the
// second otherwise; each pair is Threefry-2x64-20 of counter (pair, 0) under
// key (seed, 0) (Salmon et al. 2011), two 53-bit uniforms in (0, 1] and
// Box-Muller in Float64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pp {

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

// the two Float64 normals of counter pair `pair`
__device__ __forceinline__ void threefry_normals(uint64_t pair, uint64_t seed, double &g1, double &g2)
{
    constexpr int ROT[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    const uint64_t ks[3] = {seed, 0ull, 0x1BD11BDAA9FC1A22ull ^ seed};
    uint64_t x0 = pair + ks[0], x1 = ks[1];
#pragma unroll
    for (int blk = 0; blk < 5; blk++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            x0 += x1;
            x1 = rotl64(x1, ROT[(4 * blk + r) & 7]);
            x1 ^= x0;
        }
        const int inj = blk + 1;
        x0 += ks[inj % 3];
        x1 += ks[(inj + 1) % 3] + (uint64_t)inj;
    }
    const double u1 = ((double)(x0 >> 11) + 1.0) * 1.1102230246251565e-16;
    const double u2 = ((double)(x1 >> 11) + 1.0) * 1.1102230246251565e-16;
    const double r = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;
    g1 = r * cos(th);
    g2 = r * sin(th);
}

}  // namespace pp
