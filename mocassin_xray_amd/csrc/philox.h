// Counter-based random numbers for packet histories: Philox4x32-10
// (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11).  A packet's stream is a pure function of (seed, packet id, draw
// index, ctr_base), so a history does not depend on which lane or launch
// shape runs it.  This replaces the clock-seeded Fortran random_number of
// energyPacketDriver (photon_mod.f90:97-114).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MOC_HD __host__ __device__ inline
#else
#define MOC_HD static inline
#endif

struct MocU32x4 {
    uint32_t v[4];
};

MOC_HD MocU32x4 moc_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                  uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c0;
        const uint64_t p1 = (uint64_t)M1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    MocU32x4 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
}

// Stream of one packet: draw d of packet p is block (d, p_lo, p_hi, base).
struct MocRng {
    uint32_t k0, k1;   // seed
    uint32_t p_lo, p_hi;
    uint32_t base;     // iteration / source id
    uint32_t draw;

    MOC_HD void init(uint64_t seed, uint64_t packet, uint32_t ctr_base) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        p_lo = (uint32_t)packet;
        p_hi = (uint32_t)(packet >> 32);
        base = ctr_base;
        draw = 0;
    }
    // uniform on [0,1) with 53 random bits
    MOC_HD double u64() {
        MocU32x4 r = moc_philox4x32_10(draw++, p_lo, p_hi, base, k0, k1);
        const uint64_t m = ((uint64_t)r.v[0] << 21) ^ (uint64_t)(r.v[1] >> 11);
        return (double)m * (1.0 / 9007199254740992.0);
    }
    // uniform on [0,1) with 24 random bits
    MOC_HD float u32() {
        MocU32x4 r = moc_philox4x32_10(draw++, p_lo, p_hi, base, k0, k1);
        return (float)(r.v[0] >> 8) * (1.0f / 16777216.0f);
    }
    template <typename T>
    MOC_HD T uniform();
};

template <>
MOC_HD double MocRng::uniform<double>() { return u64(); }
template <>
MOC_HD float MocRng::uniform<float>() { return u32(); }
